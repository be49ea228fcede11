"""FITS loaders and writers (Core/FitsIO.cs, SingleReconstruction/MeasurementData.cs).

The reference ships no .fits file and reads FITS through the CSharpFITS
package, which is not in its tree, so the readers are pinned three ways:

  * files written byte by byte here from literal 80-column cards and
    struct-packed big-endian values (FITS Standard 4.0), read by both the
    oracle restatement and the native library;
  * array contents that make every index distinct, so a transposed axis, a
    wrong polarization or a missed v flip changes the result, plus the
    reference's own quirks (flags whose polarizations cancel count as
    unflagged; the flag sum runs in order, so inf - inf flags);
  * the native writer's bytes against the standard's fixed-format cards.

These run on the CPU: the FITS code is host code in libp9idg.so.
"""
import ctypes
import os
import struct

import numpy as np
import pytest

import oracle_fits as OF
import p9idg
from p9idg import FitsIO, MeasurementData, _lib


def literal_file(path, cards, payload):
    """A FITS file from literal cards (each padded to 80 columns) and raw data bytes."""
    head = "".join(c.ljust(80) for c in cards + ["END"]).encode("ascii")
    head += b" " * (-len(head) % 2880)
    body = payload + b"\0" * (-len(payload) % 2880)
    with open(path, "wb") as f:
        f.write(head + body)
    return path


def bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


# ---- header parsing and the codec ------------------------------------------


def test_literal_frequencies_file(tmp_path):
    # fixed-format and free-format values, comments, COMMENT / HISTORY / blank
    # cards and a string value all appear in real headers
    cards = [
        "SIMPLE  =                    T / file does conform to FITS standard",
        "BITPIX  = -64",
        "NAXIS   =                    1",
        "NAXIS1  =                    4 / channels",
        "EXTEND  =                    T",
        "OBJECT  = 'LMC     '           / name",
        "COMMENT frequencies in Hz",
        "HISTORY written byte by byte",
        "",
        "BSCALE  =                  2.0 / not applied by the reference",
    ]
    vals = [1.0e9, 1.25e9, -0.0, 1e300]
    p = literal_file(tmp_path / "freq.fits", cards, struct.pack(">4d", *vals))
    want = np.array(vals)
    for got in (OF.read_frequencies(p), FitsIO.ReadFrequencies(p)):
        assert np.array_equal(bits(got), bits(want))
    assert FitsIO.ReadInfo(p) == (-64, (4,))


def test_header_spanning_two_records(tmp_path):
    cards = ["SIMPLE  =                    T", "BITPIX  =                  -32", "NAXIS   =                    2",
             "NAXIS1  =                    3", "NAXIS2  =                    2"]
    cards += ["COMMENT %d" % i for i in range(40)]  # 46 cards with END: data at byte 5760
    vals = [1.5, -2.0, 3.25, 0.0, 7.0, -8.5]
    p = literal_file(tmp_path / "img.fits", cards, struct.pack(">6f", *vals))
    assert os.path.getsize(p) == 3 * 2880
    want = np.array(vals, np.float32).reshape(2, 3)
    assert np.array_equal(OF.read_image(p, np.float32), want)
    assert np.array_equal(FitsIO.ReadImage(p), want)


def test_oracle_encoder_matches_literal_cards(tmp_path):
    cards = ["SIMPLE  =                    T", "BITPIX  =                  -64", "NAXIS   =                    2",
             "NAXIS1  =                    2", "NAXIS2  =                    1"]
    p = literal_file(tmp_path / "a.fits", cards, struct.pack(">2d", 0.5, -3.0))
    assert OF.encode(np.array([[0.5, -3.0]]), -64) == open(p, "rb").read()


@pytest.mark.parametrize("img", [np.arange(15, dtype=np.float32).reshape(3, 5) - 7.25,
                                 np.linspace(-1, 1, 12).reshape(4, 3)])
def test_native_writer_bytes_are_the_standard_fixed_format(tmp_path, img):
    p = str(tmp_path / "w.fits")
    FitsIO.Write(img, p)
    bitpix = -32 if img.dtype == np.float32 else -64
    expected = OF.encode(img, bitpix, [OF.card("EXTEND", "T")])
    assert open(p, "rb").read() == expected
    assert os.path.getsize(p) % 2880 == 0


def test_write_complex_real_and_imag_parts_and_cubes(tmp_path):
    rng = np.random.default_rng(3)
    z = rng.standard_normal((6, 7)) + 1j * rng.standard_normal((6, 7))
    FitsIO.Write(z, str(tmp_path / "re.fits"))
    FitsIO.WriteImag(z, str(tmp_path / "im.fits"))
    assert np.array_equal(OF.read_image(tmp_path / "re.fits", np.float64), z.real)
    assert np.array_equal(OF.read_image(tmp_path / "im.fits", np.float64), z.imag)
    cube = rng.standard_normal((3, 4, 5))
    FitsIO.Write(cube, str(tmp_path / "cube"))
    for k in range(3):
        assert np.array_equal(FitsIO.ReadImageDouble(str(tmp_path / ("cube%d.fits" % k))), cube[k])
    with pytest.raises(TypeError):
        FitsIO.Write(np.zeros((2, 2), np.int32), str(tmp_path / "i.fits"))


def test_images_beam_and_casa(tmp_path):
    rng = np.random.default_rng(5)
    f32 = rng.standard_normal((9, 11)).astype(np.float32)
    f64 = rng.standard_normal((8, 8))
    OF.write(tmp_path / "f32.fits", f32, -32)
    OF.write(tmp_path / "f64.fits", f64, -64)
    assert np.array_equal(FitsIO.ReadImage(str(tmp_path / "f32.fits")), f32)
    assert np.array_equal(FitsIO.ReadImageDouble(str(tmp_path / "f64.fits")), f64)
    assert np.array_equal(FitsIO.ReadBeam(str(tmp_path / "f64.fits")), f64)
    # CASA images are float[Stokes][frequency][rows][cols]; the reference takes plane [0][0]
    casa = rng.standard_normal((2, 3, 6, 5)).astype(np.float32)
    OF.write(tmp_path / "casa.fits", casa, -32)
    assert np.array_equal(OF.read_casa(tmp_path / "casa.fits"), casa[0, 0])
    assert np.array_equal(FitsIO.ReadCASAFits(str(tmp_path / "casa.fits")), casa[0, 0])
    # the reference's casts: a double file is not a float image and vice versa
    for fn, name in ((FitsIO.ReadImage, "f64.fits"), (FitsIO.ReadImageDouble, "f32.fits"),
                     (FitsIO.ReadCASAFits, "f32.fits")):
        with pytest.raises(p9idg.P9IDGError) as e:
            fn(str(tmp_path / name))
        assert e.value.status == _lib.ERR_IO


# ---- measurement files -------------------------------------------------------


def indexed(shape):
    """Values that differ in every index: sum of index * distinct scales."""
    a = np.zeros(shape)
    for ax, n in enumerate(shape):
        idx = np.arange(n).reshape([-1 if i == ax else 1 for i in range(len(shape))])
        a = a + idx * (0.5 + 1.37 ** (len(shape) - ax)) * (1 + ax)
    return a


def test_uvw_flips_v_and_ranges(tmp_path):
    a = indexed((5, 7, 4)) - 3.0  # a fourth component per row is ignored
    a[1, 2, 1] = 0.0  # -> -0.0
    OF.write(tmp_path / "uvw.fits", a, -64)
    p = str(tmp_path / "uvw.fits")
    want = a[:, :, :3].copy()
    want[..., 1] = -want[..., 1]
    got = FitsIO.ReadUVW(p)
    assert np.array_equal(bits(got), bits(want))
    assert np.array_equal(bits(OF.read_uvw(p)), bits(want))
    assert np.signbit(got[1, 2, 1])
    assert np.array_equal(FitsIO.ReadUVW(p, 2, 4), want[2:4])
    assert FitsIO.CountBaselines(p) == OF.count_baselines(p) == 5
    with pytest.raises(p9idg.P9IDGError) as e:
        FitsIO.ReadUVW(p, 3, 6)
    assert e.value.status == _lib.ERR_ARGUMENT


def test_visibilities_sum_xx_and_yy(tmp_path):
    B, T, C, P = 4, 5, 3, 4
    a = indexed((B, T, C, P, 2))
    # one channel with literal polarizations: XX = 1+2i, XY and YX large, YY = 3+4i
    a[2, 3, 1] = [[1.0, 2.0], [1e3, 1e4], [1e5, 1e6], [3.0, 4.0]]
    OF.write(tmp_path / "vis.fits", a, -64)
    p = str(tmp_path / "vis.fits")
    got = FitsIO.ReadVisibilities(p, B, T, C, 2.0)
    ref = OF.read_visibilities(p, 0, B, T, C, 2.0)
    assert got[2, 3, 1] == 2.0 + 3.0j
    assert np.array_equal(bits(got), bits(ref))
    assert np.array_equal(got, (a[..., 0, 0] + a[..., 3, 0]) / 2.0 + 1j * ((a[..., 0, 1] + a[..., 3, 1]) / 2.0))
    # leading blocks and baseline ranges read the same numbers
    assert np.array_equal(FitsIO.ReadVisibilities(p, 1, 3, 2, 2, 2.0), ref[1:3, :2, :2])
    assert np.array_equal(FitsIO.ReadVisibilities(p, 2, T, C, 2.0), ref[:2])
    # complex64 output rounds the double result
    c64 = FitsIO.ReadVisibilities(p, B, T, C, 2.0, dtype=np.complex64)
    assert c64.dtype == np.complex64 and np.array_equal(c64, ref.astype(np.complex64))
    with pytest.raises(p9idg.P9IDGError) as e:
        FitsIO.ReadVisibilities(p, B, T, C + 1, 2.0)
    assert e.value.status == _lib.ERR_ARGUMENT


def test_flags_follow_the_ordered_polarization_sum(tmp_path):
    rows = [
        ([0.0, 0.0, 0.0, 0.0], False),
        ([1.0, -1.0, 0.0, 0.0], False),  # cancelling polarizations read as unflagged
        ([-0.0, 0.0, 0.0, 0.0], False),
        ([0.0, 0.0, 0.0, 1.0], True),
        ([np.nan, 0.0, 0.0, 0.0], True),
        ([1e308, 1e308, -1e308, -1e308], True),  # in order: inf, inf, inf
        ([1e308, -1e308, 1e308, -1e308], False),
        ([1e308, 1e308, -np.inf, 0.0], True),  # inf - inf = NaN
    ]
    a = np.array([r for r, _ in rows]).reshape(2, 2, 2, 4)
    want = np.array([f for _, f in rows]).reshape(2, 2, 2)
    OF.write(tmp_path / "flags.fits", a, -64)
    p = str(tmp_path / "flags.fits")
    assert np.array_equal(OF.read_flags(p, 0, 2, 2, 2), want)
    assert np.array_equal(FitsIO.ReadFlags(p, 2, 2, 2), want)
    assert np.array_equal(FitsIO.ReadFlags(p, 1, 2, 2, 1), want[1:2, :, :1])


def test_random_files_native_equals_oracle(tmp_path):
    rng = np.random.default_rng(11)
    B, T, C, P = 6, 9, 5, 4
    vis = rng.standard_normal((B, T, C, P, 2))
    flags = (rng.random((B, T, C, P)) < 0.2).astype(np.float64)
    uvw = rng.standard_normal((B, T, 3)) * 1e3
    for name, arr in (("vis", vis), ("flags", flags), ("uvw", uvw)):
        OF.write(tmp_path / (name + ".fits"), arr, -64)
    for threads in (1, 3, 0):
        FitsIO.SetThreads(threads)
        v = FitsIO.ReadVisibilities(str(tmp_path / "vis.fits"), B, T, C, 1.5)
        assert np.array_equal(bits(v), bits(OF.read_visibilities(tmp_path / "vis.fits", 0, B, T, C, 1.5)))
        f = FitsIO.ReadFlags(str(tmp_path / "flags.fits"), B, T, C)
        assert np.array_equal(f, OF.read_flags(tmp_path / "flags.fits", 0, B, T, C))
        u = FitsIO.ReadUVW(str(tmp_path / "uvw.fits"))
        assert np.array_equal(bits(u), bits(OF.read_uvw(tmp_path / "uvw.fits")))
    FitsIO.SetThreads(0)


# ---- errors ------------------------------------------------------------------


def test_error_paths(tmp_path):
    def status(fn, *a):
        with pytest.raises(p9idg.P9IDGError) as e:
            fn(*a)
        return e.value.status, str(e.value)

    st, msg = status(FitsIO.ReadFrequencies, str(tmp_path / "missing.fits"))
    assert st == _lib.ERR_IO and "missing.fits" in msg
    (tmp_path / "text.fits").write_bytes(b"hello world\n" * 400)
    assert status(FitsIO.ReadFrequencies, str(tmp_path / "text.fits"))[0] == _lib.ERR_IO
    good = OF.encode(np.arange(1000.0), -64)
    (tmp_path / "trunc.fits").write_bytes(good[:2880 + 4000])
    st, msg = status(FitsIO.ReadFrequencies, str(tmp_path / "trunc.fits"))
    assert st == _lib.ERR_IO and "truncated" in msg
    noend = good[:2880].replace(b"END" + b" " * 77, b"COMMENT" + b" " * 73)
    (tmp_path / "noend.fits").write_bytes(noend)
    assert "END" in status(FitsIO.ReadFrequencies, str(tmp_path / "noend.fits"))[1]
    OF.write(tmp_path / "f32.fits", np.zeros(4, np.float32), -32)
    assert status(FitsIO.ReadFrequencies, str(tmp_path / "f32.fits"))[0] == _lib.ERR_IO
    OF.write(tmp_path / "vis2.fits", np.zeros((2, 2, 2, 2, 2)), -64)  # only 2 polarizations
    assert status(FitsIO.ReadVisibilities, str(tmp_path / "vis2.fits"), 2, 2, 2, 2.0)[0] == _lib.ERR_IO


# ---- MeasurementData ---------------------------------------------------------


def write_part(folder, suffix, rng, B, T, C):
    vis = rng.standard_normal((B, T, C, 4, 2))
    flags = (rng.random((B, T, C, 4)) < 0.1).astype(np.float64)
    uvw = rng.standard_normal((B, T, 3)) * 500.0
    OF.write(os.path.join(folder, "vis%s.fits" % suffix), vis, -64)
    OF.write(os.path.join(folder, "flags%s.fits" % suffix), flags, -64)
    OF.write(os.path.join(folder, "uvw%s.fits" % suffix), uvw, -64)


def test_load_lmc_stitches_eight_parts(tmp_path):
    rng = np.random.default_rng(21)
    T, C = 4, 3
    freq = np.linspace(1.0e9, 1.1e9, C)
    OF.write(tmp_path / "freq.fits", freq, -64)
    sizes = [2, 3, 1, 2, 2, 1, 3, 2]
    for i, b in enumerate(sizes):
        write_part(str(tmp_path), str(i), rng, b, T, C)
    d = MeasurementData.LoadLMC(str(tmp_path))
    uvw = np.concatenate([OF.read_uvw(tmp_path / ("uvw%d.fits" % i)) for i in range(8)])
    flags = np.concatenate([OF.read_flags(tmp_path / ("flags%d.fits" % i), 0, b, T, C) for i, b in enumerate(sizes)])
    vis = np.concatenate([OF.read_visibilities(tmp_path / ("vis%d.fits" % i), 0, b, T, C, 2.0)
                          for i, b in enumerate(sizes)])
    assert np.array_equal(d.Frequencies, freq)
    assert np.array_equal(bits(d.UVW), bits(uvw))
    assert np.array_equal(d.Flags, flags)
    assert np.array_equal(bits(d.Visibilities), bits(vis))
    assert d.VisibilitiesCount == int((~flags).sum()) and 0 < d.VisibilitiesCount < flags.size
    assert np.array_equal(FitsIO.Stitch(uvw[:2], uvw[2:]), uvw)


def test_simulated_points_round_trip_of_an_observation(tmp_path):
    """An observation written in the measurement layout (v stored negated,
    XX = YY = V so that (XX + YY) / 2 = V exactly) loads back bit for bit."""
    from p9idg import observation as obs

    ob = obs.custom(G=128, S=16, K=4, L=1, n_ant=8, T=12, C=4, seed=4, core_diam=800.0, cell_arcsec=20.0,
                    sources=[(1.0, 0.0, 0.0, 0.0)], noise=0.01)
    stored = ob.uvw.copy()
    stored[..., 1] = -stored[..., 1]
    pols = np.zeros(ob.vis.shape + (4, 2))
    for p in (0, 3):
        pols[..., p, 0] = ob.vis.real
        pols[..., p, 1] = ob.vis.imag
    OF.write(tmp_path / "freq.fits", ob.freq, -64)
    OF.write(tmp_path / "uvw.fits", stored, -64)
    OF.write(tmp_path / "vis.fits", pols, -64)
    d = MeasurementData.LoadSimulatedPoints(str(tmp_path))
    assert np.array_equal(d.Frequencies, ob.freq)
    assert np.array_equal(bits(d.UVW), bits(ob.uvw))
    assert np.array_equal(bits(d.Visibilities), bits(ob.vis))
    assert not d.Flags.any() and d.VisibilitiesCount == ob.vis.size
    c = MeasurementData.CreateSimulatedStandardParams(d.VisibilitiesCount)
    assert (c.GridSize, c.SubgridSize, c.KernelSize, c.MaxTimestepsPerSubgrid, c.WLayerCount) == (256, 8, 4, 1024, 1)
    assert c.ScaleArcSec == np.float32(1.0 / 3600.0 * np.pi / 180.0)


def test_corrupted_files_fail_cleanly(tmp_path):
    """Truncated files, flipped header bytes, absurd or non-numeric axis
    values and random bytes: every reader returns OK, P9IDG_ERR_IO or
    P9IDG_ERR_ARGUMENT (never reads past the file or writes past the
    caller's buffer, whose tail is guarded)."""
    lib = _lib.idg()
    rng = np.random.default_rng(2024)
    valid = [OF.encode(np.arange(12, dtype=np.float64), -64),
             OF.encode(rng.standard_normal((3, 4)).astype(np.float32), -32),
             OF.encode(rng.standard_normal((2, 5, 3)), -64),
             OF.encode(rng.standard_normal((1, 1, 3, 4)).astype(np.float32), -32)]
    bad_values = [b"-1", b"0", b"999999999", b"2147483647", b"-2147483648", b"1e300", b"abc", b"", b"T", b"9" * 40]
    ok = {_lib.OK, _lib.ERR_IO, _lib.ERR_ARGUMENT}
    guard = 1e30
    for trial in range(400):
        data = bytearray(valid[trial % len(valid)])
        kind = trial % 5
        if kind == 0:  # truncation
            data = data[:int(rng.integers(0, len(data)))]
        elif kind == 1:  # random bytes flipped in the header record
            for _ in range(int(rng.integers(1, 20))):
                data[int(rng.integers(0, 2880))] = int(rng.integers(0, 256))
        elif kind == 2:  # an axis or BITPIX value replaced
            card = int(rng.integers(1, 6))
            v = bad_values[int(rng.integers(0, len(bad_values)))]
            field = v.rjust(20)[:70]
            data[card * 80 + 10:card * 80 + 10 + len(field)] = field
        elif kind == 3:  # END removed
            i = bytes(data).find(b"END" + b" " * 77)
            if i >= 0:
                data[i:i + 3] = b"XYZ"
        else:  # random bytes
            data = bytearray(rng.integers(0, 256, int(rng.integers(0, 6000)), dtype=np.uint8).tobytes())
        path = str(tmp_path / ("f%d.fits" % trial))
        with open(path, "wb") as f:
            f.write(bytes(data))
        p = path.encode()
        info = _lib.FitsInfo()
        assert lib.p9fits_read_info(p, ctypes.byref(info)) in ok, trial
        freq = np.full(17, guard)
        n = ctypes.c_int64()
        assert lib.p9fits_read_frequencies(p, freq.ctypes.data, 16, ctypes.byref(n)) in ok, trial
        assert freq[16] == guard
        b, t = ctypes.c_int32(), ctypes.c_int32()
        assert lib.p9fits_count_baselines(p, ctypes.byref(b), ctypes.byref(t)) in ok, trial
        img = np.full(13, guard, dtype=np.float64)
        assert lib.p9fits_read_image(p, _lib.FITS_F64, img.ctypes.data, 3, 4) in ok, trial
        assert img[12] == guard
        img32 = np.full(13, guard, dtype=np.float32)
        assert lib.p9fits_read_casa(p, img32.ctypes.data, 3, 4) in ok, trial
        assert img32[12] == np.float32(guard)
