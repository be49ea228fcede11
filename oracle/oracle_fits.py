"""The reference's FITS readers (Core/FitsIO.cs) restated in numpy on the CPU.

TEST INFRASTRUCTURE ONLY: tests import this module as the checker of the
native readers and writers in libp9idg.so.  The product package never
imports it.

The reference reads FITS through the CSharpFITS NuGet package, version 1.1.0
(Core/Core.csproj:8), a C# port of the Java nom.tam.fits library; it is not
part of the reference tree and no .fits file ships with it.  What the
reference relies on is the FITS Standard 4.0 primary HDU:

  * 2880-byte records; the header is a sequence of 80-character cards
    (keyword in columns 1-8, "= " in columns 9-10 for a value, a "/" starting
    the comment), opened by SIMPLE = T and closed by END, padded with blanks
    to a record boundary;
  * BITPIX (8, 16, 32, 64 integers, -32 / -64 IEEE floats), NAXIS and
    NAXIS1..NAXISn; the data follow the header, big-endian, NAXIS1 varying
    fastest, padded to a record boundary;
  * ImageHDU.Kernel hands back the raw array as nested jagged arrays with
    NAXISn outermost (so kernel[i][j]... indexes NAXISn, NAXISn-1, ...), and
    BSCALE / BZERO are not applied.

The reference then casts the kernel to double[] / float[] leaves, so a file of
any other element type fails with an InvalidCastException; this restatement
raises ValueError instead.
"""
import numpy as np

RECORD = 2880
CARD = 80
_DTYPES = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


def _value(card):
    """Value field of a card, comment and blanks removed (None without '= ')."""
    if card[8:10] != "= ":
        return None
    return card[10:].split("/", 1)[0].strip()


def read_primary(path):
    """(bitpix, shape in numpy order, array) of the primary HDU; the array is
    a big-endian view of the file bytes."""
    raw = open(path, "rb").read()
    if len(raw) < RECORD:
        raise ValueError("not a FITS file")
    cards = [raw[i:i + CARD].decode("latin-1") for i in range(0, len(raw) - CARD + 1, CARD)]
    if cards[0][:8].rstrip() != "SIMPLE" or _value(cards[0]) != "T":
        raise ValueError("first card is not SIMPLE = T")
    keys = {}
    end = None
    for i, card in enumerate(cards[1:], start=1):
        k = card[:8].rstrip()
        if k == "END":
            end = i
            break
        v = _value(card)
        if v is not None and k not in keys:
            keys[k] = v
    if end is None:
        raise ValueError("no END card")
    bitpix, naxis = int(keys["BITPIX"]), int(keys["NAXIS"])
    axes = [int(keys["NAXIS%d" % (a + 1)]) for a in range(naxis)]  # NAXIS1 first
    shape = tuple(reversed(axes))
    offset = -(-(end + 1) * CARD // RECORD) * RECORD
    count = int(np.prod(shape)) if naxis else 0
    data = np.frombuffer(raw, dtype=_DTYPES[bitpix], count=count, offset=offset).reshape(shape)
    return bitpix, shape, data


def _cast(path, bitpix, naxis, what):
    b, shape, data = read_primary(path)
    if b != bitpix or len(shape) != naxis:
        raise ValueError("%s needs BITPIX %d with NAXIS %d, file has BITPIX %d, shape %s" % (what, bitpix, naxis, b, shape))
    return data


def read_frequencies(path):
    """FitsIO.cs:184-189."""
    return _cast(path, -64, 1, "ReadFrequencies").astype(np.float64)


def read_uvw(path, from_bl=0, to_bl=None):
    """FitsIO.cs:196-245: u, -v, w of baselines [from_bl, to_bl), every timestep."""
    a = _cast(path, -64, 3, "ReadUVW")
    to_bl = a.shape[0] if to_bl is None else to_bl
    if not 0 <= from_bl <= to_bl <= a.shape[0]:
        raise IndexError("baseline range")
    out = np.empty((to_bl - from_bl, a.shape[1], 3), dtype=np.float64)
    for i in range(from_bl, to_bl):
        for j in range(a.shape[1]):
            out[i - from_bl, j, 0] = a[i, j, 0]
            out[i - from_bl, j, 1] = -a[i, j, 1]
            out[i - from_bl, j, 2] = a[i, j, 2]
    return out


def count_baselines(path):
    """FitsIO.cs:247-255."""
    return _cast(path, -64, 3, "CountBaselines").shape[0]


def read_visibilities(path, from_bl, to_bl, timesteps, channels, norm):
    """FitsIO.cs:267-325: ((XX.re + YY.re) / norm, (XX.im + YY.im) / norm) with
    XX = polarization 0 and YY = polarization 3 of double[B][T][C][P][2]."""
    a = _cast(path, -64, 5, "ReadVisibilities")
    if not 0 <= from_bl <= to_bl <= a.shape[0] or timesteps > a.shape[1] or channels > a.shape[2]:
        raise IndexError("range")
    out = np.empty((to_bl - from_bl, timesteps, channels), dtype=np.complex128)
    for i in range(from_bl, to_bl):
        for j in range(timesteps):
            for k in range(channels):
                xx, yy = a[i, j, k, 0], a[i, j, k, 3]
                out[i - from_bl, j, k] = complex((float(xx[0]) + float(yy[0])) / norm,
                                                 (float(xx[1]) + float(yy[1])) / norm)
    return out


def read_flags(path, from_bl, to_bl, timesteps, channels):
    """FitsIO.cs:327-390: flagged unless the polarization values, summed in
    order from 0.0, equal 0."""
    a = _cast(path, -64, 4, "ReadFlags")
    if not 0 <= from_bl <= to_bl <= a.shape[0] or timesteps > a.shape[1] or channels > a.shape[2]:
        raise IndexError("range")
    out = np.empty((to_bl - from_bl, timesteps, channels), dtype=np.bool_)
    for i in range(from_bl, to_bl):
        for j in range(timesteps):
            for k in range(channels):
                s = 0.0
                for p in a[i, j, k]:
                    s += float(p)
                out[i - from_bl, j, k] = not (s == 0)
    return out


def read_image(path, dtype):
    """FitsIO.cs:152-182: float (BITPIX -32) or double (BITPIX -64) rows x cols."""
    bitpix = -32 if np.dtype(dtype) == np.float32 else -64
    return _cast(path, bitpix, 2, "ReadImage").astype(dtype)


def read_casa(path):
    """FitsIO.cs:134-150: plane [0][0] of float[.][.][rows][cols]."""
    return _cast(path, -32, 4, "ReadCASAFits")[0, 0].astype(np.float32)


def card(key, value=None, comment=None):
    """One 80-column card; a value is right-justified to column 30."""
    if value is None:
        text = "%-8s" % key
    else:
        text = "%-8s= %20s" % (key, value)
        if comment:
            text += " / " + comment
    if len(text) > CARD:
        raise ValueError("card too long")
    return text.ljust(CARD)


def encode(array, bitpix, extra_cards=()):
    """Bytes of a FITS file whose primary array is `array` (numpy order:
    NAXISn first) stored with the given BITPIX, extra cards after the
    mandatory ones."""
    a = np.ascontiguousarray(array, dtype=_DTYPES[bitpix])
    cards = [card("SIMPLE", "T"), card("BITPIX", str(bitpix)), card("NAXIS", str(a.ndim))]
    cards += [card("NAXIS%d" % (i + 1), str(n)) for i, n in enumerate(reversed(a.shape))]
    cards += list(extra_cards) + [card("END")]
    head = "".join(cards).encode("ascii")
    head += b" " * (-len(head) % RECORD)
    body = a.tobytes()
    body += b"\0" * (-len(body) % RECORD)
    return head + body


def write(path, array, bitpix, extra_cards=()):
    with open(path, "wb") as f:
        f.write(encode(array, bitpix, extra_cards))
