"""The FP64 reconstruction fixture of the large workload
(tests/golden/cycle_large.npz, written by tests/golden/make_cycle_fixture.py
with the oracle's MajorCycle.ReconstructSerialCD, MajorCycle.cs:33-134):
loadable without pickles, written with the reference's LMC settings
(RunningMethod.cs:41-64), and self-consistent (the model image it stores
hashes to its recorded digest; the per-cycle figures follow the reference's
rules)."""
import hashlib
import os

import numpy as np
import pytest

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cycle_large.npz")


@pytest.fixture(scope="module")
def fx():
    return np.load(FIXTURE, allow_pickle=False)


def model(fx):
    G = int(fx["grid_size"])
    x = np.zeros(G * G, np.float32)
    x[fx["model_index"]] = fx["model_value"]
    return x.reshape(G, G)


def test_settings_are_the_reference_lmc_run(fx):
    assert int(fx["grid_size"]) == 3072
    assert float(fx["setting_psf_cut"]) == 16
    assert float(fx["setting_lambda"]) == 1.0
    assert float(fx["setting_alpha"]) == 0.01
    assert float(fx["setting_iterations"]) == 30000
    assert float(fx["setting_epsilon"]) == 1e-5
    assert float(fx["setting_cycles"]) == 10


def test_model_image_matches_its_digest(fx):
    x = model(fx)
    assert hashlib.sha1(np.ascontiguousarray(x).tobytes()).hexdigest() == str(fx["digest"])
    idx = fx["model_index"]
    assert np.all(np.diff(idx) > 0) and np.all(fx["model_value"] != 0)
    assert np.all(x >= 0)  # the elastic-net prox with positivity (ElasticNet.cs) never goes negative


def test_per_cycle_figures_follow_the_reference_rules(fx):
    s = fx["stats"]
    assert len(s) == 10
    # MajorCycle.cs:98-100: lambda = max(sidelobe * correction / alpha, lambda *
    # maxLipschitz(psfCut)); on this data the sidelobe term sets the first
    # cycle's value and the Lipschitz floor every later one
    lam = s["lambda_"]
    assert np.all(lam > 0)
    assert lam[0] > lam[1] and np.all(lam[1:] == lam[1])
    # every cycle but the last stops at the iteration cap; the last converges
    assert np.all(s["iterations"][:-1] == 30000) and s["converged"][-1] == 1
    assert np.all(s["converged"][:-1] == 0) and 0 < s["iterations"][-1] < 30000
    # the residual's maximum falls by an order of magnitude after the first cycle
    assert s["max_dirty"][1] < 0.2 * s["max_dirty"][0]
