"""The committed golden fixture (literal six-deep-loop results on a 3x3x3 mesh) against the
factorised oracle, and the model generator against the fixture's inputs."""
import numpy as np

from oracle import oracle as O
from tests.golden_fixture import load


def test_factorised_oracle_reproduces_golden():
    m, qm, ph, d = load()
    inp = O.Inputs(m, qm, ph)
    lw = O.lw3_gauss_all(inp, qm.irr_idx, float(d["sigma"]))
    assert np.allclose(lw, d["lw"], rtol=1e-12, atol=1e-14 * np.abs(d["lw"]).max())
    _, W = O.lw3_gauss(inp, int(qm.irr_idx[1]), float(d["sigma"]), want_w=True)
    assert np.allclose(W, d["W_q0_1"], rtol=1e-11, atol=1e-13 * np.abs(W).max())


def test_generator_still_produces_fixture_inputs(rocksalt):
    m, qm, ph, d = load()
    assert np.array_equal(rocksalt.fc3.atm_indx, d["atm_indx3"])
    assert np.allclose(rocksalt.fc3.fc, d["fc3"], rtol=0, atol=1e-12)
