"""The committed README-problem fixture: regenerating it with the oracle gives
the same numbers, and it agrees with the brute-force dense-Newton optimum."""
import os

import numpy as np

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    return np.load(os.path.join(HERE, "golden", "readme_c1.npz"))


def test_fixture_matches_oracle():
    z = load()
    fw = oracle.splr_bwd(z["M"], z["a"], z["b"], float(z["eta"]), tol=float(z["tol"]))
    assert fw["niter"] == int(z["niter"])
    np.testing.assert_allclose(fw["plan"], z["plan"], rtol=0, atol=1e-15)
    bw = oracle.backward(z["M"], fw["dual"], z["G"], float(z["eta"]), ctx=fw)
    np.testing.assert_allclose(bw["grad"], z["grad"], rtol=0, atol=1e-13)


def test_fixture_near_optimum():
    z = load()
    err = np.abs(z["plan"] - z["plan_opt"]).max() / z["plan_opt"].max()
    assert err < 1e-5
    np.testing.assert_allclose(z["plan"].sum(1), z["a"], atol=1e-8)
