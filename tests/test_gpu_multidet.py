"""Several determinants on the device (k_inv_md, k_sweep_md, k_local<double, true>): the
reference's own N2 CI wave function (tests/n2: 28 determinants over 8 orbitals, BFD
pseudopotential), without and with a Jastrow factor.

- fixed configurations against the reference's golden values (ln|Psi|, sign, gradients,
  laplacians, kinetic, Coulomb and nonlocal energies) at 1e-10;
- VMC and DMC trajectories against the oracle, move by move;
- the reference's known answer (tests/n2/run_test.py:20-23): VMC of the CI wave function gives
  the GAMESS CI energy -19.3080264611 and kinetic energy 12.2497549662 within 3 sigma.
"""
import math
import sys

import numpy as np
import pytest

from mainline_amd import engine as E
from oracle import oracle as O
from tests._util import golden, relerr, scaled_err

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["n2_ci", "n2_cij", "n2_csf"])
def test_multidet_eval_matches_reference_golden(name):
    g, deck = golden(name)
    _say("multidet eval golden", name)
    eng = E.Engine(deck, "fp64")
    r = eng.eval(g["pos"])
    assert np.array_equal(r["sign"], g["sign"])
    assert relerr(r["logpsi"], g["logpsi"]) < TOL
    for k in range(len(g["logpsi"])):
        assert scaled_err(r["grad"][k], g["grad"][k]) < TOL
        assert scaled_err(r["lap"][k], g["lap"][k]) < TOL
    assert relerr(r["eparts"][:, 0], g["kinetic"]) < TOL
    assert relerr(r["eparts"][:, 1], g["vloc"]) < TOL
    assert np.max(np.abs(r["eparts"][:, 2] - g["vnonloc"])) <= TOL * max(1.0, np.max(np.abs(g["vnonloc"])))


@pytest.mark.parametrize("method", ["VMC", "DMC"])
def test_multidet_trajectory_matches_oracle(method):
    g, deck = golden("n2_cij")
    nw = 12
    pos = g["pos"][:nw]
    tau = 0.3 if method == "VMC" else 0.02
    _say("multidet trajectory", method)
    eng = E.Engine(deck, "fp64", seed=7373)
    eng.set_walkers(pos)
    res = eng.run(E.method(method, nw, nblock=2, nstep=5, timestep=tau))
    gp, gw = eng.get_walkers(nw)
    o = O.Oracle(deck).run(pos, method, nblock=2, nstep=5, timestep=tau, seed=7373, step0=0)
    _say("energies", res["eblock"][:, 0], o["eblock"][:, 0])
    assert np.max(np.abs(gp - o["pos"])) < 1e-8
    assert np.max(np.abs(gw - o["weights"])) < 1e-8
    assert np.max(np.abs(res["eblock"][:, :4] - o["eblock"][:, :4])) < 1e-7 * max(1, np.max(np.abs(o["eblock"][:, 0])))


def test_n2_ci_vmc_known_answer():
    g, deck = golden("n2_ci")
    nw = 4096
    eng = E.Engine(deck, "fp64", seed=31)
    pos = np.concatenate([g["pos"]] * (nw // len(g["pos"]) + 1))[:nw]
    pos = pos + np.random.default_rng(4).normal(scale=0.05, size=pos.shape)
    eng.set_walkers(pos)
    eng.run(E.method("VMC", nw, nblock=2, nstep=20, timestep=0.5))
    res = eng.run(E.method("VMC", nw, nblock=20, nstep=10, timestep=0.5))
    e, ke = res["eblock"][:, 0], res["eblock"][:, 1]
    err, kerr = e.std(ddof=1) / math.sqrt(len(e)), ke.std(ddof=1) / math.sqrt(len(ke))
    _say("N2 CI VMC", e.mean(), "+-", err, "kinetic", ke.mean(), "+-", kerr, "acceptance", res["acceptance"])
    assert abs(e.mean() - (-19.3080264611)) < 3 * err + 2e-3
    assert abs(ke.mean() - 12.2497549662) < 3 * kerr + 2e-2
