"""More than 64 electrons of one spin (80 + 80, and 144 + 144): one wavefront per walker with
the inverse in device memory, the register-grid recompute (up to 128 of a spin) and the panel
recompute (above), against the FP64 oracle."""
import sys

import numpy as np
import pytest

from mainline_amd import engine as E
from mainline_amd import synth
from oracle import oracle as O
from tests._util import relerr, scaled_err, synth_deck

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_big_fp64_eval_matches_oracle():
    deck, mol = synth_deck("ae40")
    _say("big eval", mol.nelec, "electrons", mol.nbasis, "functions")
    eng = E.Engine(deck, "fp64")
    orc = O.Oracle(deck)
    pos = synth.random_configs(mol, 4, seed=41)
    a = eng.eval(pos)
    b = orc.eval(pos)
    assert np.array_equal(a["sign"], b["sign"])
    assert relerr(a["logpsi"], b["logpsi"]) < TOL
    assert scaled_err(a["grad"], b["grad"]) < 1e-9
    assert scaled_err(a["lap"], b["lap"]) < 1e-9
    assert relerr(a["eparts"][:, :2], b["eparts"][:, :2]) < TOL


def test_big_fp64_trajectory_matches_oracle():
    deck, mol = synth_deck("ae40")
    nw = 8
    pos = synth.random_configs(mol, nw, seed=42)
    eng = E.Engine(deck, "fp64", seed=777)
    eng.set_walkers(pos)
    res = eng.run(E.method("DMC", nw, nblock=1, nstep=2, timestep=0.01))
    gp, gw = eng.get_walkers(nw)
    o = O.Oracle(deck).run(pos, "DMC", nblock=1, nstep=2, timestep=0.01, seed=777, step0=0)
    _say("big trajectory max |dpos|", np.max(np.abs(gp - o["pos"])))
    assert np.max(np.abs(gp - o["pos"])) < 1e-8
    assert np.max(np.abs(gw - o["weights"])) < 1e-8
    assert abs(res["eblock"][0, 0] - o["eblock"][0, 0]) < 1e-7 * abs(o["eblock"][0, 0])


def test_big_mixed_fp32_eval_and_run():
    deck, mol = synth_deck("ae40")
    eng = E.Engine(deck, "mixed_fp32")
    pos = synth.random_configs(mol, 4, seed=43)
    a = eng.eval(pos)
    b = O.Oracle(deck).eval(pos)
    err = np.max(np.abs(a["logpsi"] - b["logpsi"]))
    kin = np.max(np.abs(a["eparts"][:, 0] - b["eparts"][:, 0]) / np.abs(b["eparts"][:, 0]))
    _say("big mixed logpsi abs err", err, "kinetic rel err", kin)
    assert np.array_equal(a["sign"], b["sign"])
    assert err < 1e-3 and kin < 5e-3
    eng.set_walkers(synth.random_configs(mol, 16, seed=44))
    res = eng.run(E.method("DMC", 16, nblock=1, nstep=3, timestep=0.01))
    assert np.isfinite(res["eblock"][0, 0]) and res["acceptance"] > 0.5


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("mixed_fp32", 2e-3)])
def test_above_128_per_spin_eval_matches_oracle(prec, tol):
    """144 electrons of each spin: the panel-by-panel Gauss-Jordan recompute (FP64 arithmetic
    on the FP64 or FP32 MO cache) gives ln|Psi|, sign, gradients and energies of the oracle."""
    deck, mol = synth_deck("ae72")
    _say("ae72 eval", prec, mol.nelec, "electrons", mol.nbasis, "functions")
    eng = E.Engine(deck, prec)
    pos = synth.random_configs(mol, 3, seed=72)
    a = eng.eval(pos)
    b = O.Oracle(deck).eval(pos)
    err = dict(logpsi=np.max(np.abs(a["logpsi"] - b["logpsi"]) / np.abs(b["logpsi"])),
               grad=max(scaled_err(a["grad"][k], b["grad"][k]) for k in range(3)),
               kin=relerr(a["eparts"][:, 0], b["eparts"][:, 0]))
    _say("ae72", prec, {k: "%.2e" % v for k, v in err.items()})
    assert np.array_equal(a["sign"], b["sign"])
    for k, v in err.items():
        assert v < tol, (k, v)


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("mixed_fp32", 2e-3)])
def test_panel_recompute_partial_panel_matches_oracle(prec, tol):
    """130 electrons of each spin: the last panel holds 2 columns, the last row group of the
    whole-matrix passes 2 rows, and the MO rows (5 x 130 values) and output rows are not
    16-byte aligned, so the recompute takes its scalar-access paths."""
    deck, mol = synth_deck("ae65")
    _say("ae65 eval", prec, mol.nelec, "electrons")
    eng = E.Engine(deck, prec)
    pos = synth.random_configs(mol, 2, seed=65)
    a = eng.eval(pos)
    b = O.Oracle(deck).eval(pos)
    err = dict(logpsi=np.max(np.abs(a["logpsi"] - b["logpsi"]) / np.abs(b["logpsi"])),
               grad=max(scaled_err(a["grad"][k], b["grad"][k]) for k in range(len(pos))),
               kin=relerr(a["eparts"][:, 0], b["eparts"][:, 0]))
    _say("ae65", prec, {k: "%.2e" % v for k, v in err.items()})
    assert np.array_equal(a["sign"], b["sign"])
    for k, v in err.items():
        assert v < tol, (k, v)


def test_above_128_per_spin_trajectory_matches_oracle():
    deck, mol = synth_deck("ae72")
    nw = 4
    pos = synth.random_configs(mol, nw, seed=73)
    eng = E.Engine(deck, "fp64", seed=778)
    eng.set_walkers(pos)
    res = eng.run(E.method("DMC", nw, nblock=1, nstep=1, timestep=0.01))
    gp, gw = eng.get_walkers(nw)
    o = O.Oracle(deck).run(pos, "DMC", nblock=1, nstep=1, timestep=0.01, seed=778, step0=0, nthreads=4)
    _say("ae72 trajectory max |dpos|", np.max(np.abs(gp - o["pos"])))
    assert np.max(np.abs(gp - o["pos"])) < 1e-8
    assert np.max(np.abs(gw - o["weights"])) < 1e-8
    assert abs(res["eblock"][0, 0] - o["eblock"][0, 0]) < 1e-7 * abs(o["eblock"][0, 0])
