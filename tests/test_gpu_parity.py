"""Parity of the HIP engine with the oracle and the reference's golden vectors (MI355X)."""
import math
import sys

import numpy as np
import pytest

from mainline_amd import engine as E
from mainline_amd import synth
from oracle import oracle as O
from tests._util import golden, relerr, scaled_err, synth_deck

pytestmark = pytest.mark.gpu

CASES = ["h_atom", "h2_jast", "n2_pp", "n2_cen", "h4", "pp16", "ae32", "ae32_vmc", "fg3"]
TOL = 1e-10  # FP64 path: relative tolerance against the reference's FP64 evaluation


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", CASES)
def test_fp64_eval_matches_reference_golden(name):
    g, deck = golden(name)
    _say("eval golden", name, g["pos"].shape)
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


def test_blas_mo_deck_matches_reference_golden():
    """ORBITALS { BLAS_MO }: the reference evaluates the same orbitals densely and leaves out the
    basis functions whose value at the point is below 1e-8 (src/orbitals/MO_matrix_blas.h:
    411-491); the engine evaluates them exactly (as CUTOFF_MO), so the values agree to the size
    of the dropped terms (measured: ln|Psi| 7e-10, gradients 2e-8, laplacians 8.5e-7, kinetic
    2.6e-7 relative), against 1e-10 for the other decks."""
    g, deck = golden("n2_blas")
    eng = E.Engine(deck, "fp64")
    r = eng.eval(g["pos"])
    err = dict(logpsi=relerr(r["logpsi"], g["logpsi"]),
               grad=max(scaled_err(r["grad"][k], g["grad"][k]) for k in range(len(g["logpsi"]))),
               lap=max(scaled_err(r["lap"][k], g["lap"][k]) for k in range(len(g["logpsi"]))),
               kin=relerr(r["eparts"][:, 0], g["kinetic"]),
               vnl=float(np.max(np.abs(r["eparts"][:, 2] - g["vnonloc"])) / max(1.0, np.max(np.abs(g["vnonloc"])))))
    _say("blas_mo", {k: "%.1e" % v for k, v in err.items()})
    assert np.array_equal(r["sign"], g["sign"])
    tol = dict(logpsi=1e-8, grad=1e-7, lap=5e-6, kin=2e-6, vnl=1e-7)
    for k, v in err.items():
        assert v < tol[k], (k, v)
    assert relerr(r["eparts"][:, 1], g["vloc"]) < TOL


@pytest.mark.parametrize("name", ["h4", "pp16", "ae32"])
def test_fp64_eval_matches_oracle_random_rotation_and_stochastic_pp(name):
    deck, mol = synth_deck(name)
    _say("eval oracle", name)
    eng = E.Engine(deck, "fp64")
    orc = O.Oracle(deck)
    pos = synth.random_configs(mol, 6, seed=77)
    rng = np.random.default_rng(9)
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    rot = q.T.copy()
    rand = rng.uniform(size=(6, mol.natoms, mol.nelec))
    for mode in (0, 1):
        a = eng.eval(pos, rot=rot, pp_mode=mode, pp_rand=rand)
        b = orc.eval(pos, rot=rot, pp_mode=mode, pp_rand=rand)
        assert np.array_equal(a["sign"], b["sign"])
        assert relerr(a["logpsi"], b["logpsi"]) < TOL
        assert scaled_err(a["grad"], b["grad"]) < TOL
        assert scaled_err(a["lap"], b["lap"]) < TOL
        assert relerr(a["eparts"][:, :2], b["eparts"][:, :2]) < TOL
        assert np.max(np.abs(a["eparts"][:, 2] - b["eparts"][:, 2])) < TOL * max(1, np.max(np.abs(b["eparts"][:, 2])))


def test_singular_determinant_is_flagged_like_the_reference():
    """Two same-spin electrons on one point: det = 0, ln|det| = -1e201 (log_value(0))."""
    deck, mol = synth_deck("h4")
    eng = E.Engine(deck, "fp64")
    pos = synth.random_configs(mol, 2, seed=5)
    pos[1, 1] = pos[1, 0]
    r = eng.eval(pos)
    assert r["logpsi"][1] < -1e200
    assert np.isfinite(r["logpsi"][0])


@pytest.mark.parametrize("name,method", [("h4", "VMC"), ("h4", "DMC"), ("pp16", "DMC"), ("pp16", "VMC")])
def test_trajectory_matches_oracle(name, method):
    """Same seed, same counter-based random streams: the device run and the oracle's CPU run
    take the same moves, so positions, weights and block energies agree to round-off."""
    deck, mol = synth_deck(name)
    _say("trajectory", name, method)
    nw = 16
    pos = synth.random_configs(mol, nw, seed=21)
    tau = 0.3 if name == "h4" else 0.02
    eng = E.Engine(deck, "fp64", seed=4242)
    eng.set_walkers(pos)
    m = E.method(method, nw, nblock=2, nstep=5, timestep=tau)
    res = eng.run(m)
    gp, gw = eng.get_walkers(nw)
    orc = O.Oracle(deck)
    o = orc.run(pos, method, nblock=2, nstep=5, timestep=tau, seed=4242, step0=0)
    assert np.max(np.abs(gp - o["pos"])) < 1e-8
    assert np.max(np.abs(gw - o["weights"])) < 1e-8
    assert np.max(np.abs(res["eblock"][:, :4] - o["eblock"][:, :4])) < 1e-7 * max(1, np.max(np.abs(o["eblock"][:, 0])))


def test_h_atom_vmc_and_dmc_known_answers():
    """tests/h: VMC reproduces the GAMESS Hartree-Fock energy -0.4998098113 and DMC the
    exact -0.5 (tests/h/run_test.py:20-23,57), each within 3 sigma."""
    g, deck = golden("h_atom")
    eng = E.Engine(deck, "fp64", seed=99)
    rng = np.random.default_rng(2)
    nw = 2048
    eng.set_walkers(rng.normal(size=(nw, 1, 3)))
    eng.run(E.method("VMC", nw, nblock=2, nstep=20, timestep=0.5))
    res = eng.run(E.method("VMC", nw, nblock=20, nstep=20, timestep=0.5))
    e = res["eblock"][:, 0]
    err = e.std(ddof=1) / math.sqrt(len(e))
    _say("H VMC", e.mean(), err)
    assert abs(e.mean() - (-0.4998098113)) < 3 * err + 1e-5
    res = eng.run(E.method("DMC", nw, nblock=2, nstep=200, timestep=0.005))
    res = eng.run(E.method("DMC", nw, nblock=20, nstep=100, timestep=0.005))
    e = res["eblock"][:, 0]
    err = e.std(ddof=1) / math.sqrt(len(e))
    _say("H DMC", e.mean(), err)
    assert abs(e.mean() - (-0.5)) < 3 * err + 2e-4


@pytest.mark.parametrize("prec", ["fp64", "mixed_fp32"])
def test_graph_replayed_dmc_equals_direct_launches(prec, monkeypatch):
    """DMC feedback intervals recorded once into a HIP graph and replayed (several replays per
    block, host synchronisation between blocks) give exactly the directly launched result."""
    deck, mol = synth_deck("pp16")
    nw = 64
    pos = synth.random_configs(mol, nw, seed=23)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("QW_GRAPHS", mode)
        eng = E.Engine(deck, prec, seed=5151)
        eng.set_walkers(pos)
        res = eng.run(E.method("DMC", nw, nblock=2, nstep=15, timestep=0.01))
        p, w = eng.get_walkers(nw)
        out[mode] = (res["eblock"], p, w)
        eng.close()
    _say("graph vs direct", out["0"][0][:, 0], out["1"][0][:, 0])
    for a, b in zip(out["0"], out["1"]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec,name,choices", [
    ("mixed_fp32", "ae32", ["4", "8", "16"]),
    ("mixed_fp32", "pp16", ["8", "104", "108", "116"]),  # the LDS layout fits this system
    ("fp64", "ae32", ["2", "4", "8"]),
])
def test_sweep_launch_choices_give_identical_results(prec, name, choices, monkeypatch):
    """Every sweep launch the tuner may pick (coefficient loads in flight; coefficients read
    from L2 or staged in LDS for eight walkers per block) gives bit-identical VMC and DMC."""
    deck, mol = synth_deck(name)
    nw = 40  # not a multiple of the walkers per block
    pos = synth.random_configs(mol, nw, seed=29)
    out = {}
    for c in choices:
        monkeypatch.setenv("QW_SWEEP", c)
        eng = E.Engine(deck, prec, seed=777)
        assert eng.tuning()["sweep"] == int(c)
        eng.set_walkers(pos)
        rv = eng.run(E.method("VMC", nw, nblock=1, nstep=2, timestep=0.3))
        rd = eng.run(E.method("DMC", nw, nblock=2, nstep=4, timestep=0.01))
        p, w = eng.get_walkers(nw)
        out[c] = (rv["eblock"], rd["eblock"], p, w)
        eng.close()
        _say("sweep", prec, name, c, out[c][1][:, 0])
    ref = out[choices[0]]
    for c in choices[1:]:
        for a, b in zip(ref, out[c]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", ["fp64", "mixed_fp32"])
def test_block_and_row_inversions_are_bitwise_equal(prec, monkeypatch):
    """The 8x8-block inversion (default for 33..64 electrons per spin) and the row-per-lane
    inversion give the same log|psi|, inverse-derived gradients and DMC trajectory bit for bit."""
    deck, mol = synth_deck("ae32")
    nw = 24
    pos = synth.random_configs(mol, nw, seed=31)
    out = {}
    for mode in ("rows", "blocks"):
        monkeypatch.setenv("QW_INV", mode)
        eng = E.Engine(deck, prec, seed=4242)
        r = eng.eval(pos)
        eng.set_walkers(pos)
        rd = eng.run(E.method("DMC", nw, nblock=1, nstep=3, timestep=0.01))
        p, w = eng.get_walkers(nw)
        out[mode] = (r["logpsi"], r["sign"], r["grad"], r["lap"], r["eparts"], rd["eblock"], p, w)
        eng.close()
    _say("inversions", prec, out["rows"][0][:3], out["blocks"][0][:3])
    for a, b in zip(out["rows"], out["blocks"]):
        assert np.array_equal(a, b)


def test_recorded_dmc_interval_follows_walker_reallocation():
    """A DMC interval recorded into a graph holds the walker buffers' addresses.  A fixed-
    configuration evaluation with more configurations than walkers reallocates those buffers;
    the next DMC run on the original population must record the interval again and take
    exactly the moves of an engine that never reallocated."""
    deck, mol = synth_deck("h4")
    nw = 64
    pos = synth.random_configs(mol, nw, seed=31)
    more = synth.random_configs(mol, 2 * nw, seed=32)
    m = E.method("DMC", nw, nblock=1, nstep=6, timestep=0.2)
    out = []
    for realloc in (False, True):
        _say("graph after reallocation:", realloc)
        eng = E.Engine(deck, "fp64", seed=515)
        assert eng.tuning()["graphs"] == 1
        eng.set_walkers(pos)
        eng.run(m)
        if realloc:
            eng.eval(more)
        eng.set_walkers(pos)
        r = eng.run(m)
        p, w = eng.get_walkers(nw)
        out.append((p, w, r["eblock"]))
        eng.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])
