"""Casula T-moves on the device (k_local's negative nonlocal sums + k_tmove) against the
oracle's restatement of Dmc_method::doTmove (src/method/Dmc_method.cpp:850-906): the same
counter-based random streams give the same DMC trajectory, move by move, for TMOVES (one
draw per walker) and TMOVESSC (one draw per electron), with one determinant and with the
reference's 28-determinant N2 CI expansion; graph replay equals direct launches;
the mixed path tracks FP64 with common random numbers."""
import sys

import numpy as np
import pytest

from mainline_amd import engine as E
from mainline_amd import synth
from oracle import oracle as O
from tests._util import golden, synth_deck

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(name, nw):
    if name in ("n2_pp", "n2_cij"):
        g, deck = golden(name)
        pos = np.concatenate([g["pos"]] * (nw // len(g["pos"]) + 1))[:nw]
        return deck, pos, 0.05
    deck, mol = synth_deck(name)
    return deck, synth.random_configs(mol, nw, seed=31), 0.02


@pytest.mark.parametrize("tm", [1, 2])
@pytest.mark.parametrize("name", ["n2_pp", "pp16", "n2_cij"])
def test_tmove_trajectory_matches_oracle(name, tm):
    nw = 16
    deck, pos, tau = _case(name, nw)
    _say("tmoves trajectory", name, tm)
    out = {}
    for t in (0, tm):
        eng = E.Engine(deck, "fp64", seed=6161)
        eng.set_walkers(pos)
        res = eng.run(E.method("DMC", nw, nblock=2, nstep=5, timestep=tau, tmoves=t))
        p, w = eng.get_walkers(nw)
        out[t] = (res["eblock"], p, w)
        eng.close()
    eb, gp, gw = out[tm]
    o = O.Oracle(deck).run(pos, "DMC", nblock=2, nstep=5, timestep=tau, seed=6161, step0=0, tmoves=tm)
    moved = np.max(np.abs(gp - out[0][1]))
    _say("max position change from T-moves", moved, "energies", eb[:, 0], o["eblock"][:, 0])
    assert moved > 1e-3  # the T-moves did move electrons
    assert np.max(np.abs(gp - o["pos"])) < 1e-8
    assert np.max(np.abs(gw - o["weights"])) < 1e-8
    assert np.max(np.abs(eb[:, :4] - o["eblock"][:, :4])) < 1e-7 * max(1, np.max(np.abs(o["eblock"][:, 0])))


@pytest.mark.parametrize("prec", ["fp64", "mixed_fp32"])
def test_tmove_graph_replay_equals_direct_launches(prec, monkeypatch):
    deck, pos, tau = _case("pp16", 64)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("QW_GRAPHS", mode)
        eng = E.Engine(deck, prec, seed=5252)
        eng.set_walkers(pos)
        res = eng.run(E.method("DMC", 64, nblock=2, nstep=15, timestep=tau, tmoves=2))
        p, w = eng.get_walkers(64)
        out[mode] = (res["eblock"], p, w)
        eng.close()
    _say("tmoves graph vs direct", out["0"][0][:, 0], out["1"][0][:, 0])
    for a, b in zip(out["0"], out["1"]):
        assert np.array_equal(a, b)


def test_mixed_tmoves_track_fp64():
    """Common random numbers: the FP32 path takes the same T-moves as FP64 until a decision
    flips, so the block energies of a short run agree closely."""
    deck, pos, tau = _case("n2_pp", 256)
    ref = {}
    for prec in ("fp64", "mixed_fp32"):
        eng = E.Engine(deck, prec, seed=919)
        eng.set_walkers(pos)
        ref[prec] = eng.run(E.method("DMC", 256, nblock=3, nstep=10, timestep=tau, tmoves=2))["eblock"][:, 0]
        eng.close()
    d = np.abs(ref["mixed_fp32"] - ref["fp64"])
    _say("tmoves dmc fp64", ref["fp64"], "mixed", ref["mixed_fp32"], "max |dE|", d.max())
    assert d.max() / np.abs(ref["fp64"]).max() < 1e-3
