"""T-moves (Casula) in the oracle, pinned before the device is compared with it.

- The candidate list (electron, vxx < 0, displacement onto a quadrature point, in the order the
  reference builds it) and the nonlocal energy left without those terms equal the reference's
  own Pseudopotential::calcNonlocTmove on the golden configurations (src/system/
  Pseudopotential.cpp:371-389,752-763).
- The two lists add back to the full nonlocal energy (the identity Dmc_method::doTmove relies on,
  src/method/Dmc_method.cpp:861-866).
- Every displacement keeps the electron on the sphere around one atom (a rotated quadrature
  point at the same radius).
- The selection rule of doTmove (src/method/Dmc_method.cpp:867-905) against brute-force
  frequencies: TMOVES moves candidate m with probability tau|vxx_m| / (1 + tau sum|vxx|),
  TMOVESSC moves electron e with probability tau S_e / (1 + tau S_e), S_e its own sum.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests._util import golden

PP_CASES = ["n2_pp", "n2_ci", "pp16"]


@pytest.mark.parametrize("name", PP_CASES + ["h_atom", "ae32"])
def test_candidates_match_the_reference(name):
    g, deck = golden(name)
    orc = O.Oracle(deck)
    ref = g["tmove"]
    for k in range(len(g["logpsi"])):
        t = orc.tmoves(g["pos"][k])
        rk = ref[ref[:, 0] == k]
        assert len(t["e"]) == len(rk)
        assert np.array_equal(t["e"], rk[:, 1].astype(np.int64))
        assert np.max(np.abs(t["vxx"] - rk[:, 2]), initial=0) <= 1e-10 * max(1.0, np.max(np.abs(rk[:, 2]), initial=0))
        assert np.max(np.abs(t["disp"] - rk[:, 3:6]), initial=0) <= 1e-12 * max(1.0, np.max(np.abs(rk[:, 3:6]), initial=0))
        assert abs(t["vnl"] - g["vnl_tmove"][k]) <= 1e-10 * max(1.0, abs(g["vnl_tmove"][k]))
        # the negative terms restore the full nonlocal energy
        assert abs(t["vnl"] + t["vxx"].sum() - g["vnonloc"][k]) <= 1e-10 * max(1.0, abs(g["vnonloc"][k]))
    if name in PP_CASES:
        assert len(ref) > 0 and np.all(ref[:, 2] < 0)


@pytest.mark.parametrize("name", PP_CASES)
def test_displacements_stay_on_the_atom_sphere(name):
    g, deck = golden(name)
    orc = O.Oracle(deck)
    atoms = _atoms(deck, orc.natoms)
    for k in range(len(g["logpsi"])):
        t = orc.tmoves(g["pos"][k])
        for e, d in zip(t["e"], t["disp"]):
            r0 = g["pos"][k][e]
            r1 = r0 + d
            dr = np.abs(np.linalg.norm(r0 - atoms, axis=1) - np.linalg.norm(r1 - atoms, axis=1))
            assert dr.min() < 1e-12


def _atoms(deck, na):
    """Atom positions from the deck's SYSTEM section (ATOM { label Z COOR x y z })."""
    import re
    import os
    text = open(deck).read()
    inc = re.findall(r"include\s+(\S+)", text, flags=re.I)
    for f in inc:
        text += open(os.path.join(os.path.dirname(deck), f)).read()
    pos = [list(map(float, m)) for m in re.findall(r"coor\s+(\S+)\s+(\S+)\s+(\S+)", text, flags=re.I)]
    assert len(pos) >= na
    return np.array(pos[:na])


@pytest.mark.parametrize("sc", [0, 1])
def test_selection_frequencies_brute_force(sc):
    g, deck = golden("n2_pp")
    orc = O.Oracle(deck)
    pos = g["pos"][4]  # 19 candidates on 8 electrons
    t = orc.tmoves(pos)
    assert len(np.unique(t["e"])) >= 6
    tau = 0.5  # large, so that moves are frequent
    s = -tau * t["vxx"]
    ne = orc.ne
    trials = 4000
    counts = np.zeros(len(s))
    ecount = np.zeros(ne)
    for step in range(trials):
        new, n = orc.tmove_step(pos, tau, sc, seed=77, walker=3, step=step)
        moved = np.where(np.any(new != pos, axis=1))[0]
        assert n == len(moved)
        if not sc:
            assert n <= 1
        for e in moved:
            ecount[e] += 1
            cand = np.where(t["e"] == e)[0]
            hit = [c for c in cand if np.allclose(new[e], pos[e] + t["disp"][c], atol=1e-14, rtol=0)]
            assert len(hit) == 1
            counts[hit[0]] += 1
    if not sc:
        p = s / (1 + s.sum())
        assert abs(counts.sum() / trials - p.sum()) < 5 * np.sqrt(p.sum() * (1 - p.sum()) / trials)
        big = p > 0.005
        assert np.all(np.abs(counts[big] / trials - p[big]) < 5 * np.sqrt(p[big] * (1 - p[big]) / trials))
    else:
        for e in range(ne):
            se = s[t["e"] == e].sum()
            pe = se / (1 + se)
            assert abs(ecount[e] / trials - pe) <= 5 * np.sqrt(max(pe * (1 - pe), 1e-4) / trials)
            cand = np.where(t["e"] == e)[0]
            if ecount[e] > 200:
                # within an electron, candidates in proportion to their |vxx|
                q = s[cand] / se
                f = counts[cand] / ecount[e]
                assert np.all(np.abs(f - q) < 5 * np.sqrt(q * (1 - q) / ecount[e]) + 1e-3)


def test_no_moves_without_pseudopotential():
    g, deck = golden("h2_jast")
    orc = O.Oracle(deck)
    new, n = orc.tmove_step(g["pos"][0], 0.5, 1, seed=1, walker=0, step=0)
    assert n == 0 and np.array_equal(new, g["pos"][0])
