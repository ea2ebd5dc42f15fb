"""The oracle pinned before it is trusted: golden vectors from the compiled reference, and
independent checks (library routines, closed forms, finite differences, known answers)."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests._util import golden, relerr, scaled_err, synth_deck

CASES = ["h_atom", "h2_jast", "n2_pp", "n2_ci", "n2_cij", "n2_csf", "h4", "pp16", "ae32", "ae32_vmc", "fg3"]
TOL = 1e-10


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_golden(name):
    g, deck = golden(name)
    orc = O.Oracle(deck)
    r = orc.eval(g["pos"])
    assert np.array_equal(r["sign"], g["sign"])
    assert relerr(r["logpsi"], g["logpsi"]) < TOL
    for k in range(len(g["logpsi"])):
        assert scaled_err(r["grad"][k], g["grad"][k]) < TOL
        assert scaled_err(r["lap"][k], g["lap"][k]) < TOL
    assert relerr(r["eparts"][:, 0], g["kinetic"]) < TOL
    assert relerr(r["eparts"][:, 1], g["vloc"]) < TOL
    assert np.max(np.abs(r["eparts"][:, 2] - g["vnonloc"])) <= TOL * max(1.0, np.max(np.abs(g["vnonloc"])))


def test_spline_against_scipy_clamped_spline():
    """The AOSPLINE table is a clamped cubic spline of the normalised radial function."""
    from scipy.interpolate import CubicSpline
    deck, mol = synth_deck("h4")
    orc = O.Oracle(deck)
    c, meta = orc.spline(0)  # contracted S3 shell
    h = meta[0]
    n = 2500
    x = np.arange(n) * h
    y = c[:n, 0]
    # clamped end derivatives as the reference estimates them (one-sided differences)
    cs = CubicSpline(x, y, bc_type=((1, (y[1] - y[0]) / h), (1, (y[-1] - y[-2]) / h)))
    assert np.max(np.abs(cs.c[::-1].T[: n - 1] - c[: n - 1])) < 1e-9 * np.max(np.abs(c[:, 0]))


@pytest.mark.parametrize("shell,l", [(0, 0), (1, 0), (2, 1)])
def test_basis_normalisation_closed_form(shell, l):
    """GAMESSNORM + Riemann renormalisation give int r^(2l+2) f(r)^2 dr = 1."""
    deck, _ = synth_deck("h4")
    c, meta = O.Oracle(deck).spline(shell)
    h = meta[0]
    r = np.arange(2500) * h
    f = c[:2500, 0]
    assert abs(np.sum(h * r ** (2 * l + 2) * f * f) - 1.0) < 1e-5


def test_basis_matches_analytic_gaussian_and_derivatives():
    """Uncontracted S shell on the spline vs the analytic normalised Gaussian; gradient and
    laplacian of every basis function vs central finite differences."""
    deck, mol = synth_deck("ae32")
    orc = O.Oracle(deck)
    rng = np.random.default_rng(3)
    pts = mol.coords[:4] + rng.normal(scale=0.9, size=(4, 3))
    b = orc.basis(pts)
    h = 1e-4
    for d in range(3):
        dp = np.zeros(3)
        dp[d] = h
        fd = (orc.basis(pts + dp)[:, :, 0] - orc.basis(pts - dp)[:, :, 0]) / (2 * h)
        assert scaled_err(b[:, :, 1 + d], fd) < 1e-5
    lap = np.zeros_like(b[:, :, 0])
    for d in range(3):
        dp = np.zeros(3)
        dp[d] = h
        lap += (orc.basis(pts + dp)[:, :, 0] - 2 * b[:, :, 0] + orc.basis(pts - dp)[:, :, 0]) / h ** 2
    assert scaled_err(b[:, :, 4], lap) < 1e-4
    # S shell of atom 0 (first function): N * exp(-a r^2) with N from the closed form
    a = mol.shells[0][1][0][0]
    norm = math.sqrt(2 * 4 * a * math.sqrt(2 * a / math.pi))
    r = np.linalg.norm(pts[0] - mol.coords[0])
    assert abs(b[0, 0, 0] - norm * math.exp(-a * r * r)) < 2e-6 * norm


def test_lu_logdet_inverse_vs_lapack():
    rng = np.random.default_rng(5)
    for n in [1, 2, 7, 32, 64]:
        a = rng.standard_normal((n, n))
        ok, a1, ld, sg = O.transpose_inverse(a)
        s_np, ld_np = np.linalg.slogdet(a)
        assert ok == 1 and sg == s_np and abs(ld - ld_np) < 1e-11 * max(1, abs(ld_np))
        assert np.max(np.abs(a1 - np.linalg.inv(a).T)) < 1e-10 * np.max(np.abs(a1))
    ok, a1, ld, sg = O.transpose_inverse(np.zeros((3, 3)))
    assert ok == 0 and ld == -1e201


def test_sherman_morrison_matches_fresh_inverse_and_det_ratio():
    rng = np.random.default_rng(6)
    n = 16
    a = rng.standard_normal((n, n))
    _, a1, ld, _ = O.transpose_inverse(a)
    for row in [0, 5, n - 1]:
        v = rng.standard_normal(n)
        upd, f = O.inverse_update_row(a1, v, row)
        b = a.copy()
        b[row] = v
        assert np.max(np.abs(upd - np.linalg.inv(b).T)) < 1e-10 * np.max(np.abs(upd))
        assert abs(f - np.linalg.det(a) / np.linalg.det(b)) < 1e-10 * abs(f)


def test_philox_known_answers():
    """Random123 known-answer vectors for Philox4x32-10."""
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    f = 0xFFFFFFFF
    assert O.philox([f, f, f, f], [f, f]) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


@pytest.mark.parametrize("name", ["h4", "n2_pp"])
def test_gradient_laplacian_vs_finite_differences(name):
    """grad Psi/Psi and lap Psi/Psi from the caches vs finite differences of ln|Psi|."""
    g, deck = golden(name)
    orc = O.Oracle(deck)
    pos = g["pos"][:2]
    r = orc.eval(pos)
    h = 1e-4
    ne = pos.shape[1]
    for e in range(ne):
        lap_fd = np.zeros(2)
        gl = np.zeros((2, 3))
        for d in range(3):
            p1 = pos.copy()
            p1[:, e, d] += h
            p2 = pos.copy()
            p2[:, e, d] -= h
            lp = orc.eval(p1)["logpsi"]
            lm = orc.eval(p2)["logpsi"]
            gl[:, d] = (lp - lm) / (2 * h)
            lap_fd += (lp - 2 * r["logpsi"] + lm) / h ** 2
        # lap Psi / Psi = lap ln|Psi| + |grad ln|Psi||^2
        lap_fd += np.sum(gl ** 2, axis=1)
        assert np.max(np.abs(gl - r["grad"][:, e])) < 1e-5 * max(1, np.max(np.abs(gl)))
        assert np.max(np.abs(lap_fd - r["lap"][:, e])) < 1e-3 * max(1, np.max(np.abs(lap_fd)))


def test_electron_electron_cusp_closed_form():
    """CUTOFF_CUSP with unlike coefficient 0.5: d ln Psi / d r12 -> 1/2 as r12 -> 0 (Kato)."""
    g, deck = golden("h2_jast")
    orc = O.Oracle(deck)
    base = np.array([[0.3, -0.2, 0.4], [0.3, -0.2, 0.4]])
    rs = [1e-3, 2e-3]
    slopes = []
    for sgn in (1, -1):
        vals = []
        for r in rs:
            p = base.copy()
            p[1, 2] += sgn * r
            vals.append(orc.eval(p[None])["logpsi"][0])
        slopes.append((vals[1] - vals[0]) / (rs[1] - rs[0]))
    # approaching from +z and -z: the smooth Slater slope cancels in the mean, the cusp stays
    assert abs(0.5 * (slopes[0] + slopes[1]) - 0.5) < 0.02


def test_coulomb_closed_form():
    g, deck = golden("h2_jast")
    orc = O.Oracle(deck)
    p = np.array([[[0.0, 0.0, 0.0], [0.0, 0.0, 2.0]]])
    z = 1.322808191
    r = orc.eval(p)["eparts"][0, 1]
    ions = np.array([[0, 0, -z], [0, 0, z]])
    ei = sum(1 / np.linalg.norm(e - i) for e in p[0] for i in ions)
    assert abs(r - (-ei + 1 / (2 * z) + 1 / 2.0)) < 1e-13


def test_icosahedral_quadrature_integrates_legendre_exactly():
    """The 12-point rule integrates P_l(q.u) over the sphere exactly for l <= 5."""
    fi0 = math.pi / 5
    pts = [(0, 0, 1), (0, 0, -1)]
    for i in range(5):
        pts.append((2 / math.sqrt(5) * math.cos(2 * i * fi0), 2 / math.sqrt(5) * math.sin(2 * i * fi0), 1 / math.sqrt(5)))
        pts.append((2 / math.sqrt(5) * math.cos((2 * i + 1) * fi0), 2 / math.sqrt(5) * math.sin((2 * i + 1) * fi0), -1 / math.sqrt(5)))
    q = np.array(pts)
    from numpy.polynomial import legendre as L
    rng = np.random.default_rng(1)
    for _ in range(5):
        u = rng.standard_normal(3)
        u /= np.linalg.norm(u)
        c = q @ u
        for l in range(6):
            coef = np.zeros(l + 1)
            coef[l] = 1
            assert abs(np.mean(L.legval(c, coef)) - (1.0 if l == 0 else 0.0)) < 1e-14


def test_pp_nonlocal_brute_force():
    """V_NL of one configuration by brute force: move each electron to each quadrature
    point, recompute the whole wavefunction, and apply the Legendre sums."""
    g, deck = golden("n2_pp")
    orc = O.Oracle(deck)
    pos = g["pos"][0].copy()
    base = orc.eval(pos[None])
    # N2 pseudopotential of tests/n2/qw.sys: one nonlocal channel (l=0) and the local one
    terms_nl = [(2 - 2, 6.99537, 31.6972)]
    ions = np.array([[0, 0, -1.60626709], [0, 0, 1.60626709]])
    q = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=float)
    # static cutoff: the reference scans inward in 0.05 steps for |v_0| > 1e-5
    r = 20.0
    cut = 20.0
    while r > 0:
        if abs(sum(c * r ** n * math.exp(-min(a * r * r, 60)) for n, a, c in terms_nl)) > 1e-5:
            cut = r - 0.05
            break
        r -= 0.05
    vnl = 0.0
    for I in ions:
        for e in range(pos.shape[0]):
            d = pos[e] - I
            rr = np.linalg.norm(d)
            if rr >= cut:
                continue
            v0 = sum(c * rr ** n * math.exp(-min(a * rr * rr, 60)) for n, a, c in terms_nl)
            acc = 0
            for k in range(6):
                p = pos.copy()
                p[e] = I + rr * q[k]
                t = orc.eval(p[None])
                ratio = t["sign"][0] * base["sign"][0] * math.exp(t["logpsi"][0] - base["logpsi"][0])
                acc += v0 * ratio / 6.0
            vnl += acc
    # local channel: 1/r term (n=1), r^1 term (n=3), r^0 term (n=2)
    loc = [(1 - 2, 9.235010000000001, 5), (3 - 2, 7.6683, 46.17505), (2 - 2, 7.34486, -30.18894)]
    for I in ions:
        for e in range(pos.shape[0]):
            rr = np.linalg.norm(pos[e] - I)
            vnl += sum(c * rr ** n * math.exp(-min(a * rr * rr, 60)) for n, a, c in loc)
    assert abs(vnl - base["eparts"][0, 2]) < 1e-9 * max(1, abs(vnl))


def test_oracle_vmc_h_atom_known_answer():
    """VMC on the reference's H-atom deck reproduces the GAMESS Hartree-Fock energy
    -0.4998098113 (tests/h/run_test.py:20-23) within 3 sigma."""
    g, deck = golden("h_atom")
    orc = O.Oracle(deck)
    rng = np.random.default_rng(1)
    pos = rng.normal(scale=1.0, size=(256, 1, 3))
    pos = orc.run(pos, "VMC", nblock=2, nstep=20, timestep=0.5, seed=11)["pos"]  # equilibrate
    res = orc.run(pos, "VMC", nblock=20, nstep=50, timestep=0.5, seed=12, nthreads=4)
    e = res["eblock"][:, 0]
    err = e.std(ddof=1) / math.sqrt(len(e))
    assert abs(e.mean() - (-0.4998098113)) < 3 * err + 1e-4
