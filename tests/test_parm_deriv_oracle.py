"""The oracle's linear-parameter derivatives (Wavefunction::getParmDeriv of SLATER-JASTROW with
JASTROW2, and the CSF coefficients of OPTIMIZE_DET) pinned by the reference's own values and by
finite differences in the parameters."""
import numpy as np
import pytest

from oracle import oracle as O
from tests._util import golden

CASES = ["h_atom", "h2_jast", "n2_pp", "n2_cij", "n2_cio", "n2_csfj", "h4", "pp16", "ae32", "fg3"]


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if b.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", CASES)
def test_parm_deriv_matches_reference_golden(name):
    g, deck = golden(name)
    o = O.Oracle(deck)
    # (the reference keeps DETWT weights as logarithms: their values come back within an ulp)
    assert np.allclose(o.parms(), g["parms"], rtol=1e-14, atol=0)
    r = o.parm_deriv(g["pos"])
    assert r["gradient"].shape == g["pd_gradient"].shape
    assert _rel(r["gradient"], g["pd_gradient"]) < 1e-12
    assert _rel(r["gradderiv"], g["pd_gradderiv"]) < 1e-12
    assert _rel(r["val_gradient"], g["pd_valgrad"]) < 1e-12


@pytest.mark.parametrize("name", ["h4", "n2_pp", "n2_cio", "n2_csfj"])
def test_parm_deriv_finite_differences(name):
    """gradient = d ln|Psi|/dp and gradderiv[:3] = d grad ln Psi / dp by central differences of
    the evaluation; gradderiv[3] = d (lap Psi / Psi)/dp minus 2 gradderiv[:3] . grad J_frozen,
    where J_frozen is the part of the Jastrow without parameters (FREEZE): the reference forms
    the cross term with the gradient of the pieces that have parameters and of the determinant
    only (its val_gradient), which the test recovers as the evaluation's gradient minus
    val_gradient."""
    g, deck = golden(name)
    o = O.Oracle(deck)
    pos = g["pos"][:3]
    p0 = o.parms()
    r = o.parm_deriv(pos)
    ev0 = o.eval(pos)
    frozen_grad = ev0["grad"] - r["val_gradient"]
    for q in range(len(p0)):
        h = 1e-5 * max(1.0, abs(p0[q]))
        out = []
        for sgn in (1, -1):
            p = p0.copy()
            p[q] += sgn * h
            o.set_parms(p)
            out.append(o.eval(pos))
        o.set_parms(p0)
        d_logpsi = (out[0]["logpsi"] - out[1]["logpsi"]) / (2 * h)
        d_grad = (out[0]["grad"] - out[1]["grad"]) / (2 * h)
        d_lap = (out[0]["lap"] - out[1]["lap"]) / (2 * h)
        gd = r["gradderiv"][:, q]
        assert np.allclose(r["gradient"][:, q], d_logpsi, rtol=1e-6, atol=1e-7)
        assert np.allclose(gd[..., :3], d_grad, rtol=1e-6, atol=1e-6)
        expect = d_lap - 2 * np.sum(gd[..., :3] * frozen_grad, axis=-1)
        assert np.allclose(gd[..., 3], expect, rtol=1e-5, atol=1e-5)


def test_parm_hessian_and_val_gradient_identities():
    """val_gradient = grad ln D + sum_p c_p gradderiv[:3] (parameters' pieces are linear in c)."""
    g, deck = golden("pp16")
    o = O.Oracle(deck)
    pos = g["pos"][:2]
    r = o.parm_deriv(pos)
    c = o.parms()
    lin = np.einsum("p,kpej->kej", c, r["gradderiv"][..., :3])
    # determinant part: the same quantity with every parameter set to zero
    o.set_parms(np.zeros_like(c))
    r0 = o.parm_deriv(pos)
    o.set_parms(c)
    assert np.allclose(r["val_gradient"], r0["val_gradient"] + lin, rtol=1e-12, atol=1e-12)
