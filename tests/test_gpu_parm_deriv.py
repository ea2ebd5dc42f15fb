"""Device linear-parameter derivatives (qw_parm_deriv, Wavefunction::getParmDeriv) against the
reference's own values and against the oracle."""
import sys

import numpy as np
import pytest

from mainline_amd import engine as E
from mainline_amd import synth
from oracle import oracle as O
from tests._util import golden, synth_deck

pytestmark = pytest.mark.gpu

TOL = 1e-10  # FP64 path, relative to the largest magnitude of each quantity


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if b.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", ["h_atom", "h2_jast", "n2_pp", "h4", "pp16", "ae32", "fg3", "n2_cij", "n2_cio",
                                  "n2_csfj"])
def test_fp64_parm_deriv_matches_reference_golden(name):
    """Including several determinants (n2_cij) and OPTIMIZE_DET's CSF coefficients, alone
    (n2_cio) and before the Jastrow parameters (n2_csfj)."""
    g, deck = golden(name)
    _say("parm deriv golden", name, g["pos"].shape)
    eng = E.Engine(deck, "fp64")
    # (the reference keeps DETWT weights as logarithms: their values come back within an ulp)
    assert np.allclose(eng.parms(), g["parms"], rtol=1e-14, atol=0)
    r = eng.parm_deriv(g["pos"])
    eng.close()
    errs = (_rel(r["gradient"], g["pd_gradient"]), _rel(r["gradderiv"], g["pd_gradderiv"]),
            _rel(r["val_gradient"], g["pd_valgrad"]))
    _say("  rel errors", errs)
    assert max(errs) < TOL
    gg = r["gradient"]
    hh = gg[:, :, None] * gg[:, None, :]
    ns = CSF_PARMS.get(name, 0)
    hh[:, :ns, :ns] = 0.0  # the determinant's own block is zero
    assert np.allclose(r["hessian"], hh, rtol=0, atol=0)


# CSF coefficients (OPTIMIZE_DET) at the front of the parameter list of these fixtures
CSF_PARMS = {"n2_cio": 27, "n2_csfj": 11}


@pytest.mark.parametrize("name", ["pp16", "ae40"])
def test_fp64_parm_deriv_matches_oracle(name):
    """Random configurations, including the block-per-walker path (80 electrons per spin)."""
    deck, mol = synth_deck(name)
    pos = synth.random_configs(mol, 6, seed=71)
    _say("parm deriv oracle", name, pos.shape)
    eng = E.Engine(deck, "fp64")
    r = eng.parm_deriv(pos)
    eng.close()
    o = O.Oracle(deck).parm_deriv(pos)
    errs = (_rel(r["gradient"], o["gradient"]), _rel(r["gradderiv"], o["gradderiv"]),
            _rel(r["val_gradient"], o["val_gradient"]))
    _say("  rel errors", errs)
    assert max(errs) < TOL


def test_mixed_parm_deriv_close_to_oracle():
    """Mixed engine: FP64 Jastrow sums, FP32 MO cache and inverse in the determinant gradient."""
    deck, mol = synth_deck("pp16")
    pos = synth.random_configs(mol, 4, seed=72)
    eng = E.Engine(deck, "mixed_fp32")
    r = eng.parm_deriv(pos)
    eng.close()
    o = O.Oracle(deck).parm_deriv(pos)
    _say("mixed rel", _rel(r["gradient"], o["gradient"]), _rel(r["val_gradient"], o["val_gradient"]))
    assert _rel(r["gradient"], o["gradient"]) < 1e-12  # Jastrow sums only: FP64 throughout
    assert _rel(r["val_gradient"], o["val_gradient"]) < 1e-3
    assert _rel(r["gradderiv"], o["gradderiv"]) < 1e-3
