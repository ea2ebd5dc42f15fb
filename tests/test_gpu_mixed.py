"""Mixed-precision paths against the FP64 oracle: errors per quantity at fixed configurations,
and correlated-sampling agreement of VMC/DMC energies with the FP64 engine (same random
streams, so the trajectories stay together until a precision-dependent decision flips)."""
import sys

import numpy as np
import pytest

from mainline_amd import engine as E
from mainline_amd import synth
from oracle import oracle as O
from tests._util import synth_deck

pytestmark = pytest.mark.gpu

# Tolerances against FP64 at fixed configurations (random Gaussian-cloud configurations of
# the synthetic molecules).  log|Psi| is compared as an absolute error (it is a sum of
# O(N) logarithms); gradients and laplacians as max error over max magnitude per config;
# energies relative.  fp32: FP32 LCAO + FP32 inverse; bf16: 3-product split; bf16x6: 6 products.
TOL = {
    "mixed_fp32": dict(logpsi=2e-4, grad=2e-4, lap=2e-3, kin=2e-3, vloc=1e-9),
    "mixed_bf16": dict(logpsi=2e-3, grad=2e-3, lap=2e-2, kin=2e-2, vloc=1e-9),
    "mixed_bf16x6": dict(logpsi=2e-4, grad=2e-4, lap=2e-3, kin=2e-3, vloc=1e-9),
}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", ["mixed_fp32", "mixed_bf16", "mixed_bf16x6"])
@pytest.mark.parametrize("name", ["h4", "pp16", "ae32"])
def test_mixed_eval_errors_against_fp64_oracle(name, prec):
    deck, mol = synth_deck(name)
    eng = E.Engine(deck, prec)
    orc = O.Oracle(deck)
    pos = synth.random_configs(mol, 8, seed=31)
    a = eng.eval(pos)
    b = orc.eval(pos)
    t = TOL[prec]
    err = dict(
        logpsi=np.max(np.abs(a["logpsi"] - b["logpsi"])),
        grad=max(np.max(np.abs(a["grad"][k] - b["grad"][k])) / np.max(np.abs(b["grad"][k])) for k in range(8)),
        lap=max(np.max(np.abs(a["lap"][k] - b["lap"][k])) / np.max(np.abs(b["lap"][k])) for k in range(8)),
        kin=np.max(np.abs(a["eparts"][:, 0] - b["eparts"][:, 0]) / np.abs(b["eparts"][:, 0])),
        vloc=np.max(np.abs(a["eparts"][:, 1] - b["eparts"][:, 1]) / np.abs(b["eparts"][:, 1])),
    )
    _say("mixed-vs-fp64", name, prec, {k: "%.2e" % v for k, v in err.items()})
    assert np.array_equal(a["sign"], b["sign"])
    for k, v in err.items():
        assert v < t[k], (k, v)


def _dmc_blocks(deck, mol, prec, nw=256, nblock=4):
    pos = synth.random_configs(mol, nw, seed=8)
    eng = E.Engine(deck, prec, seed=777)
    eng.set_walkers(pos)
    return eng.run(E.method("DMC", nw, nblock=nblock, nstep=10, timestep=0.01))["eblock"][:, 0]


@pytest.mark.parametrize("prec", ["mixed_fp32", "mixed_bf16x6"])
def test_mixed_dmc_energy_tracks_fp64_with_common_random_numbers(prec):
    deck, mol = synth_deck("pp16")
    ref = _dmc_blocks(deck, mol, "fp64")
    got = _dmc_blocks(deck, mol, prec)
    d = np.abs(got - ref)
    _say("dmc fp64", ref, prec, got, "max |dE|", d.max())
    assert d.max() / np.abs(ref).max() < 1e-3


def test_bf16x3_dmc_first_block_tracks_fp64():
    """The 3-product BF16 split (relative product error ~2^-17) follows the FP64 trajectory
    only over short times: its first block agrees, later blocks drift (all-electron and
    pseudopotential molecules alike), which is why the FP32 and 6-product modes are the
    production mixed modes."""
    deck, mol = synth_deck("pp16")
    ref = _dmc_blocks(deck, mol, "fp64", nblock=1)
    got = _dmc_blocks(deck, mol, "mixed_bf16", nblock=1)
    _say("dmc first block fp64", ref, "bf16x3", got)
    assert abs(got[0] - ref[0]) / abs(ref[0]) < 1e-3
