"""VMC and DMC energies of the engine against the reference program's own runs of the same
decks and methods (tests/golden/ref_energies.json, written by tests/golden/ref_energies.py
with the reference built from its sources): the FP64 engine's energy, with the same walkers,
blocks, steps and timestep (the first blocks dropped as equilibration, as there), agrees with
the reference's within two combined standard errors."""
import json
import os
import sys

import numpy as np
import pytest

from mainline_amd import engine as E
from tests._util import GOLDEN, synth_deck

pytestmark = pytest.mark.gpu

REF = json.load(open(os.path.join(GOLDEN, "ref_energies.json")))


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _deck(spec):
    if spec.startswith("synth:"):
        return synth_deck(spec[6:])[0]
    return os.path.join(GOLDEN, "decks", spec[7:])


@pytest.mark.parametrize("case", sorted(REF))
def test_energy_agrees_with_the_reference_program(case):
    c = REF[case]
    eng = E.Engine(_deck(c["deck"]), "fp64", seed=20261203)
    nw = c["walkers"]
    eng.set_walkers(eng.random_walkers(nw, seed=5))
    # the reference's generate_sample warms its walkers up with VMC before the method
    eng.run(E.method("VMC", nw, nblock=1, nstep=50, timestep=0.3 if c["method"] == "DMC" else c["timestep"]))
    m = E.method(c["method"], nw, nblock=c["nblock"], nstep=c["nstep"], timestep=c["timestep"])
    eb = eng.run(m)["eblock"]
    blocks = eb[c["drop"]:, 0]
    mean, err = blocks.mean(), blocks.std(ddof=1) / np.sqrt(len(blocks))
    comb = np.hypot(err, c["error"])
    _say(case, "engine %.5f +- %.5f" % (mean, err), "reference %.5f +- %.5f" % (c["energy"], c["error"]),
         "difference %.2f combined sigma" % (abs(mean - c["energy"]) / comb))
    assert abs(mean - c["energy"]) <= 2 * comb
