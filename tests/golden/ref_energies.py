"""Writes tests/golden/ref_energies.json: VMC and DMC energies of the reference program itself.

Runs in the build container only (it needs oracle/_ref/qwalk, built from the reference sources
by oracle/Makefile.ref).  Each case runs P independent reference processes (distinct
RANDOMSEED, the walkers split between them, as MPI ranks would hold them); each process's
block energies come from its own .log; each block's energies are combined over the
processes by their total weights (the record an MPI run writes), the first `drop` blocks are
discarded as equilibration, and the mean and standard error are taken over the blocks.  Stored per case: the deck text, METHOD words, walkers, blocks and steps, and the
combined energy and error bar.

Cases:
  h4_vmc    the 4-electron H4 chain (the workload's plain FP64 VMC): VMC, 256 walkers,
            50 blocks x 20 steps, timestep 0.5, NDECORR 2
  pp16_vmc  the 64-electron pseudopotential molecule (stochastic nonlocal pseudopotential):
            VMC, 256 walkers, 150 blocks x 10 steps, timestep 0.3, the first half dropped (its
            energy still drifts over the first few hundred steps).  Its trial function (random
            orbitals) has no stationary DMC energy within a run of this length -- DMC keeps
            projecting down -- so its energies are compared in VMC, where both sample the same
            |Psi|^2; the stationary DMC comparison is the N2 deck
  n2_dmc    the reference's N2 deck (tests/n2 basis, orbitals and BFD pseudopotential) with the
            POLYPADE Jastrow of n2_pp: DMC at timestep 0.02, 512 walkers, 40 blocks x 20 steps

Usage: python tests/golden/ref_energies.py [case ...]   (from the repository root; ~15 min on 8 cores)
"""
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from mainline_amd import synth  # noqa: E402

QWALK = os.path.join(ROOT, "oracle", "_ref", "qwalk")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_energies.json")
PROCS = 8

CASES = {
    "h4_vmc": dict(deck="synth:h4", method="VMC", walkers=256, nblock=50, nstep=20, timestep=0.5, drop=5),
    "pp16_vmc": dict(deck="synth:pp16", method="VMC", walkers=256, nblock=150, nstep=10, timestep=0.3, drop=75),
    "n2_dmc": dict(deck="golden:n2_pp/n2.wfeval", method="DMC", walkers=512, nblock=40, nstep=20, timestep=0.02,
                   drop=10),
}


def deck_dir(spec, work):
    """Copies the case's deck files into work/ and returns the deck file name."""
    if spec.startswith("synth:"):
        mol = getattr(synth, spec[6:])()
        path = mol.write(work)
        return os.path.basename(path)
    src = os.path.join(ROOT, "tests", "golden", "decks", os.path.dirname(spec[7:]))
    for f in os.listdir(src):
        shutil.copy(os.path.join(src, f), work)
    return os.path.basename(spec[7:])


def run_case(name, c):
    work = tempfile.mkdtemp(prefix="refE_%s_" % name)
    deck = deck_dir(c["deck"], work)
    body = open(os.path.join(work, deck)).read()
    per = c["walkers"] // PROCS
    words = "%s NCONFIG %d NBLOCK %d NSTEP %d TIMESTEP %g" % (c["method"], per, c["nblock"], c["nstep"],
                                                             c["timestep"])
    procs = []
    for p in range(PROCS):
        f = os.path.join(work, "r%d.in" % p)
        with open(f, "w") as fh:
            fh.write("METHOD { %s }\nRANDOMSEED { %d %d }\n%s" % (words, 31 + 7 * p, 1001 + 13 * p, body))
        procs.append(subprocess.Popen([QWALK, os.path.basename(f)], cwd=work, stdout=subprocess.DEVNULL,
                                      stderr=subprocess.DEVNULL))
    codes = [q.wait() for q in procs]
    if any(codes):
        raise RuntimeError("%s: reference exited with %s (%s)" % (name, codes, work))
    # per block: the processes' energies combined by their total weights (what an MPI run's
    # block record holds), then the mean and standard error over blocks after `drop`
    E = []
    for p in range(PROCS):
        e, w = read_blocks(os.path.join(work, "r%d.in.log" % p))
        E.append((e, w))
    nb = min(len(e) for e, _ in E)
    blocks = np.array([sum(E[p][0][b] * E[p][1][b] for p in range(PROCS)) / sum(E[p][1][b] for p in range(PROCS))
                       for b in range(nb)])[c["drop"]:]
    mean = float(blocks.mean())
    err = float(blocks.std(ddof=1) / np.sqrt(len(blocks)))
    print("%s: %.6f +- %.6f over %d blocks" % (name, mean, err, len(blocks)), flush=True)
    return dict(c, method_words=words, processes=PROCS, energy=mean, error=err, blocks=blocks.tolist(),
                deck_text=body)


def read_blocks(path):
    """Block energies and total weights of a reference .log (src/properties/Properties_block.cpp:73-138)."""
    e, w = [], []
    for line in open(path):
        t = line.split()
        if len(t) >= 2 and t[0] == "totweight":
            w.append(float(t[1]))
        elif len(t) >= 3 and t[0] == "total_energy" and t[1] == "{":
            e.append(float(t[2]))
    return e, w


def main():
    """python ref_energies.py [case ...]: runs the named cases (all by default) and updates them
    in the json."""
    out = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in [k for k in out if k not in CASES]:
        del out[name]
    for name, c in CASES.items():
        if len(sys.argv) > 1 and name not in sys.argv[1:]:
            continue
        out[name] = run_case(name, c)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
