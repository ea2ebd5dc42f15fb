"""Writes tests/golden/configs/: walker files (.config) written by the reference program itself.

Runs in the build container only (it needs oracle/_ref/qwalk, built from the reference sources
by oracle/Makefile.ref).  Each case runs the reference on a deck with a short METHOD section and
STORECONFIG, and keeps the file it wrote:

  h4_vmc.config   the 4-electron H4 chain after VMC (NCONFIG 12, 2 blocks of 5 steps):
                  Config_save_point records only
  n2_dmc.config   the reference's N2 pseudopotential deck after DMC (NCONFIG 8, 2 blocks of
                  5 steps, timestep 0.02): Dmc_point records with weights and signs

Usage: python tests/golden/ref_configs.py   (from the repository root; a few seconds)
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from mainline_amd import synth  # noqa: E402

QWALK = os.path.join(ROOT, "oracle", "_ref", "qwalk")
OUT = os.path.join(HERE, "configs")

CASES = {
    "h4_vmc": dict(deck="synth:h4", words="VMC NCONFIG 12 NBLOCK 2 NSTEP 5 TIMESTEP 0.5"),
    "n2_dmc": dict(deck="golden:n2_pp/n2.wfeval", words="DMC NCONFIG 8 NBLOCK 2 NSTEP 5 TIMESTEP 0.02"),
}


def deck_body(spec, work):
    if spec.startswith("synth:"):
        path = getattr(synth, spec[6:])().write(work)
        return open(path).read()
    src = os.path.join(HERE, "decks", os.path.dirname(spec[7:]))
    for f in os.listdir(src):
        shutil.copy(os.path.join(src, f), work)
    return open(os.path.join(work, os.path.basename(spec[7:]))).read()


def main():
    os.makedirs(OUT, exist_ok=True)
    for name, c in CASES.items():
        work = tempfile.mkdtemp(prefix="refcfg_%s_" % name)
        body = deck_body(c["deck"], work)
        with open(os.path.join(work, "run.in"), "w") as fh:
            fh.write("METHOD { %s STORECONFIG %s.config }\nRANDOMSEED { 17 4242 }\n%s" % (c["words"], name, body))
        subprocess.run([QWALK, "run.in"], cwd=work, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        shutil.copy(os.path.join(work, name + ".config"), os.path.join(OUT, name + ".config"))
        print(name, "->", os.path.join(OUT, name + ".config"))
        shutil.rmtree(work)


if __name__ == "__main__":
    main()
