"""Writes the golden vectors under tests/golden/ from the compiled reference.

Runs in the build container only: it needs oracle/_ref/qw_golden, which oracle/Makefile.ref
builds from the reference sources.  For each case it evaluates fixed electron
configurations through the reference's own Wavefunction / System / Pseudopotential objects
(deterministic pseudopotential, identity quadrature rotation) and stores inputs and outputs
as an .npz fixture:

  pos[K][Ne][3], logpsi[K], sign[K], grad[K][Ne][3], lap[K][Ne], kinetic[K], vloc[K],
  vnonloc[K], and from Wavefunction::getParmDeriv: parms[np] (getVarParms),
  pd_gradient[K][np], pd_gradderiv[K][np][Ne][4], pd_valgrad[K][Ne][3],
  and from Pseudopotential::calcNonlocTmove: vnl_tmove[K] and the candidate list
  tmove[M][6] = configuration, electron, vxx, displacement (3), in the reference's order
  (+ orb_sha256 for generated decks, to detect generator drift)

Cases:
  h_atom    the reference's tests/h deck (H atom, AOSPLINE S shells, CUTOFF_MO)
  h2_jast   the reference's tests/h2 deck with its two-group Jastrow (cusp + POLYPADE)
  n2_pp     the reference's tests/n2 basis, orbitals and BFD pseudopotential (AIP 6), one
            determinant, with a POLYPADE Jastrow of nonzero coefficients
  n2_cen    n2_pp with its basis on two centers read from a file (CENTERS { READ }),
            displaced from the nuclei
  n2_blas   n2_pp with BLAS_MO orbitals (the reference drops basis values below 1e-8)
  n2_ci     the reference's tests/n2 CI wave function (28 determinants, qw.ci + cidet);
  n2_cij    the same with the POLYPADE Jastrow of n2_pp
  n2_csf    the n2_ci expansion written as CSF sections
  n2_cio    n2_ci with OPTIMIZE_DET (the determinant weights 2.. as parameters)
  n2_csfj   n2_csf with OPTIMIZE_DET and the Jastrow (CSF coefficients, then Jastrow parameters)
  h4, pp16, ae32   the seeded synthetic molecules of mainline_amd/synth.py (4, 64, 128 e)
  ae32_vmc  16 configurations of the 128-electron molecule from a seeded FP64 VMC
            equilibration (the oracle's VMC: 30 steps at timestep 0.3 from Gaussian clouds),
            the kind of configuration the mixed-precision error report uses
  fg3       three atoms with every supported shell symmetry (S P 6D 7F 7F_crystal 10F 9G
            9G_pyscf 15G), 12 electrons

Output files: the reference program (oracle/_ref/qwalk) runs a short DMC of the tests/h deck
and its block log and per-block JSON are kept as ref_output/h_dmc.qwlog and h_dmc.json (the
layout of the files the qwalk_amd driver writes is checked against them).

Usage: python tests/golden/make_golden.py [case ...]   (from the repository root)
"""
import hashlib
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from mainline_amd import synth  # noqa: E402
from oracle.oracle import REF_GOLDEN, REF_QWALK, parse_golden  # noqa: E402

REF_TESTS = "/root/reference/tests"
OUT = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(OUT, "decks")

N2_SLATER = """SLATER
ORBITALS {
  CUTOFF_MO
  MAGNIFY 1
  NMO 8
  ORBFILE qw.orb
  INCLUDE qw.basis
  CENTERS { USEGLOBAL }
}
DETWT { 1.0 }
STATES {
  1 2 3 4 5
  1 2 3 4 5
}
"""
N2_JAST = """JASTROW2
GROUP {
  EEBASIS { EE CUTOFF_CUSP GAMMA 24.0 CUSP 1.0 CUTOFF 7.5 }
  EEBASIS { EE CUTOFF_CUSP GAMMA 24.0 CUSP 1.0 CUTOFF 7.5 }
  TWOBODY_SPIN { FREEZE LIKE_COEFFICIENTS { 0.25 0.0 } UNLIKE_COEFFICIENTS { 0.0 0.5 } }
}
GROUP {
  EEBASIS { EE POLYPADE BETA0 0.5 NFUNC 3 RCUT 7.5 }
  EIBASIS { N POLYPADE BETA0 0.2 NFUNC 4 RCUT 7.5 }
  ONEBODY { COEFFICIENTS { N -0.3 0.1 -0.05 0.02 } }
  TWOBODY { COEFFICIENTS { 0.12 -0.04 0.01 } }
}
"""


def run_ref(deck_dir, deck, pos):
    cfg = os.path.join(deck_dir, "golden_cfg.txt")
    synth.write_configs(cfg, pos)
    out = subprocess.run([REF_GOLDEN, deck, "golden_cfg.txt"], cwd=deck_dir, capture_output=True,
                         text=True, check=True).stdout
    os.remove(cfg)
    g = parse_golden(out)
    K = len(g)
    return dict(
        pos=np.asarray(pos),
        logpsi=np.array([c["logpsi"] for c in g]),
        sign=np.array([c["sign"] for c in g], dtype=np.float64),
        grad=np.array([c["elec"][:, :3] for c in g]),
        lap=np.array([c["elec"][:, 3] for c in g]),
        kinetic=np.array([c["kinetic"] for c in g]),
        vloc=np.array([c["vloc"] for c in g]),
        vnonloc=np.array([c["vnonloc"] for c in g]),
        parms=g[0]["parm"],
        pd_gradient=np.array([c["pdgrad"] for c in g]),
        pd_gradderiv=np.array([c["pdgd"] for c in g]),
        pd_valgrad=np.array([c["pdvg"] for c in g]),
        vnl_tmove=np.array([c["vnl_tmove"] for c in g]),
        tmove=np.array([[k] + list(t) for k, c in enumerate(g) for t in c["tmove"]]).reshape(-1, 6),
    ) if K else None


def clouds(centers, counts, k, seed, spread):
    rng = np.random.default_rng(seed)
    idx = np.concatenate([np.arange(n) % len(centers) for n in counts])
    centers = np.asarray(centers, dtype=np.float64)
    return centers[idx][None] + spread * rng.standard_normal((k, len(idx), 3))


def csf_text(cidet):
    """CSF sections for a DETWT + STATES text: consecutive determinants of equal |weight|
    grouped into CSF { |w| sign ... }, STATES kept."""
    import re
    body = re.sub(r"#[^\n]*", "", cidet)
    w = [float(x) for x in re.search(r"DETWT\s*\{([^}]*)\}", body).group(1).split()]
    states = re.search(r"STATES\s*\{[^}]*\}", body).group(0)
    out, i = [], 0
    while i < len(w):
        j = i
        while j + 1 < len(w) and abs(abs(w[j + 1]) - abs(w[i])) < 1e-15:
            j += 1
        out.append("CSF { %r %s }" % (abs(w[i]), " ".join("1" if w[k] > 0 else "-1" for k in range(i, j + 1))))
        i = j + 1
    return "\n".join(out) + "\n" + states + "\n"


def ref_deck_cases(tmp):
    cases = {}
    # H atom (tests/h): qw.sys + qw.slater (+ qw.basis, qw.orb)
    d = os.path.join(DECKS, "h_atom")
    os.makedirs(d, exist_ok=True)
    for f in ["qw.sys", "qw.slater", "qw.basis", "qw.orb"]:
        shutil.copy(os.path.join(REF_TESTS, "h", f), d)
    with open(os.path.join(d, "h.wfeval"), "w") as f:
        f.write("INCLUDE qw.sys\nTRIALFUNC { INCLUDE qw.slater }\n")
    cases["h_atom"] = (d, "h.wfeval", clouds([[0, 0, 0]], [1, 0], 16, 11, 1.3))
    # H2 with Jastrow (tests/h2)
    d = os.path.join(DECKS, "h2_jast")
    os.makedirs(d, exist_ok=True)
    for f in ["qw.sys", "qw.slater", "qw.basis", "qw.orb", "jast2.wfin"]:
        shutil.copy(os.path.join(REF_TESTS, "h2", f), d)
    with open(os.path.join(d, "h2.wfeval"), "w") as f:
        f.write("INCLUDE qw.sys\nTRIALFUNC { SLATER-JASTROW WF1 { INCLUDE qw.slater } "
                "WF2 { INCLUDE jast2.wfin } }\n")
    cases["h2_jast"] = (d, "h2.wfeval", clouds([[0, 0, -1.3228], [0, 0, 1.3228]], [1, 1], 16, 12, 1.0))
    # N2 with the BFD pseudopotential (tests/n2), one determinant
    d = os.path.join(DECKS, "n2_pp")
    os.makedirs(d, exist_ok=True)
    for f in ["qw.sys", "qw.basis", "qw.orb"]:
        shutil.copy(os.path.join(REF_TESTS, "n2", f), d)
    with open(os.path.join(d, "n2.slater"), "w") as f:
        f.write(N2_SLATER)
    with open(os.path.join(d, "n2.jast2"), "w") as f:
        f.write(N2_JAST)
    with open(os.path.join(d, "n2.wfeval"), "w") as f:
        f.write("INCLUDE qw.sys\nTRIALFUNC { SLATER-JASTROW WF1 { INCLUDE n2.slater } "
                "WF2 { INCLUDE n2.jast2 } }\n")
    cases["n2_pp"] = (d, "n2.wfeval", clouds([[0, 0, -1.60626709], [0, 0, 1.60626709]], [5, 5], 12, 13, 0.9))
    # the same wave function with its basis on centers read from a file, displaced from the
    # nuclei (CENTERS { READ }), and with BLAS_MO orbitals
    with open(os.path.join(d, "n2.centers"), "w") as f:
        f.write("2\nN 0.05 -0.03 -1.50626709\nN -0.04 0.02 1.70626709\n")
    with open(os.path.join(d, "n2cen.slater"), "w") as f:
        f.write(N2_SLATER.replace("CENTERS { USEGLOBAL }", "CENTERS { READ n2.centers }"))
    with open(os.path.join(d, "n2cen.wfeval"), "w") as f:
        f.write("INCLUDE qw.sys\nTRIALFUNC { SLATER-JASTROW WF1 { INCLUDE n2cen.slater } "
                "WF2 { INCLUDE n2.jast2 } }\n")
    cases["n2_cen"] = (d, "n2cen.wfeval", clouds([[0, 0, -1.60626709], [0, 0, 1.60626709]], [5, 5], 12, 19, 0.9))
    with open(os.path.join(d, "n2blas.slater"), "w") as f:
        f.write(N2_SLATER.replace("CUTOFF_MO", "BLAS_MO"))
    with open(os.path.join(d, "n2blas.wfeval"), "w") as f:
        f.write("INCLUDE qw.sys\nTRIALFUNC { SLATER-JASTROW WF1 { INCLUDE n2blas.slater } "
                "WF2 { INCLUDE n2.jast2 } }\n")
    cases["n2_blas"] = (d, "n2blas.wfeval", clouds([[0, 0, -1.60626709], [0, 0, 1.60626709]], [5, 5], 12, 23, 0.9))
    # N2 CI (tests/n2/qw.ci): 28 determinants over 8 MOs, BFD pseudopotential; without and
    # with the Jastrow above
    d = os.path.join(DECKS, "n2_ci")
    os.makedirs(d, exist_ok=True)
    for f in ["qw.sys", "qw.basis", "qw.orb", "cidet"]:
        shutil.copy(os.path.join(REF_TESTS, "n2", f), d)
    with open(os.path.join(d, "n2_jast.jast2"), "w") as f:
        f.write(N2_JAST)
    slater = ("SLATER\n ORBITALS {\n  CUTOFF_MO\n  MAGNIFY 1\n  NMO 8\n  ORBFILE qw.orb\n"
              "  INCLUDE qw.basis\n  CENTERS { USEGLOBAL }\n }\n include cidet\n")
    with open(os.path.join(d, "n2ci.wfeval"), "w") as f:
        f.write("INCLUDE qw.sys\nTRIALFUNC {\n" + slater + "}\n")
    with open(os.path.join(d, "n2cij.wfeval"), "w") as f:
        f.write("INCLUDE qw.sys\nTRIALFUNC { SLATER-JASTROW\nWF1 {\n" + slater + "}\nWF2 { INCLUDE n2_jast.jast2 }\n}\n")
    n2c = clouds([[0, 0, -1.60626709], [0, 0, 1.60626709]], [5, 5], 12, 17, 0.9)
    cases["n2_ci"] = (d, "n2ci.wfeval", n2c)
    cases["n2_cij"] = (d, "n2cij.wfeval", n2c)
    # the same expansion as CSFs: runs of equal |weight| become one CSF { |w| +-1 ... }
    with open(os.path.join(d, "cidet_csf"), "w") as f:
        f.write(csf_text(open(os.path.join(d, "cidet")).read()))
    with open(os.path.join(d, "n2csf.wfeval"), "w") as f:
        f.write("INCLUDE qw.sys\nTRIALFUNC {\n" + slater.replace("include cidet", "include cidet_csf") + "}\n")
    cases["n2_csf"] = (d, "n2csf.wfeval", n2c)
    # OPTIMIZE_DET: the CSF coefficients as parameters, with DETWT alone and with CSFs and a Jastrow
    with open(os.path.join(d, "n2cio.wfeval"), "w") as f:
        f.write("INCLUDE qw.sys\nTRIALFUNC {\n" + slater.replace("SLATER\n", "SLATER\n OPTIMIZE_DET\n") + "}\n")
    with open(os.path.join(d, "n2csfj.wfeval"), "w") as f:
        f.write("INCLUDE qw.sys\nTRIALFUNC { SLATER-JASTROW\nWF1 {\n"
                + slater.replace("SLATER\n", "SLATER\n OPTIMIZE_DET\n").replace("include cidet", "include cidet_csf")
                + "}\nWF2 { INCLUDE n2_jast.jast2 }\n}\n")
    cases["n2_cio"] = (d, "n2cio.wfeval", n2c)
    cases["n2_csfj"] = (d, "n2csfj.wfeval", n2c)
    return cases


OUTPUT_DECK = "method { DMC timestep 0.02 nblock 2 nstep 5 nconfig 16 }\ninclude qw.sys\ntrialfunc { include qw.slater }\n"


def ref_output_files(tmp):
    """The reference's own .log and .json for a short DMC of the H atom."""
    d = os.path.join(tmp, "h_out")
    os.makedirs(d)
    for f in ["qw.sys", "qw.slater", "qw.basis", "qw.orb"]:
        shutil.copy(os.path.join(REF_TESTS, "h", f), d)
    open(os.path.join(d, "qw.run"), "w").write(OUTPUT_DECK)
    subprocess.run([REF_QWALK, "qw.run"], cwd=d, check=True, capture_output=True, timeout=300)
    out = os.path.join(OUT, "ref_output")
    os.makedirs(out, exist_ok=True)
    open(os.path.join(out, "h_dmc.deck"), "w").write(OUTPUT_DECK)
    # the block log is kept as .qwlog (".log" files are treated as run logs elsewhere)
    for ext, keep in [("log", "qwlog"), ("json", "json")]:
        shutil.copy(os.path.join(d, "qw.run." + ext), os.path.join(out, "h_dmc." + keep))


def main():
    """python make_golden.py [case ...]: all fixtures by default, else only the named cases."""
    if not os.path.exists(REF_GOLDEN):
        sys.exit("build the reference first: make -f oracle/Makefile.ref -j8")
    only = set(sys.argv[1:])
    with tempfile.TemporaryDirectory() as tmp:
        if not only:
            ref_output_files(tmp)
        cases = ref_deck_cases(tmp)
        for name, (d, deck, pos) in cases.items():
            if only and name not in only:
                continue
            res = run_ref(d, deck, pos)
            np.savez_compressed(os.path.join(OUT, name + ".npz"), deck=os.path.relpath(os.path.join(d, deck), OUT), **res)
            print(name, res["pos"].shape)
        for mol, k in [(synth.h4(), 16), (synth.pp16(), 8), (synth.ae32(), 4), (synth.fg_molecule(), 12)]:
            if only and mol.name not in only:
                continue
            d = os.path.join(tmp, mol.name)
            deck = os.path.basename(mol.write(d))
            orb = open(os.path.join(d, mol.name + ".orb"), "rb").read()
            pos = synth.random_configs(mol, k, seed=101)
            res = run_ref(d, deck, pos)
            np.savez_compressed(os.path.join(OUT, mol.name + ".npz"), deck="synth:" + mol.name,
                                orb_sha256=hashlib.sha256(orb).hexdigest(), **res)
            print(mol.name, res["pos"].shape)
        if not only or "ae32_vmc" in only:
            from oracle.oracle import Oracle
            mol = synth.ae32()
            d = os.path.join(tmp, "ae32_vmc")
            deck = os.path.basename(mol.write(d))
            orb = open(os.path.join(d, mol.name + ".orb"), "rb").read()
            start = synth.random_configs(mol, 16, seed=202)
            pos = Oracle(os.path.join(d, deck)).run(start, "VMC", nblock=1, nstep=30, timestep=0.3, seed=20261230,
                                                   nthreads=8)["pos"]
            res = run_ref(d, deck, pos)
            np.savez_compressed(os.path.join(OUT, "ae32_vmc.npz"), deck="synth:" + mol.name,
                                orb_sha256=hashlib.sha256(orb).hexdigest(), **res)
            print("ae32_vmc", res["pos"].shape)


if __name__ == "__main__":
    main()
