"""Seeded synthetic molecules written as QWalk input decks.

The generator emits the same text formats a QWalk user writes by hand, so that one deck
drives both this engine and the reference program:

  <name>.sys     SYSTEM { MOLECULE NSPIN { nu nd } ATOM { X Z COOR x y z } ... } [+ PSEUDO]
  <name>.basis   BASIS { X AOSPLINE GAMESS { ... } }      (src/basis_function/Cubic_spline_build.cpp:140)
  <name>.orb     "mo basis center label" lines, COEFFICIENTS, values (src/orbitals/MO_matrix.h:173)
  <name>.slater  SLATER ORBITALS { CUTOFF_MO ... } DETWT STATES  (src/wavefunction/Slat_wf_data.cpp:34)
  <name>.jast2   JASTROW2 with the group layout of tests/h2/jast2.wfin
  <name>.wfeval  SYSTEM + PSEUDO + TRIALFUNC, used for fixed-configuration evaluation
  <name>.vmc / <name>.dmc   METHOD section + the above, runnable by the reference program

Workloads (sizes as named in the benchmark definitions):
  h4      linear H4 chain at 1.4 bohr, 2+2 electrons, 20 basis functions, all-electron.
  pp16    16 carbon-like pseudo-atoms (Z_eff=4), 32+32 electrons, 208 functions, s/p
          nonlocal channels + local channel, 12-point quadrature.
  ae32    32 Z=4 all-electron atoms, 64+64 electrons, 416 functions (the 128-electron case).
  aeN     N Z=4 atoms, 2N+2N electrons, 13 functions per atom (uneven-size batches).

Choices the workload text leaves open are fixed here and recorded in the design notes:
contraction coefficients of the H4 S3 shell are seeded U(0.1, 1); MO coefficient columns are
Lowdin-orthonormalised in the Euclidean metric, C <- C (C^T C)^(-1/2); the box edge for N
atoms is 12 * (N/16)^(1/3) bohr; all atoms of a molecule share one species and one basis.
"""
import os
import numpy as np

# functions per AOSPLINE shell symmetry
SHELL_NF = {"S": 1, "P": 3, "5D": 5, "6D": 6, "10F": 10, "7F": 7, "7F_crystal": 7, "9G": 9, "9G_pyscf": 9,
            "15G": 15}

JASTROW_CUSP_GROUP = """GROUP {
  EEBASIS { EE CUTOFF_CUSP GAMMA 24.0 CUSP 1.0 CUTOFF 7.5 }
  EEBASIS { EE CUTOFF_CUSP GAMMA 24.0 CUSP 1.0 CUTOFF 7.5 }
  TWOBODY_SPIN { FREEZE LIKE_COEFFICIENTS { 0.25 0.0 } UNLIKE_COEFFICIENTS { 0.0 0.5 } }
}
"""


def _fmt(x):
    return "%.17g" % x


class Molecule:
    """One synthetic system: geometry, basis shells, MOs, Jastrow and optional PP."""

    def __init__(self, name, label, Z, coords, nup, ndn, shells, mo_up, mo_dn,
                 jast_ei, jast_ee, jast_nfunc, pseudo=None, seed=(1, 1)):
        self.name = name
        self.label = label
        self.Z = Z
        self.coords = np.asarray(coords, dtype=np.float64)
        self.nup = nup
        self.ndn = ndn
        self.shells = shells          # list of (sym, [(alpha, coeff), ...])
        self.mo_up = mo_up            # (nbasis, nup)
        self.mo_dn = mo_dn            # (nbasis, ndn)
        self.jast_ei = jast_ei        # one-body POLYPADE coefficients
        self.jast_ee = jast_ee        # two-body POLYPADE coefficients
        self.jast_nfunc = jast_nfunc
        self.pseudo = pseudo          # None or dict(aip, channels=[[(n, alpha, c), ...], ...])
        self.seed = seed

    @property
    def natoms(self):
        return self.coords.shape[0]

    @property
    def nelec(self):
        return self.nup + self.ndn

    @property
    def nfunc_atom(self):
        return sum(SHELL_NF[s] for s, _ in self.shells)

    @property
    def nbasis(self):
        return self.nfunc_atom * self.natoms

    # ------------------------------------------------------------------ text
    def sys_text(self):
        out = ["SYSTEM { MOLECULE", "  NSPIN { %d %d }" % (self.nup, self.ndn)]
        for a in range(self.natoms):
            x, y, z = self.coords[a]
            out.append("  ATOM { %s %s COOR %s %s %s }" % (self.label, _fmt(self.Z), _fmt(x), _fmt(y), _fmt(z)))
        out.append("}")
        if self.pseudo is not None:
            ch = self.pseudo["channels"]
            out.append("PSEUDO {")
            out.append("  %s" % self.label)
            out.append("  AIP %d" % self.pseudo["aip"])
            out.append("  BASIS { %s" % self.label)
            out.append("   RGAUSSIAN")
            out.append("   OLDQMC {")
            out.append("    0.0 %d" % len(ch))
            out.append("    " + " ".join(str(len(c)) for c in ch))
            for c in ch:
                for (n, alpha, coeff) in c:
                    out.append("    %d %s %s" % (n, _fmt(alpha), _fmt(coeff)))
            out.append("   }")
            out.append("  }")
            out.append("}")
        return "\n".join(out) + "\n"

    def basis_text(self):
        out = ["BASIS {", self.label, "AOSPLINE", "GAMESS {"]
        for sym, prims in self.shells:
            out.append("%s %d" % (sym, len(prims)))
            for i, (alpha, c) in enumerate(prims):
                out.append("  %d %s %s" % (i + 1, _fmt(alpha), _fmt(c)))
        out.append("}")
        out.append("}")
        return "\n".join(out) + "\n"

    def orb_text(self):
        """All MOs (up first, then down) over all centers; one coefficient label per entry."""
        nfa = self.nfunc_atom
        mos = np.concatenate([self.mo_up, self.mo_dn], axis=1)   # (nbasis, nmo)
        nmo = mos.shape[1]
        lines = []
        label = 1
        vals = []
        for mo in range(nmo):
            for a in range(self.natoms):
                for f in range(nfa):
                    lines.append("%d %d %d %d" % (mo + 1, f + 1, a + 1, label))
                    vals.append(mos[a * nfa + f, mo])
                    label += 1
        lines.append("COEFFICIENTS")
        for i in range(0, len(vals), 4):
            lines.append(" ".join(_fmt(v) for v in vals[i:i + 4]))
        return "\n".join(lines) + "\n"

    def slater_text(self):
        nmo = self.nup + self.ndn
        up = " ".join(str(i + 1) for i in range(self.nup))
        dn = " ".join(str(self.nup + i + 1) for i in range(self.ndn))
        return ("SLATER\nORBITALS {\n  CUTOFF_MO\n  MAGNIFY 1\n  NMO %d\n  ORBFILE %s.orb\n"
                "  INCLUDE %s.basis\n  CENTERS { USEATOMS }\n}\nDETWT { 1.0 }\nSTATES {\n  %s\n  %s\n}\n"
                % (nmo, self.name, self.name, up, dn))

    def jast_text(self):
        ei = " ".join(_fmt(c) for c in self.jast_ei)
        ee = " ".join(_fmt(c) for c in self.jast_ee)
        return ("JASTROW2\n" + JASTROW_CUSP_GROUP +
                "GROUP {\n"
                "  EIBASIS { %s POLYPADE BETA0 0.2 NFUNC %d RCUT 7.5 }\n"
                "  ONEBODY { COEFFICIENTS { %s %s } }\n"
                "  EEBASIS { EE POLYPADE BETA0 0.5 NFUNC %d RCUT 7.5 }\n"
                "  TWOBODY { COEFFICIENTS { %s } }\n"
                "}\n" % (self.label, self.jast_nfunc, self.label, ei, self.jast_nfunc, ee))

    def trialfunc_text(self):
        return ("TRIALFUNC {\n  SLATER-JASTROW\n  WF1 { INCLUDE %s.slater }\n"
                "  WF2 { INCLUDE %s.jast2 }\n}\n" % (self.name, self.name))

    def method_text(self, method, **kw):
        opts = " ".join("%s %s" % (k.upper(), v) for k, v in kw.items())
        return ("METHOD { %s %s }\nINCLUDE %s.sys\n%sRANDOMSEED { %d %d }\n"
                % (method.upper(), opts, self.name, self.trialfunc_text(), self.seed[0], self.seed[1]))

    def write(self, directory, methods=None):
        """Write all deck files into `directory`; returns the path of the .wfeval deck."""
        os.makedirs(directory, exist_ok=True)
        files = {
            "%s.sys" % self.name: self.sys_text(),
            "%s.basis" % self.name: self.basis_text(),
            "%s.orb" % self.name: self.orb_text(),
            "%s.slater" % self.name: self.slater_text(),
            "%s.jast2" % self.name: self.jast_text(),
            "%s.wfeval" % self.name: "INCLUDE %s.sys\n%s" % (self.name, self.trialfunc_text()),
        }
        for meth, kw in (methods or {}).items():
            files["%s.%s" % (self.name, meth.lower())] = self.method_text(meth, **kw)
        for fn, txt in files.items():
            with open(os.path.join(directory, fn), "w") as f:
                f.write(txt)
        return os.path.join(directory, "%s.wfeval" % self.name)


# ---------------------------------------------------------------------- builders
def _lowdin(c):
    s = c.T @ c
    w, v = np.linalg.eigh(s)
    return c @ (v @ np.diag(w ** -0.5) @ v.T)


def _place_atoms(rng, natoms, box, minsep=2.5):
    pts = []
    tries = 0
    while len(pts) < natoms:
        p = rng.uniform(-box / 2, box / 2, size=3)
        tries += 1
        if tries > 200000:
            raise RuntimeError("could not place %d atoms in a %.2f bohr box" % (natoms, box))
        if all(np.linalg.norm(p - q) >= minsep for q in pts):
            pts.append(p)
    return np.array(pts)


def h4(seed=20261202):
    """Linear H4, 1.4 bohr spacing, 2+2 electrons, S3/S1/P1 per atom (20 functions)."""
    rng = np.random.default_rng(seed)
    z = (np.arange(4) - 1.5) * 1.4
    coords = np.stack([np.zeros(4), np.zeros(4), z], axis=1)
    c3 = rng.uniform(0.1, 1.0, size=3)
    shells = [("S", [(13.0, c3[0]), (1.96, c3[1]), (0.44, c3[2])]),
              ("S", [(0.12, 1.0)]),
              ("P", [(0.8, 1.0)])]
    nb = 20
    mo_up = _lowdin(rng.standard_normal((nb, 2)))
    mo_dn = _lowdin(rng.standard_normal((nb, 2)))
    ei = rng.uniform(-0.1, 0.1, size=3)
    ee = rng.uniform(-0.1, 0.1, size=3)
    return Molecule("h4", "H", 1.0, coords, 2, 2, shells, mo_up, mo_dn, ei, ee, 3,
                    seed=(seed % 100000007, 1))


def _z4_shells(rng):
    a = np.sort(np.exp(rng.uniform(np.log(0.1), np.log(10.0), size=5)))[::-1]
    return [("S", [(a[0], 1.0)]), ("S", [(a[3], 1.0)]),
            ("P", [(a[1], 1.0)]), ("P", [(a[4], 1.0)]),
            ("5D", [(a[2], 1.0)])]


def z4_molecule(natoms, seed, name=None, pseudo=False):
    """natoms Z=4 atoms (all-electron, or carbon-like pseudo-atoms), 2*natoms electrons per spin."""
    rng = np.random.default_rng(seed)
    box = 12.0 * (natoms / 16.0) ** (1.0 / 3.0)
    coords = _place_atoms(rng, natoms, box)
    shells = _z4_shells(rng)
    nb = 13 * natoms
    nup = ndn = 2 * natoms
    mo_up = _lowdin(rng.standard_normal((nb, nup)))
    mo_dn = _lowdin(rng.standard_normal((nb, ndn)))
    ei = rng.uniform(-0.1, 0.1, size=4)
    ee = rng.uniform(-0.1, 0.1, size=4)
    pp = None
    label = "Be"
    if pseudo:
        label = "C"
        # nonlocal s, nonlocal p, then the local channel (last), RGAUSSIAN OLDQMC
        # terms (n, alpha, c) meaning c * r^(n-2) * exp(-alpha r^2)
        pp = {"aip": 12, "channels": [
            [(2, rng.uniform(3.0, 8.0), rng.uniform(10.0, 40.0)),
             (2, rng.uniform(3.0, 8.0), -rng.uniform(5.0, 15.0))],
            [(2, rng.uniform(3.0, 8.0), rng.uniform(5.0, 20.0))],
            [(1, rng.uniform(5.0, 10.0), 4.0),
             (3, rng.uniform(4.0, 8.0), rng.uniform(10.0, 30.0)),
             (2, rng.uniform(4.0, 8.0), -rng.uniform(5.0, 15.0))],
        ]}
    return Molecule(name or ("pp%d" % natoms if pseudo else "ae%d" % natoms), label, 4.0,
                    coords, nup, ndn, shells, mo_up, mo_dn, ei, ee, 4, pseudo=pp,
                    seed=(seed % 100000007, 7))


def fg_molecule(seed=20261206):
    """Three Z=4 atoms whose basis holds every supported shell symmetry (S, P, 6D, 7F,
    7F_crystal, 10F, 9G, 9G_pyscf, 15G: 67 functions per atom), 6 + 6 electrons: the fixture
    system for the F and G angular functions."""
    rng = np.random.default_rng(seed)
    coords = np.array([[0.0, 0.0, 0.0], [2.4, 0.3, -0.2], [-0.7, 2.1, 0.9]])
    syms = ["S", "P", "6D", "7F", "7F_crystal", "10F", "9G", "9G_pyscf", "15G"]
    shells = [(sy, [(float(a), 1.0)]) for sy, a in zip(syms, np.exp(rng.uniform(np.log(0.3), np.log(1.5), len(syms))))]
    nb = 3 * sum(SHELL_NF[sy] for sy in syms)
    mo_up = _lowdin(rng.standard_normal((nb, 6)))
    mo_dn = _lowdin(rng.standard_normal((nb, 6)))
    ei = rng.uniform(-0.1, 0.1, size=4)
    ee = rng.uniform(-0.1, 0.1, size=4)
    return Molecule("fg3", "Be", 4.0, coords, 6, 6, shells, mo_up, mo_dn, ei, ee, 4, seed=(seed % 100000007, 7))


def pp16(seed=20261203):
    return z4_molecule(16, seed, name="pp16", pseudo=True)


def ae32(seed=20261204):
    return z4_molecule(32, seed, name="ae32", pseudo=False)


def batch_molecules(seed=20261205):
    """Eight molecules with 8..192 electrons and seeded walker counts in [512, 4096]."""
    rng = np.random.default_rng(seed)
    sizes = [8, 16, 32, 48, 64, 96, 128, 192]
    mols = []
    for i, ne in enumerate(sizes):
        m = z4_molecule(ne // 4, seed + 17 * (i + 1), name="ae%d_b" % (ne // 4))
        mols.append((m, int(rng.integers(512, 4097))))
    return mols


def random_configs(mol, k, seed=7, spread=1.2):
    """k electron configurations near the atoms (Gaussian clouds), shape (k, ne, 3)."""
    rng = np.random.default_rng(seed)
    ne = mol.nelec
    out = np.empty((k, ne, 3))
    for c in range(k):
        idx = np.concatenate([np.arange(mol.nup) % mol.natoms, np.arange(mol.ndn) % mol.natoms])
        out[c] = mol.coords[idx] + spread * rng.standard_normal((ne, 3))
    return out


def nonsingular_configs(mol, k, logpsi, seed=7, spread=1.2, tries=20):
    """random_configs with every configuration whose Slater determinant is exactly zero (an
    electron beyond every basis function's cutoff: ln|Psi| = -1e201) drawn again from the next
    seed; logpsi(pos) -> ln|Psi| per configuration."""
    out = random_configs(mol, k, seed=seed, spread=spread)
    for t in range(1, tries + 1):
        bad = np.where(~(logpsi(out) > -1e200))[0]
        if bad.size == 0:
            return out
        out[bad] = random_configs(mol, bad.size, seed=seed + 1000 * t, spread=spread)
    raise RuntimeError("no nonsingular configurations for %s" % mol.name)


def write_configs(path, cfg):
    cfg = np.asarray(cfg)
    with open(path, "w") as f:
        f.write("%d %d\n" % (cfg.shape[0], cfg.shape[1]))
        for c in cfg:
            for e in c:
                f.write("%s %s %s\n" % (_fmt(e[0]), _fmt(e[1]), _fmt(e[2])))
