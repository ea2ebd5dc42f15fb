// Host-side setup of the walker engine: reads the QWalk deck (SYSTEM, PSEUDO, TRIALFUNC,
// METHOD) and flattens it into the tables the device kernels use.
#pragma once
#include <string>
#include <vector>

namespace qw {

constexpr int kMaxJF = 8;     // functions per Jastrow radial object
constexpr int kMaxPPL = 6;    // pseudopotential channels
constexpr int kMaxPPT = 6;    // Gaussian terms per channel
constexpr int kMaxAIP = 12;   // quadrature points
constexpr int kMaxDet = 256;  // determinants of a multi-determinant expansion

// Shell symmetries (src/basis_function/Cubic_spline_build.cpp:26-140,664-800).  7F_crystal and
// 9G_pyscf are the 7F and 9G functions in the orders of those programs.
enum SymKind { SK_S = 0, SK_P = 1, SK_5D = 2, SK_6D = 3, SK_10F = 4, SK_7F = 5, SK_7FC = 6, SK_9G = 7, SK_9GP = 8,
               SK_15G = 9 };

// One radial spline shell placed on an atom.
struct ShellDesc {
  int table;     // spline table index
  int sym;       // SymKind
  int f0;        // first basis function (global index)
  int nf;        // number of angular functions
  int atom;
  int pad;
  double rcut;   // cutoff of this shell's functions
  double objthr; // cutoff of the basis object the shell belongs to
};

// A Jastrow radial object (CUTOFF_CUSP: kind 0, nf 1; POLYPADE: kind 1) with its
// coefficients folded in: cl = like-spin (or one-body) coefficient, cu = unlike-spin.
struct JObj {
  int kind;
  int nf;
  double rcut, rcutinv, cusp, gamma, pade0;
  double beta[kMaxJF];
  double cl[kMaxJF];
  double cu[kMaxJF];
};

// Semilocal pseudopotential of one atom type (RGAUSSIAN OLDQMC).
struct PPType {
  int L;                     // channels; the last is the local one
  int nterm[kMaxPPL];
  double n[kMaxPPL][kMaxPPT];  // power of r (already decremented by 2)
  double a[kMaxPPL][kMaxPPT];
  double c[kMaxPPL][kMaxPPT];
  int aip;
  int addzeff;
  double q[kMaxAIP][3];
  double w[kMaxAIP];
  double cutoff;
};

// A linear Jastrow parameter, in the reference's order: group by group, the one-body
// COEFFICIENTS sections in input order (each coefficient of its atom type), then the two-body
// coefficients (TWOBODY in order; TWOBODY_SPIN like ones then unlike ones).  Pieces marked
// FREEZE have none (src/wavefunction/Jastrow2_wf.cpp:477-520, Jastrow2_one.cpp:25-100,
// Jastrow2_two.h:96,148).
struct JParm {
  int body;           // 1: electron-ion (one-body), 2: electron-electron
  int spin;           // two-body pairs it applies to: 0 all, 1 like spins, 2 unlike spins
  std::string label;  // one-body: the atom label
  double value;       // the coefficient (after SCALE)
  JObj unit;          // the basis object with only this coefficient, set to 1
};

struct SysTables {
  int nup = 0, ndn = 0, na = 0, nb = 0;
  std::vector<double> atom_pos;   // [na][3]
  std::vector<double> atom_Z;     // [na]
  std::vector<ShellDesc> shells;  // grouped by atom
  int ntab = 0, nint = 0;         // spline tables, intervals per table (padded to the max)
  std::vector<double> spline;     // [ntab][nint][4]
  std::vector<double> tab_invspacing, tab_spacing;
  std::vector<double> C[2];       // per spin [nb][nmo[s]]: the spin's MO list
  // determinants: weights, and each determinant's orbitals per spin as indices into that
  // spin's MO list ([ndet][nspin]); one determinant: nmo = nspin, occupation 0..n-1
  int ndet = 1;
  std::vector<double> detwt;
  std::vector<int> occ[2];
  int nmo[2] = {0, 0};
  // CSFs over consecutive determinants (DETWT input: one CSF {detwt, 1} per determinant):
  // coefficient, first determinant ([ncsf+1]) and each determinant's weight in its CSF;
  // OPTIMIZE_DET makes the coefficients of CSFs 1.. parameters
  std::vector<double> csf_c, csf_w;
  std::vector<int> csf_d0;
  bool optimize_det = false;
  std::vector<JObj> ee;           // electron-electron objects (merged over groups)
  std::vector<JObj> ei;           // electron-ion objects, atom ranges in atom_ei0
  std::vector<int> atom_ei0;      // [na+1]
  std::vector<JParm> jparms;      // linear Jastrow parameters (getParmDeriv order)
  std::vector<std::string> atom_label;  // [na]
  std::vector<PPType> pptypes;
  std::vector<int> atom_pp;       // [na], -1 = no pseudopotential
  double vii = 0;                 // ion-ion Coulomb energy
  int ne() const { return nup + ndn; }
};

struct MethodDesc {
  int method = 0;
  int nconfig = 1;
  int nblock = 100;
  int nstep = 100;
  int ndecorr = 2;
  double timestep = 0.5;
  int feedback_interval = 5;
  double feedback = 1.0;
  double max_weight = 7.0;
  double branch_start_cutoff = 10.0;
  int pp_stochastic = 1;
  int tmoves = 0;  // 1 TMOVES, 2 TMOVESSC
  double start_feedback = 1.0;  // DMC: feedback of the first trial energy
  std::string label = "vmc";    // block-log label (LABEL)
  // walker files: READCONFIG / STORECONFIG, both <deck>.config by default
  // (src/method/Vmc_method.cpp:59-62, src/method/Dmc_method.cpp:62-63,91-92)
  std::string readconfig, storeconfig;
};

// Throws std::runtime_error with a message on bad or unsupported input.
// `unsupported` is set when the input is valid QWalk but outside this engine's scope.
SysTables read_deck(const std::string& path, bool* unsupported);
MethodDesc read_method(const std::string& path);
std::vector<MethodDesc> read_methods(const std::string& path);

}  // namespace qw
