/* Test infrastructure only: the oracle.
 *
 * An FP64 CPU restatement of the QWalk VMC/DMC hot path, written from the reference's
 * algorithms (each function in qw_oracle.cpp cites the reference file:line it follows).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
 * as the checker or the timed CPU baseline; the product never links or calls it.
 */
#ifndef QW_ORACLE_H
#define QW_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_system orc_system;

/* Read SYSTEM / PSEUDO / TRIALFUNC from a QWalk deck; INCLUDE and ORBFILE paths are resolved
 * relative to the deck's directory.  Returns NULL and fills err on failure. */
orc_system* orc_create(const char* deck_path, char* err, int errlen);
void orc_destroy(orc_system* s);

/* nup, ndn, natoms, nbasis (total basis functions), has_pseudo */
void orc_info(const orc_system* s, int* out5);

/* Fixed-configuration evaluation (the parity quantity).
 *   pos      [K][Ne][3]
 *   rot      [9] quadrature rotation (row-major x,y,z axes as rows) or NULL for identity
 *   pp_mode  0 = deterministic (cutoff), 1 = stochastic with the uniforms in pp_rand
 *   pp_rand  [K][natoms][Ne] uniforms, used when pp_mode==1
 * outputs:
 *   logpsi[K], sign[K], grad[K][Ne][3], lap[K][Ne] (lap = laplacian(Psi)/Psi),
 *   eparts[K][3] = kinetic, local potential (Coulomb), nonlocal pseudopotential. */
int orc_eval(orc_system* s, int K, const double* pos, const double* rot, int pp_mode,
             const double* pp_rand, double* logpsi, double* sign, double* grad, double* lap,
             double* eparts);

/* Linear Jastrow parameters (Wavefunction::getParmDeriv order): returns their number and,
 * when values != NULL, their values. */
int orc_nparm(orc_system* s, double* values);

/* Sets the parameters (orc_nparm order); returns the number set.  For finite differences. */
int orc_set_parms(orc_system* s, const double* values);

/* d ln Psi / dp at K configurations: grad[K][np], gd[K][np][Ne][4] (d/dp of grad ln Psi and
 * of lap Psi / Psi), vg[K][Ne][3] (the val_gradient of the reference).  Returns np. */
int orc_parm_deriv(orc_system* s, int K, const double* pos, double* grad, double* gd, double* vg);

/* MO values of one spin's orbital list at n points: out[n][5][nmo_spin]
 * (value, d/dx, d/dy, d/dz, laplacian). */
int orc_mo_eval(orc_system* s, int spin, int n, const double* pts, double* out);

/* Basis values of all functions at n points: out[n][nbasis][5]. */
int orc_basis_eval(orc_system* s, int n, const double* pts, double* out);

/* Spline of basis shell `shell` (counted over all AOSPLINE objects in order):
 * returns number of intervals; fills coeff[nint][4] when coeff != NULL, plus
 * meta[3] = spacing, threshold(after padding), cutoff of the shell. */
int orc_spline(orc_system* s, int shell, double* coeff, double* meta);

/* Dense per-spin coefficient matrix C[nbasis][nmo_spin] (dropped entries are 0) and the
 * per-function cutoff radii rcut[nbasis]. */
int orc_mo_coeff(orc_system* s, int spin, double* C, double* rcut);

/* Standalone pieces used by pinning tests. */
/* LU with partial pivoting and implicit scaling; transposed inverse a1 = (a^-1)^T.
 * Returns 1 on success (logdet/sign filled), 0 on a singular matrix. */
int orc_transpose_inverse(int n, const double* a, double* a1, double* logdet, double* sign);
/* Sherman-Morrison update of the transposed inverse for new row `row` of the matrix;
 * returns det_old/det_new as the reference does. */
double orc_inverse_update_row(int n, double* a1, const double* newrow, int row);
/* Philox4x32-10 block: ctr[4], key[2] -> out[4]. */
void orc_philox(const unsigned* ctr, const unsigned* key, unsigned* out);

/* CPU walker engine (the same moves, weights and branching as the device engine).
 * method: 0 = VMC, 1 = DMC.  Walker positions are in/out: pos[nw][Ne][3].
 * Block results: eblock[nblock][5] = weighted E, KE, Vloc, Vnl, weight sum.
 * Returns acceptance ratio. */
typedef struct orc_run_opts {
  int method;
  int nwalkers;
  int nblock;
  int nstep;
  int ndecorr;         /* VMC sweeps per step */
  double timestep;
  unsigned long long seed;
  int pp_stochastic;   /* 1 = reference default stochastic PP evaluation */
  int feedback_interval;
  double feedback;
  double max_weight;
  double branch_start_cutoff;
  int nthreads;
  unsigned int step0;  /* first global step index (RNG counter) */
  int tmoves;          /* DMC: 0 none, 1 TMOVES, 2 TMOVESSC */
} orc_run_opts;

/* T-move candidates of one configuration pos[Ne][3] (Pseudopotential::calcNonlocTmove):
 * electron, vxx (< 0) and displacement per candidate, in the reference's order; vnl_out = the
 * nonlocal energy without them.  Returns the count, or -1 if it exceeds cap. */
int orc_tmoves(orc_system* s, const double* pos, const double* rot, int pp_mode, const double* pp_rand,
               int cap, int* e_out, double* vxx_out, double* disp_out, double* vnl_out);
/* One T-move step (Dmc_method::doTmove) of one configuration, in place; returns electrons moved. */
int orc_tmove_step(orc_system* s, double* pos, const double* rot, int pp_mode, const double* pp_rand,
                   double tau, int sc, unsigned long long seed, unsigned wid, unsigned step);

double orc_run(orc_system* s, const orc_run_opts* o, double* pos, double* weights,
               double* eblock, double* state /* [4]: etrial, eref, teff, cutoff sigma */);

#ifdef __cplusplus
}
#endif
#endif
