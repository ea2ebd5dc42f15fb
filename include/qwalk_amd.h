/* qwalk_amd: MI355X walker engine for QWalk's VMC/DMC hot path — C ABI.
 *
 * Plain C types only; no exceptions cross this boundary.  Inputs are host-owned and copied;
 * device state is owned by the engine.  One engine per host thread; each engine owns one
 * HIP stream on the device it was created on.  Every call returns QW_OK (0) or a negative
 * status; qw_last_error() then holds a message.
 *
 * Which reference interface each entry point replaces (paths relative to QWalk's tree):
 *   qw_engine_create     Wavefunction_data / System / Pseudopotential construction from
 *                        the TRIALFUNC / SYSTEM / PSEUDO input
 *                        (src/wavefunction/Wavefunction_data.cpp:32-59,
 *                         src/system/System.cpp:29-49,145-149)
 *   qw_eval_fixed        Wavefunction::updateLap + getVal + getLap per electron, plus
 *                        System::calcKinetic, calcLoc and Pseudopotential::calcNonlocWithTest
 *                        at fixed configurations (src/wavefunction/Wavefunction.h:117-256,
 *                        src/properties/Properties_gather.cpp:40-74,
 *                        src/method/Test_method.cpp:128-176)
 *   qw_parm_count        Wavefunction_data::nparms + getVarParms for the CSF coefficients
 *                        (OPTIMIZE_DET) and the linear Jastrow parameters
 *                        (src/wavefunction/Slat_wf_data.cpp:574-604,
 *                        src/wavefunction/Jastrow2_wf.cpp:477-520,678-687)
 *   qw_parm_deriv        Wavefunction::getParmDeriv of SLATER(-JASTROW) with JASTROW2
 *                        (src/wavefunction/Slat_wf.h:555-686, src/wavefunction/Slat_Jastrow.cpp:171-193,
 *                        src/wavefunction/Jastrow2_wf.cpp:1486-1553,1900-1945,
 *                        src/wavefunction/Wavefunction.cpp:27-113)
 *   qw_method_read       METHOD { VMC | DMC } keyword reading with the reference defaults
 *   qw_method_count/     (src/method/Vmc_method.cpp:33-132, src/method/Dmc_method.cpp:33-160);
 *   qw_method_read_at    every METHOD section in order (src/program_utils/ooqmc.cpp:148-170)
 *   qw_walkers_set/get   Config_save_point restorePos / savePos for a walker population
 *                        (src/method/Vmc_method.cpp:394-399,492)
 *   qw_run               Vmc_method::runWithVariables / Dmc_method::runWithVariables
 *                        (src/method/Vmc_method.cpp:351-558, src/method/Dmc_method.cpp:328-567)
 *   qw_kernel_times      per-kernel device timings (no reference counterpart: the
 *                        reference prints only a wall time per method,
 *                        src/program_utils/ooqmc.cpp:158-168)
 */
#ifndef QWALK_AMD_H
#define QWALK_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qw_engine qw_engine;
typedef int qw_status;

enum {
  QW_OK = 0,
  QW_ERR_ARG = -1,          /* bad argument (null pointer, size out of range) */
  QW_ERR_INPUT = -2,        /* deck could not be read or is inconsistent */
  QW_ERR_UNSUPPORTED = -3,  /* valid QWalk input outside this engine's scope */
  QW_ERR_DEVICE = -4,       /* HIP error */
  QW_ERR_NOMEM = -5         /* device allocation failed */
};

enum {
  QW_PREC_FP64 = 0,        /* everything in FP64 */
  QW_PREC_MIXED_FP32 = 1,  /* LCAO FP32, rank-1 updates FP32, FP64 LU refresh / log-det / sums */
  QW_PREC_MIXED_BF16 = 2,  /* as MIXED_FP32, LCAO on MFMA: BF16 split terms, 3 products */
  QW_PREC_MIXED_BF16X6 = 3 /* as MIXED_BF16 with three-term splits and 6 products */
};

typedef struct qw_engine_opts {
  int precision;            /* QW_PREC_* */
  int device;               /* HIP device ordinal */
  unsigned long long seed;  /* key of the counter-based random streams */
  int refresh_interval;     /* mixed precision: FP64 LU recompute every this many moves
                               (0 = the default: 32 up to 64 electrons of a spin, above that
                               half a sweep rounded up to a multiple of 32; a value >= the
                               electrons of a spin = once per sweep of that spin channel) */
} qw_engine_opts;

typedef struct qw_method_desc {
  int method;               /* 0 = VMC, 1 = DMC */
  int nconfig;              /* walkers */
  int nblock;
  int nstep;
  int ndecorr;              /* VMC: electron sweeps per step */
  double timestep;
  int feedback_interval;    /* DMC: steps between trial-energy updates and branching */
  double feedback;
  double max_weight;        /* DMC: MAX_POSS_WEIGHT */
  double branch_start_cutoff;
  int pp_stochastic;        /* 1 = the reference's default stochastic nonlocal evaluation */
  int tmoves;               /* DMC: 0 = locality approximation only, 1 = TMOVES, 2 = TMOVESSC
                               (Casula T-moves, src/method/Dmc_method.cpp:67-73,850-906) */
  double start_feedback;    /* DMC: feedback of the first trial energy, START_FEEDBACK
                               (src/method/Dmc_method.cpp:97-98,771); <= 0 = use feedback */
} qw_method_desc;

/* Engine for the SYSTEM / PSEUDO / TRIALFUNC sections of a QWalk deck.  INCLUDE and ORBFILE
 * paths are resolved relative to the deck's directory.  On failure *out is NULL and the
 * message is available from qw_last_error(NULL). */
qw_status qw_engine_create(const char* deck_path, const qw_engine_opts* opts, qw_engine** out);
/* An engine for the same deck as src (its parsed tables; the deck is not read again) with
 * other options (precision, device, seed, refresh interval); it gives exactly what
 * qw_engine_create of that deck with these options gives.  No reference counterpart (the
 * reference reads its input once per process); for servers that start many jobs of one
 * deck. */
qw_status qw_engine_clone(const qw_engine* src, const qw_engine_opts* opts, qw_engine** out);
void qw_engine_destroy(qw_engine* e);
const char* qw_last_error(const qw_engine* e);

/* out6 = nup, ndn, natoms, nbasis, has_pseudo, max walkers resident so far */
qw_status qw_engine_info(const qw_engine* e, int* out6);

/* Fixed-configuration evaluation.
 *   pos[K][Ne][3]; rot[9] quadrature rotation (rows = rotated x, y, z axes) or NULL;
 *   pp_mode 0 = deterministic (static cutoff), 1 = stochastic with uniforms pp_rand[K][natoms][Ne]
 * outputs: logpsi[K] (ln|Psi|), sign[K], grad[K][Ne][3] (grad Psi / Psi),
 *          lap[K][Ne] (laplacian Psi / Psi), eparts[K][3] = kinetic, Coulomb, nonlocal. */
qw_status qw_eval_fixed(qw_engine* e, int K, const double* pos, const double* rot, int pp_mode,
                        const double* pp_rand, double* logpsi, double* sign, double* grad,
                        double* lap, double* eparts);

/* Allocation-free qw_eval_fixed.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code recorded into a HIP graph, tight serving loops,
 * harnesses that forbid hipMalloc).  It calls no device allocator (hipMalloc, hipFree or an
 * allocating library routine) and does not synchronise: every pointer is a device pointer on
 * the engine's device, with the layouts of qw_eval_fixed (pos[K][Ne][3], rot[9] or NULL,
 * pp_rand[K][natoms][Ne] when pp_mode = 1; outputs logpsi[K], sign[K], grad[K][Ne][3],
 * lap[K][Ne], eparts[K][3]), and the work is queued on `stream` (a hipStream_t; NULL = the
 * engine's own stream).  The per-configuration state (inverses, MO caches, energies) lives in
 * a fixed device array of QW_SCRATCH_BYTES per GPU, cleared at the start of each call;
 * K above what fits (qw_eval_fixed_static_max) is refused with QW_ERR_ARG.  Every call on one
 * GPU shares that array: calls that overlap on one GPU (two streams, two host threads) race.
 * Systems with more than 128 electrons of one spin are refused (QW_ERR_UNSUPPORTED). */
#define QW_SCRATCH_BYTES (32u << 20)
qw_status qw_eval_fixed_static(qw_engine* e, int K, const double* pos, const double* rot, int pp_mode,
                               const double* pp_rand, double* logpsi, double* sign, double* grad, double* lap,
                               double* eparts, void* stream);
/* Largest K qw_eval_fixed_static accepts for this engine (0 when it accepts none). */
qw_status qw_eval_fixed_static_max(const qw_engine* e, int* kmax);

/* Linear parameters in the reference's order: with OPTIMIZE_DET and several determinants the
 * coefficients of CSFs 2.. first (DETWT input: each determinant is a CSF), then the Jastrow's
 * (group by group: one-body coefficient sections in input order, then two-body coefficients,
 * TWOBODY_SPIN like before unlike; FREEZE pieces have none): *np = their number; values[np] =
 * their values when not NULL. */
qw_status qw_parm_count(const qw_engine* e, int* np, double* values);

/* Derivatives of ln Psi with respect to the parameters of qw_parm_count at fixed
 * configurations pos[K][Ne][3] (the Parm_deriv_return of the reference):
 *   gradient[K][np]           d ln Psi / dp
 *   hessian[K][np][np]        gradient_i * gradient_j, zero between two CSF coefficients
 *                             (may be NULL)
 *   gradderiv[K][np][Ne][4]   d/dp of grad_e ln Psi (3) and of lap_e Psi / Psi
 *   val_gradient[K][Ne][3]    gradient of ln Psi_S (the determinants) plus that of the Jastrow
 *                             pieces with parameters
 * As in the reference, FREEZE pieces take no part in val_gradient or in the laplacian
 * cross term 2 gradderiv . val_gradient. */
qw_status qw_parm_deriv(qw_engine* e, int K, const double* pos, double* gradient, double* hessian,
                        double* gradderiv, double* val_gradient);

/* Read the first METHOD section of a deck into *out (reference defaults applied). */
qw_status qw_method_read(const char* deck_path, qw_method_desc* out);

/* A deck may hold several METHOD sections, run one after another on the same walkers
 * (src/program_utils/ooqmc.cpp:148-170).  *n = their number. */
qw_status qw_method_count(const char* deck_path, int* n);
/* The index-th METHOD section (0-based) and its log label (LABEL, default "vmc" / "dmc";
 * src/method/Vmc_method.cpp:65-66, src/method/Dmc_method.cpp:94-95) into label[label_len]
 * (may be NULL). */
qw_status qw_method_read_at(const char* deck_path, int index, qw_method_desc* out, char* label, int label_len);

/* Walker population: pos[nw][Ne][3], weights[nw] (NULL = all 1). */
qw_status qw_walkers_set(qw_engine* e, int nw, const double* pos, const double* weights);
qw_status qw_walkers_get(qw_engine* e, int nw, double* pos, double* weights);
/* The resident walkers' determinant state and last local energy, as the method loop holds
 * them (after Sherman-Morrison updates since the last recompute; the reference keeps the same
 * quantities in each walker's Slat_wf storage and Properties_point,
 * src/wavefunction/Slat_wf.h:293-401, src/properties/Properties_point.h:30-90):
 *   logdet[nw][2] ln|det| per spin, sign[nw][2], elast[nw][4] = kinetic, Coulomb, nonlocal,
 *   total of the last local-energy evaluation, dgl[nw][Ne][4] = grad_e ln|D| (3) and
 *   lap_e D / D of the determinants, from the inverse the last sweep of each spin ended with
 *   (src/wavefunction/Slat_wf.h:1021-1080).  Any pointer may be NULL. */
qw_status qw_walkers_state(qw_engine* e, int nw, double* logdet, double* sign, double* elast, double* dgl);

/* Walker files (<runid>.config), host only, in the reference's text layout so either program
 * restarts from the other's checkpoint.
 * qw_configs_read: read_configurations<Dmc_point> (src/program_utils/qmc_io.h:427-503,
 *   src/system/Sample_point.cpp:120-139, src/method/Dmc_method.cpp:1252-1265).  *nw = walkers
 *   in the file; the first min(*nw, max_nw) go to pos[.][Ne][3], weights[.] (1 for a walker
 *   without a DMC record) and signs[.]; any of the three may be NULL (count only).  A walker
 *   with another electron count is an error (QW_ERR_INPUT).
 * qw_configs_write: write_configurations (src/program_utils/qmc_io.h:371-423) of
 *   Config_save_point (weights NULL: VMC, src/system/Sample_point.cpp:140-148) or Dmc_point
 *   records (weights given; signs NULL = all 1; src/method/Dmc_method.cpp:1228-1250); the old
 *   file becomes <path>.backup.
 * qw_method_files: READCONFIG and STORECONFIG of the index-th METHOD section, each
 *   <deck>.config when absent (src/method/Vmc_method.cpp:59-62, src/method/Dmc_method.cpp:62-63,
 *   91-92), into buffers of len bytes.
 * qw_configs_error: message of the calling thread's last failed call of these three. */
qw_status qw_configs_read(const char* path, int ne, int max_nw, int* nw, double* pos, double* weights, int* signs);
qw_status qw_configs_write(const char* path, int nw, int ne, const double* pos, const double* weights,
                           const int* signs);
qw_status qw_method_files(const char* deck_path, int index, char* readconfig, char* storeconfig, int len);
const char* qw_configs_error(void);

/* Fields of one block record written by qw_run (the quantities of a QWalk block log,
 * src/properties/Properties.cpp:578-641, src/properties/Properties_block.cpp:73-138):
 *   0 E, 1 kinetic, 2 Coulomb, 3 nonlocal      weighted block averages
 *   4 weight sum, 5 points (walker-steps)
 *   6..9 per-point variances of E, kinetic, Coulomb, nonlocal
 *   10 mean weight, 11 weight variance */
#define QW_BLOCK_FIELDS 12

/* Random initial walkers as the reference's Molecular_sample::randomGuess
 * (src/system/Molecular_sample.cpp:194-260): electrons of each spin in cubes of half-width
 * inirange around the atoms by charge, the rest around the centroid.  pos[nw][Ne][3]. */
qw_status qw_walkers_random(const qw_engine* e, int nw, unsigned long long seed, double inirange, double* pos);

/* Run VMC or DMC on the resident walkers.
 *   eblock[nblock][QW_BLOCK_FIELDS] block records (see above);
 *   stats[4] = acceptance, electron moves done, seconds, electron-moves/s. */
qw_status qw_run(qw_engine* e, const qw_method_desc* m, double* eblock, double* stats);

/* Per-kernel device time of the last qw_run: names (comma separated) into buf, and
 * avg_ms[i], count[i] for up to n kernels; returns the number of kernels in *nk. */
qw_status qw_kernel_times(const qw_engine* e, char* buf, int buflen, double* avg_ms, int* count,
                          int n, int* nk);


/* Device bytes the engine allocates for a population of nw walkers (walker state, MO caches,
 * inverses), excluding the system tables already resident.  Used for admission by HBM
 * budget.  No reference counterpart: the reference sizes nothing on a device. */
qw_status qw_engine_walker_bytes(const qw_engine* e, int nw, unsigned long long* bytes);

/* Launch choices the engine made (no reference counterpart): out4 = LCAO coefficient loads in
 * flight chosen by the online tuner (0 while it is still measuring), HIP graphs used for DMC
 * steps (0/1), block-per-walker kernels for more than 64 electrons of a spin (0/1), register
 * tile width of the inverse. */
qw_status qw_engine_tuning(const qw_engine* e, int* out4);

/* ---------------------------------------------------------------- batches of systems
 * Several engines (systems of different sizes on one device, all fp64 or all mixed_fp32) run
 * DMC together: each step is one sweep launch per spin channel (on the mixed path per launch
 * interval, the largest common divisor of the engines' recompute intervals, each followed by
 * one recompute launch covering the systems whose own interval ends there) covering every
 * system, one local-energy launch, and per feedback interval one branching
 * and one copy launch, recorded once as a HIP graph and replayed.  The blocks of a launch read
 * their system from a slot table; the LCAO coefficient loads in flight are tuned online per
 * size bucket (systems with the same register tile).  Each engine keeps its own walkers,
 * random streams, DMC control and block records, and its results in a batch are bitwise those
 * of qw_run on that engine alone.  Not batched: several determinants, T-moves, and on the
 * mixed path systems above 128 electrons of a spin (QW_ERR_UNSUPPORTED).  No reference
 * counterpart: QWalk runs one system per process (src/program_utils/ooqmc.cpp:148-170; the
 * per-system loop is src/method/Dmc_method.cpp:328-567). */
typedef struct qw_batch qw_batch;
qw_status qw_batch_create(int n, qw_engine* const* engines, qw_batch** out);
void qw_batch_destroy(qw_batch* b);
/* Message of the last failure (b = NULL: of qw_batch_create). */
const char* qw_batch_error(const qw_batch* b);
/* DMC (m->method = 1) on every engine's resident walkers (qw_walkers_set first).
 *   eblock[n][nblock][QW_BLOCK_FIELDS] block records in engine order;
 *   stats[4 + n] = electron moves, seconds, electron-moves/s, tuner finished (0/1), then each
 *   engine's acceptance (tau_eff / tau). */
qw_status qw_batch_run(qw_batch* b, const qw_method_desc* m, double* eblock, double* stats);
/* Continuous batching (the serving loop's launches): qw_batch_open makes an empty batch;
 * engines (walkers set) join it at any time with the runs they are to do (e.g. a VMC
 * equilibration, then their method; VMC and DMC runs mix in one batch), and every
 * qw_batch_step advances every member by one step (DMC: one step of its feedback interval;
 * VMC: one step of ndecorr sweeps) in launches shared by all members of a method: per spin
 * channel and recompute interval one sweep launch and one recompute launch, then one local-
 * energy launch, and one branching launch for the DMC members whose feedback interval ends.
 * A member's results are bitwise those of qw_run on its engine alone, whichever members it
 * shares the launches with.  finished[*nfinished] (up to 64) = members whose runs ended in
 * this step; qw_batch_leave then gives the last run's eblock[nblock][QW_BLOCK_FIELDS] and
 * stats[4] (as qw_run) and removes the member (the caller then owns the engine again). */
/* refresh_interval (mixed path): the batch's launch interval (0 = 32); a member's own recompute
 * interval must be a multiple of it or cover a whole sweep of its spin channels. */
qw_status qw_batch_open(int device, int precision, int refresh_interval, qw_batch** out);
qw_status qw_batch_join(qw_batch* b, qw_engine* e, int nruns, const qw_method_desc* runs, int* member);
qw_status qw_batch_step(qw_batch* b, int* finished, int* nfinished);
qw_status qw_batch_leave(qw_batch* b, int member, double* eblock, double* stats);
/* Device time per launch kind of the last run: avg_ms4[k], count4[k] for k = sweep, inverse
 * recompute, local energy, branching (+ copies); tuning[2i], tuning[2i+1] = register tile of
 * size bucket i and its LCAO loads in flight (up to 5 buckets; may be NULL). */
qw_status qw_batch_kernel_times(const qw_batch* b, double* avg_ms4, int* count4, int* tuning);

/* ---------------------------------------------------------------- serving loop
 * A request loop over the engine: VMC/DMC jobs (decks of different sizes) are submitted at
 * any time, admitted in arrival order while the sum of their device bytes fits the HBM budget
 * (and fewer than max_active run), and run concurrently, each on its own engine, stream and
 * host thread; a finished job frees its memory and the next pending job is admitted.  A job's
 * results depend only on its own deck, seed and method, so they are the same alone and in a
 * mix.  No reference counterpart: QWalk runs one method per process
 * (src/program_utils/ooqmc.cpp:120-168); this is the batching/serving layer above it. */
typedef struct qw_server qw_server;

typedef struct qw_server_opts {
  int device;
  unsigned long long hbm_budget;  /* bytes of walker state admitted at once (0 = 64 GiB) */
  int max_active;                 /* jobs running at once (0 = 8) */
} qw_server_opts;

typedef struct qw_job_desc {
  int precision;                  /* QW_PREC_* */
  unsigned long long seed;
  int nwalkers;
  int equil_steps;                /* VMC steps from the random guess before the method (0 = none) */
  qw_method_desc method;          /* nconfig is ignored (nwalkers is used) */
} qw_job_desc;

/* stats of a finished job (seconds on the server's clock, from qw_server_create) */
#define QW_JOB_STATS 10  /* 0 submit, 1 admitted, 2 finished, 3 latency (finished - submit),
                            4 electron moves, 5 run seconds, 6 electron-moves/s, 7 device bytes,
                            8 set-up seconds (engine and walkers), 9 1 = ran in a batch */

qw_status qw_server_create(const qw_server_opts* opts, qw_server** out);
/* Queue a job; the deck is read when the job is admitted.  *job_id >= 1. */
qw_status qw_server_submit(qw_server* s, const char* deck_path, const qw_job_desc* job, long long* job_id);
/* Block until the job has finished; eblock[method.nblock][QW_BLOCK_FIELDS], stats[QW_JOB_STATS].
 * Returns the job's own status (its error message from qw_server_error). */
qw_status qw_server_wait(qw_server* s, long long job_id, double* eblock, double* stats);
const char* qw_server_error(qw_server* s, long long job_id);
/* Waits for all jobs, then releases the server. */
void qw_server_destroy(qw_server* s);

#ifdef __cplusplus
}
#endif
#endif
