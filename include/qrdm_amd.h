/*
 * qrdm_amd.h -- C ABI of the MI355X DM-QR library (libqrdm_amd.so).
 *
 * QR factorisation with Deviation Maximization column pivoting, A*P = Q*R,
 * computed on one AMD Instinct MI355X (gfx950) by hand-written HIP kernels.
 * All signatures use plain C types; lapack_int is 32-bit int.
 */
#ifndef QRDM_AMD_H_
#define QRDM_AMD_H_

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Drop-in for the reference's public entry point
 *   lapack_int dgeqrdm(int matrix_layout, lapack_int m, lapack_int n,
 *                      double* a, lapack_int lda, lapack_int* jpvt,
 *                      double* tau, lapack_int* ncols, double* thres,
 *                      int DM_procedure);
 * (include/QRDM.h:19-22, src/dgeqrdm.c:5-16, src/dgeqrdm_work.c:420-485).
 *
 * All arrays are HOST memory owned by the caller and updated in place:
 *   a      lda x n column-major; on exit R (upper trapezoid) and the
 *          Householder vectors below the diagonal;
 *   jpvt   on entry 0 for free columns (non-zero = fixed leading column);
 *          on exit 1-based: column j of A*P is column jpvt[j] of A;
 *   tau    min(m,n) reflector scalars;
 *   ncols  on entry ncols[0] in {1,2,3} enables the stopping test
 *          (eta = eps*n, eps*sqrt(n), thres[2]); on exit ncols[it] is the
 *          number of columns reduced at DM step it (room for n entries);
 *   thres  thres[0] = delta in [0,1] (cosine bound), thres[1] = tau in [0,1]
 *          (norm threshold), thres[2] = eta for stop mode 3;
 *   nb     maximum DM block (k_max), 1..256 on this device (steps wider than
 *          64 columns run host-driven, one synchronisation per step, FP64
 *          without the qrdm_opts forms).
 * Returns 0 on success; -1 for an illegal argument (a message is printed, as
 * the reference's xerbla_ does); -10 for nb > 256, which this device build
 * does not support (the reference accepts any nb > 0); -13 when the input
 * holds a NaN or an infinity (A is then left untouched);
 * <= -1000 for a device error (see qrdm_last_error).  matrix_layout 101 is
 * accepted and treated as column-major, as the reference's DM code does.
 */
int dgeqrdm(int matrix_layout, int m, int n, double *a, int lda, int *jpvt, double *tau,
            int *ncols, double *thres, int nb);

/*
 * Device-resident variant of dgeqrdm: d_a (lda x n) and d_tau (min(m,n)) are
 * DEVICE pointers on the current HIP device; jpvt, ncols, thres stay in host
 * memory.  Work is queued on `stream` (a hipStream_t, NULL = default stream)
 * and the call returns after it completes.  No counterpart in the reference:
 * it is the form the host entry point above runs after its copies, exposed
 * for callers whose matrix already lives in HBM.
 */
int qrdm_dgeqrdm_dev(int matrix_layout, int m, int n, double *d_a, int lda, int *jpvt,
                     double *d_tau, int *ncols, const double *thres, int nb, void *stream);

/*
 * qrdm_dgeqrdm_dev that allocates nothing: no counterpart in the reference.
 * For callers that cannot allocate during a call (serving loops and
 * harnesses that forbid hipMalloc).  Same arguments, arithmetic and results
 * as qrdm_dgeqrdm_dev; its device workspace is a fixed 64 MiB array inside
 * the library (one per device, held until the process ends), enough for
 * e.g. m = n = 4096 or m = 16384, n = 2048; a larger problem returns -14,
 * and nb > 64 returns -10 (messages on stderr) before any device work.  No device or pinned memory,
 * events or memset nodes are created: regions that must start at zero are
 * cleared by a kernel, and the DM loop's state is read back into host memory
 * of the library, waiting on the stream.  Calls that overlap on one device
 * share the array and race; the call synchronises the stream, so it cannot
 * be captured into a graph.
 */
int qrdm_dgeqrdm_dev_noalloc(int matrix_layout, int m, int n, double *d_a, int lda, int *jpvt,
                             double *d_tau, int *ncols, const double *thres, int nb,
                             void *stream);

/* Precision of the trailing update's two products (qrdm_opts.trailing).
 * FP32 and BF16X3 keep the FP64 products until the remaining trailing block
 * is small enough for the lower precision to stay within 1e-12 of A
 * (maxnrm * sqrt(n - j) <= 1e-12 * max_initial_column_norm / (64 u_low),
 * u_low = 2^-24 for FP32, 2^-17 for BF16X3), then use the lower one; the
 * _ALL forms use it in every step. */
#define QRDM_TRAILING_FP64 0   /* FP64 matrix cores (the plain path) */
#define QRDM_TRAILING_FP32 1   /* FP32 matrix cores, FP32 accumulation, scheduled */
#define QRDM_TRAILING_BF16X3 2 /* split BF16 (hi*hi + hi*lo + lo*hi), FP32 accumulation, scheduled */
#define QRDM_TRAILING_FP32_ALL 3   /* FP32 in every step */
#define QRDM_TRAILING_BF16X3_ALL 4 /* split BF16 in every step */

/* Options of qrdm_dgeqrdm_dev_opt; a NULL pointer means all defaults. */
typedef struct {
  int trailing;       /* QRDM_TRAILING_*: only C := C - V T^T V^T C changes
                         precision (C itself stays FP64 in memory); the panel,
                         the norm downdate/recompute and the cosine test stay
                         FP64 */
  const void *policy; /* NULL, or a qrdm_policy in DEVICE memory: the learned
                         network proposes each step's tau (candidate count)
                         and delta (cosine bound); the panel's truncation keeps
                         thres[1] */
} qrdm_opts;

/*
 * The learned (tau, delta) proposal: a perceptron 34 -> 64 -> 64 -> 2 (ReLU,
 * sigmoid outputs) with bfloat16 weights and FP32 accumulation.  Inputs: a
 * 32-bin histogram of the remaining partial norms divided by their maximum
 * (bin floor(32 r), clamped to 31; counts divided by n - j), the step
 * fraction j / min(m, n), and min((m - j) / (n - j), 8) / 8.  Outputs y0, y1
 * map to tau = tau_lo + (tau_hi - tau_lo) y0, delta = delta_lo +
 * (delta_hi - delta_lo) y1; a non-finite or out-of-range proposal falls back
 * to thres.  No counterpart in the reference.
 */
typedef struct {
  unsigned short w1[64 * 34]; /* [out][in], bfloat16 bit patterns */
  unsigned short w2[64 * 64];
  unsigned short w3[2 * 64];
  float b1[64], b2[64], b3[2];
  float tau_lo, tau_hi, delta_lo, delta_hi;
} qrdm_policy;

/* Per-step record of the DM loop (debugging and step-by-step parity). */
typedef struct {
  int j;         /* 0-based first column of the step */
  int kmax;      /* min(nb, m - j, n - j) */
  int nc;        /* number of norm candidates (> tau * max, at most kmax) */
  int fjb;       /* columns accepted by the cosine test */
  int fjb_cmp;   /* columns reduced by the panel (<= fjb) */
  int nswap;     /* column swaps made by the permutation */
  int panel;     /* 1: Gram-based panel, 0: column-by-column panel */
  double maxnrm; /* largest partial norm after the step's downdate */
  int sel[64];   /* accepted columns relative to j, in acceptance order */
  double tau;    /* the step's tau and delta (thres, or the learned proposal) */
  double delta;
  int prec;      /* the step's trailing-update precision (QRDM_TRAILING_FP64/FP32/BF16X3) */
} qrdm_step;

/*
 * qrdm_dgeqrdm_dev plus a trace: up to ntrace records are written to trace
 * and the number of steps taken to *nsteps.  Same arithmetic as above.
 * Setting the environment variable QRDM_PANEL=exact makes every step use the
 * column-by-column panel (both panel forms satisfy the same parity bar).
 */
int qrdm_dgeqrdm_dev_trace(int matrix_layout, int m, int n, double *d_a, int lda, int *jpvt,
                           double *d_tau, int *ncols, const double *thres, int nb, void *stream,
                           qrdm_step *trace, int ntrace, int *nsteps);

/*
 * qrdm_dgeqrdm_dev_trace with options (opts may be NULL).  No counterpart in
 * the reference, which has only the FP64 path.  The scheduled reduced-
 * precision forms keep the backward error within 1e-12 (they switch only
 * where the remaining block allows it, and may never switch); the _ALL
 * forms trade that agreement for speed (backward error about the lower
 * precision's unit roundoff times the matrix's norm).
 */
int qrdm_dgeqrdm_dev_opt(int matrix_layout, int m, int n, double *d_a, int lda, int *jpvt,
                         double *d_tau, int *ncols, const double *thres, int nb,
                         const qrdm_opts *opts, void *stream, qrdm_step *trace, int ntrace,
                         int *nsteps);

/*
 * Batched DM-QR: many independent factorisations (each with dgeqrdm's
 * semantics, free columns, the column-by-column panel) in ONE persistent
 * kernel launch.  No counterpart in the reference, which factorises one
 * matrix per call.
 *
 * Matrix i is column-major m x n with leading dimension lda at
 * d_a + a_off; its tau (min(m, n)) at d_tau + tau_off; its jpvt and ncols
 * (n entries each, jpvt 1-based on exit) at d_jpvt + col_off and
 * d_ncols + col_off; its info (0, or -13 for a non-finite entry, the matrix
 * then untouched) at d_info[i].  All device pointers.  Limits: 1 <= n <= 1024,
 * m >= 1, lda >= m.
 */
typedef struct {
  int m, n, lda, reserved;
  long long a_off, tau_off, col_off;
} qrdm_batch_item;

typedef struct qrdm_batch_plan qrdm_batch_plan;

/* Uploads the descriptors and the order the workgroups take the matrices in
 * (decreasing flop count; each workgroup takes the next one from a device
 * counter, so launches of one plan must not overlap).  nb: the DM
 * block (k_max) of every factorisation, 1..64; nb <= 0 lets
 * qrdm_batch_autotune choose it among 16, 32, 64 (it changes the pivots).
 * Returns NULL on error (text through qrdm_last_error). */
qrdm_batch_plan *qrdm_batch_plan_create(int count, const qrdm_batch_item *items, int nb);
void qrdm_batch_plan_destroy(qrdm_batch_plan *plan);
/* out[0..5) = {waves per workgroup, workgroups, nb, timed candidates,
 * panel rows (see qrdm_batch_plan_set_panel_rows)}; returns 5. */
int qrdm_batch_plan_query(const qrdm_batch_plan *plan, int *out, int nout);
/* Uses a learned (tau, delta) network (qrdm_policy in device memory, or
 * NULL for thres) in every factorisation of the batch. */
int qrdm_batch_plan_set_policy(qrdm_batch_plan *plan, const void *d_policy);
/* One eager run of the batch that also writes a training record per DM
 * step: 40 floats = the 34 network inputs, the label (tau, delta) scaled to
 * [0, 1] over ranges = {tau_lo, tau_hi, delta_lo, delta_hi}, the block the
 * defaults reduce, the label's block, j, n - j.  The label is the point of a
 * (tau, delta) grid whose greedy acceptance yields the largest block that
 * the panel's truncation test (default tau) would keep, judged on the
 * candidates' Cholesky factor.  Up to cap records go to d_records; *d_count
 * (device, zero on entry) counts all steps. */
int qrdm_batch_trace(qrdm_batch_plan *plan, double *d_a, double *d_tau, int *d_jpvt,
                     int *d_ncols, int *d_info, const double *thres, float *d_records, int cap,
                     unsigned *d_count, const float *ranges, void *stream);
/* Sets the launch shape by hand: waves per workgroup, 4 (the one shape
 * built; 8 is switched off and returns -1). */
int qrdm_batch_plan_set_shape(qrdm_batch_plan *plan, int waves);
/* The panel form by hand: a DM step whose panel has at least `rows` rows
 * (and two or more columns) takes the Gram-based form -- Cholesky of the
 * panel's Gram, one re-orthogonalisation (CholeskyQR2), Householder vectors
 * reconstructed from the Q factor, T from V^T V -- the others the
 * column-by-column dgeqr2 form; rows <= 0: column by column only.  A new
 * plan starts at 128 rows.  Both forms meet the same parity bar; they round
 * differently, so the choice can change pivots decided by rounding.  No
 * counterpart in the reference. */
int qrdm_batch_plan_set_panel_rows(qrdm_batch_plan *plan, int rows);
/* Times the candidate launches (nb in {16, 32, 64} when the plan was
 * created with nb <= 0, else the plan's nb; times panel rows in {0 (column
 * by column), 128, 256}; 4 waves per workgroup) on a scratch copy of the
 * batch d_a (left unchanged) and keeps the fastest.  thres = {delta, tau}.
 * Synchronous. */
int qrdm_batch_autotune(qrdm_batch_plan *plan, const double *d_a, const double *thres,
                        void *stream);
/* Factorises the whole batch in place: one kernel launch on stream,
 * asynchronous (results are valid once the stream has completed it).
 * thres = {delta, tau, eta}; stop_mode as ncols[0] of dgeqrdm (0..3).
 * use_graph != 0: the launch is captured once into a hipGraph (per set of
 * pointers and parameters) and replayed with hipGraphLaunch. */
int qrdm_dgeqrdm_batched(qrdm_batch_plan *plan, double *d_a, double *d_tau, int *d_jpvt,
                         int *d_ncols, int *d_info, const double *thres, int stop_mode,
                         int use_graph, void *stream);

/*
 * Training of the proposal network on the device (no counterpart in the
 * reference).  theta: the FP32 master parameters, W1 (64 x 34, row-major),
 * b1, W2 (64 x 64), b2, W3 (2 x 64), b3; the forward pass uses their
 * bfloat16 roundings.  x: ns x 34 inputs, y: ns x 2 targets in [0, 1];
 * loss = (1/ns) sum (sigmoid(z) - y)^2.  All pointers device memory.
 * qrdm_policy_sizes: out = {parameters, sizeof(qrdm_policy), floats per
 * training record}.
 */
int qrdm_policy_sizes(int *out, int nout);
/* d_grad = dloss/dtheta, d_loss[0] = loss. */
int qrdm_policy_grad(const float *d_theta, const float *d_x, const float *d_y, int ns,
                     float *d_grad, float *d_loss, void *stream);
/* One Adam step (beta 0.9, 0.999, eps 1e-8, step number t >= 1) on the
 * minibatch; d_m1, d_m2 the moments (zero before the first step). */
int qrdm_policy_adam(float *d_theta, float *d_m1, float *d_m2, const float *d_x,
                     const float *d_y, int ns, float lr, int t, float *d_loss, void *stream);
/* theta -> qrdm_policy (bfloat16 weights); ranges (host) = {tau_lo, tau_hi,
 * delta_lo, delta_hi}. */
int qrdm_policy_pack(const float *d_theta, const float *ranges, void *d_net, void *stream);
/* d_out[2 s + o] = the network's sigmoid outputs for the inputs d_x[34 s..]. */
int qrdm_policy_predict(const void *d_net, const float *d_x, int ns, float *d_out,
                        void *stream);

/*
 * Unpivoted Householder QR on the device (LAPACK dgeqrf semantics; the
 * reference's QRDM.QRF forwards to LAPACKE_dgeqrf, QRDM_wrapper.c:44-68).
 * d_a, d_tau device pointers.
 */
int qrdm_dgeqrf_dev(int m, int n, double *d_a, int lda, double *d_tau, void *stream);

/*
 * C := Q C (trans = 0) or Q^T C (trans = 1) from the left, Q = H(1)...H(k)
 * held in d_a/d_tau as dgeqrdm or dgeqrf leave them (the reference's
 * QRDM.DORMQR forwards to LAPACKE_dormqr(layout,'L','N',...),
 * QRDM_wrapper.c:104-126).  All matrices are device pointers.
 */
int qrdm_dormqr_dev(int trans, int m, int n, int k, const double *d_a, int lda,
                    const double *d_tau, double *d_c, int ldc, void *stream);

/* Host-pointer forms of the two routines above (copies in and out). */
int qrdm_dgeqrf(int m, int n, double *a, int lda, double *tau);
int qrdm_dormqr(int trans, int m, int n, int k, const double *a, int lda, const double *tau,
                double *c, int ldc);

/*
 * Host index walk of one DM permutation (permute_marked,
 * src/dgeqrdm_work.c:149-262, with marks 0/1): given the accepted columns
 * sel[0..nsel) (relative positions < ncols, acceptance order), writes the
 * swap sequence it performs as pairs swaps[2*i], swaps[2*i+1] and returns
 * their number (at most ncols: every swap advances the front pointer).  No counterpart in the reference's ABI: it
 * is exposed so the host side of the permutation can be checked without a
 * GPU.  Returns -1 for invalid input.
 */
int qrdm_permute_walk(int ncols, const int *sel, int nsel, int *swaps);

/*
 * Kernel-time accounting for measurement.  While enabled, the driver records
 * HIP events (on the stream it launches on) around each class of kernel and
 * accumulates, per class, the summed event time in ms, the launch count and
 * the algorithmic work of those launches.  Classes:
 *   0 trailing update, partial products V^T C (k_gemm1) work = flops
 *   1 norm downdate/recompute                          work = HBM bytes
 *   2 DM selection (top-k, candidate Gram, greedy)     work = HBM bytes
 *   3 panel (Householder + V + T)                      work = flops
 *   4 column permutation (on tall panels with the      work = HBM bytes
 *     panel's first Cholesky beside it)
 *   5 initial norms                                    work = HBM bytes
 *   6 trailing update, second half (C -= V W)          work = HBM bytes
 *   7 trailing update, W = T^T (sum of the partials)   work = flops of T^T
 * qrdm_profile_read fills out[3*c + {0,1,2}] = {ms, launches, work} for
 * c < nclass and returns the number of classes.  No counterpart in the
 * reference (it has no timing facility).
 */
void qrdm_profile_enable(int on);
/* Classes timed while enabled: bit c selects class c (default: all).  Each
 * timed region costs two event records on the stream (a few microseconds of
 * idle device time each), so a measurement of the whole factorisation should
 * time only the classes it reports. */
void qrdm_profile_select(unsigned mask);
void qrdm_profile_reset(void);
int qrdm_profile_read(double *out, int nclass);

/* Text of the last device error seen by this thread's calls ("" if none). */
const char *qrdm_last_error(void);

/* Library version string. */
const char *qrdm_version(void);

#ifdef __cplusplus
}
#endif

#endif /* QRDM_AMD_H_ */
