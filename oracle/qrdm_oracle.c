/*
 * qrdm_oracle.c -- CPU restatement of the reference's DM-QR (dgeqrdm) path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the GPU library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (qrdm_amd/) never links, calls or falls back to it.
 *
 * It restates, in plain C with its own loops (no BLAS), the arithmetic the
 * reference performs through CBLAS/LAPACKE:
 *
 *   dgeqrdm / dgeqrdm_work  src/dgeqrdm.c:5-16, src/dgeqrdm_work.c:420-838
 *   norm_update             src/dgeqrdm_work.c:36-122
 *   cmpStruct + qsort       src/dgeqrdm_work.c:126-142, :345
 *   permute_marked          src/dgeqrdm_work.c:149-262
 *   DM_perm                 src/dgeqrdm_work.c:280-418
 *   dgeqr2_mia              src/dgeqr2.c:29-195
 *   dlarfg_mia + d_sign     src/dlarfg.c:21-188
 *   dlarf_ (side = left)    src/dlarf.c:28-211
 *   LAPACKE_dlarfb_mia      src/dlarfb.c:40-151 (NaN checks, then dlarfb 'L','T','F','C')
 *   LAPACKE_dlarft ('F','C') call site src/dgeqrdm_work.c:751-754 (LAPACK semantics)
 *
 * Third-party arithmetic restated here: BLAS dnrm2 (extended-precision sum
 * of squares, as OpenBLAS's x86-64 kernel), dlapy2, iladlc, dgemv/dger, dsyrk, dlarft, dlarfb, dgeqrf
 * and dormqr semantics of LAPACK 3.x.  The reference links an unpinned system
 * BLAS/LAPACK (setup_QRDM.py:13,38); the compiled reference used to pin this
 * file links OpenBLAS 0.3.28 as bundled with scipy 1.15.3.
 *
 * Pinning (tests/test_oracle.py, tests/test_abi.py):
 *   - whole driver: against the reference compiled from its own sources
 *     (oracle/_ref, built by oracle/build_ref.sh), jpvt/ncols identical and
 *     R, V, tau within 1e-12 relative, on committed seeded fixtures;
 *   - o_nrm2/o_dlarfg/o_dlarf/o_dlarft/o_dlarfb: closed forms and LAPACK
 *     (scipy.linalg.lapack) known answers;
 *   - o_norm_update: closed form with forced cancellation (recompute branch);
 *   - o_dm_select: brute force and invariants on tiny inputs;
 *   - o_permute_marked: exhaustive small cases against the compiled reference
 *     (through whole-driver runs) and permutation invariants.
 *   - fixed-column pre-pass with nfxd > 0 (src/dgeqrdm_work.c:594-635): the
 *     reference's DM loop is undefined behaviour there (sort keys initialised
 *     at absolute offsets, read at relative ones: :678-681 vs :322,:354-357);
 *     this file defines the corrected semantics.  Pinned by mathematics only
 *     (leading block = unpivoted QR of the fixed columns).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <float.h>
#include <stdint.h>

#define OMIN(a, b) ((a) <= (b) ? (a) : (b))
#define OMAX(a, b) ((a) >= (b) ? (a) : (b))

/* LAPACK dlamch('E'): relative machine precision (rounding) = 2^-53. */
static const double O_EPS = 1.1102230246251565e-16;
/* LAPACK dlamch('S'): safe minimum = DBL_MIN. */
static const double O_SFMIN = 2.2250738585072014e-308;

typedef struct {
  double val;
  int idx;
} o_key;

/* One record per DM step; filled when the caller passes a trace buffer. */
typedef struct {
  int j;        /* 0-based first column of the step */
  int kmax;     /* min(nb, m_sub, n_sub) */
  int nc;       /* number of norm candidates */
  int fjb;      /* DM-accepted width */
  int fjb_cmp;  /* width reduced by the panel */
  int nswap;    /* column swaps made by permute_marked */
  double maxnrm;/* max partial norm after the step's downdate */
  int sel[64];  /* idxI[0..min(fjb,64)) relative to j, selection order */
} oracle_step;

/* ------------------------------------------------------------------ BLAS-1 */

/* dnrm2 as OpenBLAS's x86-64 kernel computes it: the sum of squares in
 * extended (x87 80-bit) precision, then the square root, rounded once to
 * double.  The 15-bit exponent keeps every square of a double finite and
 * normal, and sums of squares of small integers are exact, so columns with
 * equal exact norms get equal computed norms (the sort's ties survive). */
static double o_nrm2(int n, const double *x, int incx) {
  if (n < 1 || incx < 1) return 0.0;
  if (n == 1) return fabs(x[0]);
  long double s = 0.0L;
  for (int i = 0; i < n; i++) {
    long double v = (long double)x[(size_t)i * incx];
    s += v * v;
  }
  return (double)sqrtl(s);
}

/* dlapy2: sqrt(x^2 + y^2) without unnecessary overflow. */
static double o_dlapy2(double x, double y) {
  double xa = fabs(x), ya = fabs(y);
  double w = OMAX(xa, ya), z = OMIN(xa, ya);
  if (z == 0.0 || w > DBL_MAX) return w;
  double q = z / w;
  return w * sqrt(1.0 + q * q);
}

/* f2c d_sign(a, b): |a| with the sign test b >= 0 (so -0.0 counts as +). */
static double o_dsign(double a, double b) {
  double x = fabs(a);
  return b >= 0 ? x : -x;
}

/* iladlc: index (1-based) of the last non-zero column of the m x n matrix. */
static int o_iladlc(int m, int n, const double *a, int lda) {
  if (n == 0) return 0;
  if (a[(size_t)(n - 1) * lda] != 0.0 || a[(m - 1) + (size_t)(n - 1) * lda] != 0.0)
    return n;
  for (int c = n; c >= 1; c--)
    for (int r = 0; r < m; r++)
      if (a[r + (size_t)(c - 1) * lda] != 0.0) return c;
  return 0;
}

/* ------------------------------------------------------- Householder steps */

/*
 * dlarfg_mia (src/dlarfg.c:28-188): LAPACK 3.2 dlarfg with the DM truncation
 * test.  Returns 1 (and leaves alpha, x, tau untouched) when idx_house > 1 and
 * ||x|| < thres; otherwise generates H with beta = -sign(alpha)*||(alpha,x)||.
 */
int o_dlarfg(int n, double *alpha, double *x, double *tau, double thres, int idx_house) {
  if (n <= 1) {
    *tau = 0.0;
    return 0;
  }
  double xnorm = o_nrm2(n - 1, x, 1);
  if (idx_house > 1 && xnorm < thres) return 1;
  if (xnorm == 0.0) {
    *tau = 0.0;
    return 0;
  }
  double beta = -o_dsign(o_dlapy2(*alpha, xnorm), *alpha);
  double safmin = O_SFMIN / O_EPS;
  int knt = 0;
  if (fabs(beta) < safmin) {
    double rsafmn = 1.0 / safmin;
    do {
      ++knt;
      for (int i = 0; i < n - 1; i++) x[i] *= rsafmn;
      beta *= rsafmn;
      *alpha *= rsafmn;
    } while (fabs(beta) < safmin);
    xnorm = o_nrm2(n - 1, x, 1);
    beta = -o_dsign(o_dlapy2(*alpha, xnorm), *alpha);
  }
  *tau = (beta - *alpha) / beta;
  double s = 1.0 / (*alpha - beta);
  for (int i = 0; i < n - 1; i++) x[i] *= s;
  for (int k = 0; k < knt; k++) beta *= safmin;
  *alpha = beta;
  return 0;
}

/*
 * dlarf_ applied from the left (src/dlarf.c:129-185): C := C - tau v (C^T v)^T
 * with trailing zeros of v and trailing zero columns of C trimmed.
 */
void o_dlarf_left(int m, int n, const double *v, double tau, double *c, int ldc,
                  double *work) {
  if (tau == 0.0) return;
  int lastv = m;
  while (lastv > 0 && v[lastv - 1] == 0.0) --lastv;
  if (lastv <= 0) return;
  int lastc = o_iladlc(lastv, n, c, ldc);
  for (int jc = 0; jc < lastc; jc++) {
    const double *cj = c + (size_t)jc * ldc;
    double s = 0.0;
    for (int r = 0; r < lastv; r++) s += cj[r] * v[r];
    work[jc] = s;
  }
  for (int jc = 0; jc < lastc; jc++) {
    double *cj = c + (size_t)jc * ldc;
    double f = -tau * work[jc];
    for (int r = 0; r < lastv; r++) cj[r] += v[r] * f;
  }
}

/*
 * dgeqr2_mia (src/dgeqr2.c:29-195): unblocked QR of the m x n DM block with
 * in-panel truncation.  thres starts at 5e-14 and becomes tau_dm*|R(1,1)|
 * after the first column when tau_dm > 0 and n > 1.  *ncmp receives the
 * number of reflectors generated.
 */
void o_dgeqr2_dm(int m, int n, double *a, int lda, double *tau, double *work,
                 double tau_dm, int *ncmp) {
  int k = OMIN(m, n);
  double thres = 5e-14;
  int i;
  for (i = 1; i <= k; ++i) {
    int len = m - i + 1;
    double *aii = &a[(i - 1) + (size_t)(i - 1) * lda];
    double *x = &a[(OMIN(i + 1, m) - 1) + (size_t)(i - 1) * lda];
    if (o_dlarfg(len, aii, x, &tau[i - 1], thres, i)) {
      *ncmp = i - 1;
      return;
    }
    if (i < n) {
      double d = *aii;
      if (i == 1 && tau_dm > 0.0) thres = tau_dm * fabs(d);
      *aii = 1.0;
      o_dlarf_left(len, n - i, aii, tau[i - 1], &a[(i - 1) + (size_t)i * lda], lda, work);
      *aii = d;
    }
  }
  *ncmp = i - 1;
}

/*
 * dlarft, DIRECT='F', STOREV='C': upper-triangular T (k x k, ldt) with
 * H(1)...H(k) = I - V T V^T; V is m x k unit lower trapezoidal (the strict
 * upper part of the stored array is ignored).  The strict lower part of T is
 * set to zero.
 */
void o_dlarft_fc(int m, int k, const double *v, int ldv, const double *tau, double *t,
                 int ldt) {
  for (int i = 0; i < k; i++) {
    for (int r = 0; r < k; r++) t[r + (size_t)i * ldt] = 0.0;
    if (tau[i] == 0.0) continue;
    /* t(0:i, i) = -tau_i * V(i:m, 0:i)^T v_i, with v_i(i) = 1 */
    for (int p = 0; p < i; p++) {
      const double *vp = v + (size_t)p * ldv;
      const double *vi = v + (size_t)i * ldv;
      double s = vp[i]; /* row i: v_p(i) * 1 */
      for (int r = i + 1; r < m; r++) s += vp[r] * vi[r];
      t[p + (size_t)i * ldt] = -tau[i] * s;
    }
    /* t(0:i, i) = T(0:i, 0:i) * t(0:i, i)  (upper triangular, in place) */
    for (int p = 0; p < i; p++) {
      double s = 0.0;
      for (int q = p; q < i; q++) s += t[p + (size_t)q * ldt] * t[q + (size_t)i * ldt];
      t[p + (size_t)i * ldt] = s;
    }
    t[i + (size_t)i * ldt] = tau[i];
  }
}

/* LAPACKE_dge_nancheck semantics: 1 if any NaN in the m x n block. */
static int o_ge_nan(int m, int n, const double *a, int lda) {
  for (int c = 0; c < n; c++)
    for (int r = 0; r < m; r++)
      if (isnan(a[r + (size_t)c * lda])) return 1;
  return 0;
}

/*
 * LAPACKE_dlarfb_mia (src/dlarfb.c:40-151) for SIDE='L', TRANS='T',
 * DIRECT='F', STOREV='C': NaN checks on C (-13), T (-11), V (-9), then
 * C := H^T C = C - V (C^T V T)^T.
 */
int o_dlarfb_ltfc(int m, int n, int k, const double *v, int ldv, const double *t, int ldt,
                  double *c, int ldc) {
  if (o_ge_nan(m, n, c, ldc)) return -13;
  if (o_ge_nan(k, k, t, ldt)) return -11;
  for (int p = 0; p < k; p++) /* strictly lower part of the unit triangle */
    for (int r = p + 1; r < k && r < m; r++)
      if (isnan(v[r + (size_t)p * ldv])) return -9;
  if (m > k && o_ge_nan(m - k, k, v + k, ldv)) return -9;
  if (m <= 0 || n <= 0 || k <= 0) return 0;
  double *w = (double *)malloc(sizeof(double) * (size_t)n * k);
  /* W(c, i) = sum_r C(r, c) V(r, i), V unit lower trapezoidal */
  for (int jc = 0; jc < n; jc++) {
    const double *cj = c + (size_t)jc * ldc;
    for (int i = 0; i < k; i++) {
      const double *vi = v + (size_t)i * ldv;
      double s = cj[i];
      for (int r = i + 1; r < m; r++) s += cj[r] * vi[r];
      w[jc + (size_t)i * n] = s;
    }
  }
  /* W := W T (T upper triangular), column by column from the right */
  for (int jc = 0; jc < n; jc++) {
    for (int i = k - 1; i >= 0; i--) {
      double s = 0.0;
      for (int p = 0; p <= i; p++) s += w[jc + (size_t)p * n] * t[p + (size_t)i * ldt];
      w[jc + (size_t)i * n] = s;
    }
  }
  /* C := C - V W^T */
  for (int jc = 0; jc < n; jc++) {
    double *cj = c + (size_t)jc * ldc;
    for (int i = 0; i < k; i++) {
      const double *vi = v + (size_t)i * ldv;
      double f = w[jc + (size_t)i * n];
      cj[i] -= f;
      for (int r = i + 1; r < m; r++) cj[r] -= vi[r] * f;
    }
  }
  free(w);
  return 0;
}

/* ----------------------------------------------------------- DM pieces */

/*
 * norm_update (src/dgeqrdm_work.c:80-116): Drmac-Bujanovic safe downdate of
 * the partial norms after fjb rows were triangularised; recompute (and reset
 * workold) when temp2 <= tol3z.  a points at the first R12 row of the first
 * trailing column; rows fjb..fjb+m-1 lie below.  Columns with work == 0 are
 * skipped.  The square sum is the reference's naive sequential loop.
 */
void o_norm_update(int m, int n, int fjb, const double *a, int lda, double *work,
                   double *workold, double *maxnrm, double tol3z) {
  *maxnrm = 0.0;
  for (int jj = 0; jj < n; jj++) {
    if (work[jj] != 0.0) {
      const double *col = a + (size_t)jj * lda;
      double s = 0.0;
      for (int ii = 0; ii < fjb; ii++) s += col[ii] * col[ii];
      double temp = sqrt(fabs(s)) / work[jj];
      double d2 = (temp + 1.0) * (1.0 - temp);
      temp = (0.0 >= d2) ? 0.0 : d2;
      double d1 = work[jj] / workold[jj];
      double temp2 = temp * (d1 * d1);
      if (temp2 <= tol3z) {
        if (m > 0) {
          work[jj] = o_nrm2(m, col + fjb, 1);
          workold[jj] = work[jj];
        } else {
          work[jj] = 0.0;
          workold[jj] = 0.0;
        }
      } else {
        work[jj] *= sqrt(temp);
      }
      if (*maxnrm < work[jj]) *maxnrm = work[jj];
    }
  }
}

/* Descending by value; equal values keep ascending position.  This is the
 * order glibc's merge-sort qsort gives cmpStruct (src/dgeqrdm_work.c:126-142),
 * written here with an explicit tie-break so that any qsort reproduces it. */
static int o_key_cmp(const void *pa, const void *pb) {
  const o_key *a = (const o_key *)pa, *b = (const o_key *)pb;
  if (a->val > b->val) return -1;
  if (a->val < b->val) return 1;
  return (a->idx > b->idx) - (a->idx < b->idx);
}

/*
 * permute_marked (src/dgeqrdm_work.c:149-262) with nz = 0: moves the marked
 * (selected) columns into positions 0..nsel-1 by swapping whole columns
 * (m rows), jpvt and norms; selected columns that already sit in leading
 * positions stay.  The two-pointer walk (jb from the front, jt from the back)
 * and its placement order are the reference's.  Returns the swap count.
 */
int o_permute_marked(int m, int ncols, double *a, int lda, int *jpvt, double *norms,
                     int *marked, const int *idx_sel, int nsel, int nz) {
  int jb = 0, jt = ncols - 1, nsw = 0;
#define O_SWAP_COLS(p, q)                                                     \
  do {                                                                        \
    int ti_ = marked[p]; marked[p] = marked[q]; marked[q] = ti_;              \
    ti_ = jpvt[p]; jpvt[p] = jpvt[q]; jpvt[q] = ti_;                          \
    double td_ = norms[p]; norms[p] = norms[q]; norms[q] = td_;              \
    double *cp_ = a + (size_t)(p) * lda, *cq_ = a + (size_t)(q) * lda;        \
    for (int r_ = 0; r_ < m; r_++) { td_ = cp_[r_]; cp_[r_] = cq_[r_]; cq_[r_] = td_; } \
    nsw++;                                                                    \
  } while (0)
  for (int s = 0; s < nz + nsel; s++) {
    int jc = idx_sel[s];
    while (jb < jt && (marked[jt] == 1 || marked[jb] == 2)) {
      O_SWAP_COLS(jt, jb);
      while (jb < ncols && marked[jb] != 0 && marked[jb] != 2) jb++;
      while (jt >= 0 && marked[jt] != 0 && marked[jt] != 1) jt--;
    }
    if (marked[jc] == 1) {
      while (jb < ncols && marked[jb] != 0 && marked[jb] != 2) jb++;
      if (jc <= jb || jc < nsel) continue;
      if (marked[jb] == 0) {
        O_SWAP_COLS(jc, jb);
        jb++;
      }
    } else if (marked[jc] == 2) {
      while (jt >= 0 && marked[jt] != 0 && marked[jt] != 1) jt--;
      if (jc >= jt || jc >= ncols - nz) continue;
      if (marked[jt] == 0) {
        O_SWAP_COLS(jc, jt);
        jt--;
      }
    }
  }
#undef O_SWAP_COLS
  return nsw;
}

/*
 * DM_perm (src/dgeqrdm_work.c:280-418).  a0 = &A(0, j) (top of the first
 * trailing column), rows j..m-1 are the m_sub active rows (row offset j).
 * Sorts the partial norms, picks nc candidates (> tau * max, at most kmax),
 * scales them by 1/norm, forms the upper cosine Gram and accepts greedily
 * while max |cos| against earlier accepts < delta; then permute_marked.
 */
static void o_dm_perm(int mfull, int ncols, int j, double *norms, double *a0, int lda,
                      int *jpvt, double tau, double delta, int kmax, int *idx_sel,
                      int *nsel, int *mark, int *idxtmp, double *as, double *cosmat,
                      o_key *keys, oracle_step *tr) {
  int msub = mfull - j;
  for (int i = 0; i < ncols; i++) {
    keys[i].val = norms[i];
    keys[i].idx = i;
    mark[i] = 0;
  }
  qsort(keys, ncols, sizeof(o_key), o_key_cmp);
  int nc = 0;
  double lim = tau * keys[0].val;
  while (nc < ncols && nc < kmax && keys[nc].val > lim) nc++;
  *nsel = 1;
  idx_sel[0] = keys[0].idx;
  mark[keys[0].idx] = 1;
  idxtmp[0] = 0;
  if (nc > 1) {
    for (int c = 0; c < nc; c++) {
      double cc = 1.0 / keys[c].val;
      const double *src = a0 + (size_t)keys[c].idx * lda + j;
      for (int r = 0; r < msub; r++) as[(size_t)c * msub + r] = cc * src[r];
    }
    for (int q = 0; q < nc; q++)
      for (int p = 0; p <= q; p++) {
        const double *xp = as + (size_t)p * msub, *xq = as + (size_t)q * msub;
        double s = 0.0;
        for (int r = 0; r < msub; r++) s += xp[r] * xq[r];
        cosmat[p + (size_t)q * nc] = s;
      }
    for (int i = 1; i < nc; i++) {
      if (mark[keys[i].idx] != 0) continue;
      double mx = 0.0;
      for (int s = 0; s < *nsel; s++) {
        double v = fabs(cosmat[idxtmp[s] + (size_t)i * nc]);
        if (mx < v) mx = v;
      }
      if (mx < delta && *nsel < kmax) {
        idxtmp[*nsel] = i;
        idx_sel[*nsel] = keys[i].idx;
        mark[keys[i].idx] = 1;
        (*nsel)++;
      }
    }
  }
  if (tr) {
    tr->nc = nc;
    for (int s = 0; s < *nsel && s < 64; s++) tr->sel[s] = idx_sel[s];
  }
  int nsw = o_permute_marked(mfull, ncols, a0, lda, jpvt, norms, mark, idx_sel, *nsel, 0);
  if (tr) tr->nswap = nsw;
}

/* ------------------------------------------ fixed columns: dgeqrf + dormqr */

/* Unpivoted Householder QR of an m x n block (LAPACK dgeqr2 semantics). */
void o_dgeqr2(int m, int n, double *a, int lda, double *tau, double *work) {
  int k = OMIN(m, n);
  for (int i = 0; i < k; i++) {
    double *aii = &a[i + (size_t)i * lda];
    o_dlarfg(m - i, aii, &a[OMIN(i + 1, m - 1) + (size_t)i * lda], &tau[i], 0.0, 1);
    if (i + 1 < n) {
      double d = *aii;
      *aii = 1.0;
      o_dlarf_left(m - i, n - i - 1, aii, tau[i], &a[i + (size_t)(i + 1) * lda], lda, work);
      *aii = d;
    }
  }
}

/* dormqr SIDE='L': C := Q^T C (trans=1) or Q C (trans=0), Q = H(1)...H(k). */
void o_dormqr_left(int trans, int m, int n, int k, double *a, int lda, const double *tau,
                   double *c, int ldc, double *work) {
  for (int s = 0; s < k; s++) {
    int i = trans ? s : k - 1 - s;
    double *aii = &a[i + (size_t)i * lda];
    double d = *aii;
    *aii = 1.0;
    o_dlarf_left(m - i, n, aii, tau[i], c + i, ldc, work);
    *aii = d;
  }
}

/* ------------------------------------------------------------- the driver */

/*
 * dgeqrdm_work (src/dgeqrdm_work.c:420-838).  Column-major only; the DM code
 * of the reference indexes a[j*lda+i] whatever matrix_layout says, so only
 * 102 is coherent (101 passes validation there, and here).  trace (optional)
 * receives one record per step, up to ntrace records; *nsteps the count.
 */
/*
 * oracle_dgeqrdm_sched: the same loop where step it selects with
 * sched_tau[it], sched_delta[it] (it < nsched; thres afterwards) in place of
 * thres[1], thres[0]; the panel's truncation keeps thres[1].  This is how
 * the device path runs with a learned (tau, delta) proposal, which records
 * the parameters of every step.
 */
int oracle_dgeqrdm_sched(int matrix_layout, int m, int n, double *a, int lda, int *jpvt,
                         double *tau, int *ncols, const double *thres, int nb,
                         const double *sched_tau, const double *sched_delta, int nsched,
                         oracle_step *trace, int ntrace, int *nsteps);

int oracle_dgeqrdm_ex(int matrix_layout, int m, int n, double *a, int lda, int *jpvt,
                      double *tau, int *ncols, const double *thres, int nb,
                      oracle_step *trace, int ntrace, int *nsteps) {
  return oracle_dgeqrdm_sched(matrix_layout, m, n, a, lda, jpvt, tau, ncols, thres, nb, NULL,
                              NULL, 0, trace, ntrace, nsteps);
}

int oracle_dgeqrdm_sched(int matrix_layout, int m, int n, double *a, int lda, int *jpvt,
                         double *tau, int *ncols, const double *thres, int nb,
                         const double *sched_tau, const double *sched_delta, int nsched,
                         oracle_step *trace, int ntrace, int *nsteps) {
  double delta = thres[0], tau_dm = thres[1], eta = 0.0;
  int enable_stop = 0;
  double eps = O_EPS;
  double tol3z = sqrt(eps);
  if (nsteps) *nsteps = 0;
  if (ncols[0] == 1) {
    enable_stop = 1;
    eta = eps * n;
  } else if (ncols[0] == 2) {
    enable_stop = 2;
    eta = eps * sqrt((double)n);
  } else if (ncols[0] == 3) {
    enable_stop = 3;
    eta = thres[2];
  }
  int info = 0;
  if (matrix_layout != 102 && matrix_layout != 101) info = -1;
  else if (m <= 0) info = -2;
  else if (n <= 0) info = -3;
  else if (lda < OMAX(1, m)) info = -5;
  else if (delta < 0.0 || delta > 1.0) info = -9;
  else if (tau_dm < 0.0 || tau_dm > 1.0) info = -9;
  else if (nb <= 0) info = -10;
  if (info != 0) return -1;
  int minmn = OMIN(m, n);

  /* Move fixed columns up front (src/dgeqrdm_work.c:594-610). */
  int nfxd = 0;
  for (int jj = 0; jj < n; jj++) {
    if (jpvt[jj] != 0) {
      if (jj != nfxd) {
        double *p = a + (size_t)jj * lda, *q = a + (size_t)nfxd * lda;
        for (int r = 0; r < m; r++) {
          double t = p[r]; p[r] = q[r]; q[r] = t;
        }
        jpvt[jj] = jpvt[nfxd];
        jpvt[nfxd] = jj + 1;
      } else {
        jpvt[jj] = jj + 1;
      }
      nfxd++;
    } else {
      jpvt[jj] = jj + 1;
    }
  }
  double *w1 = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  if (nfxd > 0) {
    int na = OMIN(m, nfxd);
    o_dgeqr2(m, na, a, lda, tau, w1);
    if (na < n) o_dormqr_left(1, m, n - na, na, a, lda, tau, a + (size_t)na * lda, lda, w1);
  }
  if (nfxd < minmn) {
    int sm = m - nfxd;
    int ldt = nb;
    double *work = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *workold = work + n;
    int *mark = (int *)malloc(sizeof(int) * (size_t)n);
    int *idx_sel = (int *)malloc(sizeof(int) * (size_t)(n + nb));
    int *idxtmp = (int *)malloc(sizeof(int) * (size_t)(nb + 1));
    double *as = (double *)malloc(sizeof(double) * ((size_t)m * nb + 1));
    double *cosmat = (double *)malloc(sizeof(double) * ((size_t)nb * nb + 1));
    double *tmat = (double *)malloc(sizeof(double) * ((size_t)nb * nb + 1));
    o_key *keys = (o_key *)malloc(sizeof(o_key) * (size_t)n);
    double maxnrm = 0.0;
    for (int jj = nfxd; jj < n; jj++) {
      work[jj] = o_nrm2(sm, a + nfxd + (size_t)jj * lda, 1);
      if (work[jj] > maxnrm) maxnrm = work[jj];
      workold[jj] = work[jj];
    }
    eta = eta * maxnrm;
    int j = nfxd, it = -1;
    while (j < minmn) {
      int msub = m - j, nsub = n - j;
      it++;
      int kmax = OMIN(OMIN(nb, msub), nsub);
      int fjb = 0, fjb_cmp = 0;
      oracle_step *tr = (trace && it < ntrace) ? &trace[it] : NULL;
      if (tr) {
        memset(tr, 0, sizeof(*tr));
        tr->j = j;
        tr->kmax = kmax;
      }
      const double tau_s = (sched_tau && it < nsched) ? sched_tau[it] : tau_dm;
      const double delta_s = (sched_delta && it < nsched) ? sched_delta[it] : delta;
      o_dm_perm(m, nsub, j, &work[j], a + (size_t)j * lda, lda, &jpvt[j], tau_s, delta_s,
                kmax, idx_sel, &fjb, mark, idxtmp, as, cosmat, keys, tr);
      double *ajj = a + j + (size_t)j * lda;
      o_dgeqr2_dm(msub, fjb, ajj, lda, &tau[j], w1, tau_dm, &fjb_cmp);
      ncols[it] = fjb_cmp;
      o_dlarft_fc(msub, fjb_cmp, ajj, lda, &tau[j], tmat, ldt);
      info = o_dlarfb_ltfc(msub, nsub - fjb, fjb_cmp, ajj, lda, tmat, ldt,
                           a + j + (size_t)(j + fjb) * lda, lda);
      if (tr) {
        tr->fjb = fjb;
        tr->fjb_cmp = fjb_cmp;
      }
      if (nsteps) *nsteps = it + 1;
      if (info != 0) break;
      o_norm_update(msub - fjb_cmp, nsub - fjb_cmp, fjb_cmp,
                    a + j + (size_t)(j + fjb_cmp) * lda, lda, &work[j + fjb_cmp],
                    &workold[j + fjb_cmp], &maxnrm, tol3z);
      if (tr) tr->maxnrm = maxnrm;
      j += fjb_cmp;
      if (enable_stop >= 1 && maxnrm * sqrt((double)(nsub - fjb_cmp)) <= eta) break;
    }
    free(keys);
    free(tmat);
    free(cosmat);
    free(as);
    free(idxtmp);
    free(idx_sel);
    free(mark);
    free(work);
  }
  free(w1);
  return info;
}

/* Same signature as the reference's dgeqrdm (include/QRDM.h:19-22). */
int oracle_dgeqrdm(int matrix_layout, int m, int n, double *a, int lda, int *jpvt,
                   double *tau, int *ncols, double *thres, int nb) {
  return oracle_dgeqrdm_ex(matrix_layout, m, n, a, lda, jpvt, tau, ncols, thres, nb,
                           NULL, 0, NULL);
}

/* Explicit thin Q (m x k) from the reflectors: Q = H(1)...H(k) [I; 0]. */
void oracle_form_q(int m, int k, double *a, int lda, const double *tau, double *q, int ldq) {
  double *work = (double *)malloc(sizeof(double) * (size_t)(k + 1));
  for (int c = 0; c < k; c++)
    for (int r = 0; r < m; r++) q[r + (size_t)c * ldq] = (r == c) ? 1.0 : 0.0;
  o_dormqr_left(0, m, k, k, a, lda, tau, q, ldq, work);
  free(work);
}

/* Exposed pieces, for the unit tests of the restatement. */
double oracle_nrm2(int n, const double *x) { return o_nrm2(n, x, 1); }
int oracle_sizeof_step(void) { return (int)sizeof(oracle_step); }
