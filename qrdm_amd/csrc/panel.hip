// panel.hip -- the DM panel (dgeqr2_mia, src/dgeqr2.c:144-191) for tall
// panels, computed without a per-column grid-wide reduction.
//
// Householder QR of the m_sub x fjb panel P is determined by P's Gram matrix
// and its top fjb x fjb block, so it is computed as:
//   G   = P^T P: the DM step's candidate Gram (same columns, same rows j..m),
//         gathered in panel order -- no pass over P;
//   k_chol1  R1 = chol(G) and R1^-1 (one workgroup); the factorisation stops
//         at a column whose pivot proves the DM truncation test fires there;
//   pass 2   G2 = (P R1^-1)^T (P R1^-1): CholeskyQR2 re-orthogonalisation
//         (matrix cores, 64-row slabs);
//   k_recon  R = chol(G2) R1; Q11 = P_top R^-1; LU without pivoting of
//         Q11 D - I with the sign D chosen per column as dlarfg's
//         beta = -sign(alpha) ||.|| does (Householder reconstruction,
//         Ballard et al., "Reconstructing Householder vectors from
//         tall-skinny QR", 2014): L = V1, U = -T V1^T, diag(U) = -tau;
//         T from T V1^T = -U; M = R^-1 D U^-1.  The DM truncation test of
//         column i uses ||x_i|| = R_ii sqrt(1 - a_i^2), a_i = the LU's
//         unsigned pivot, against thres = tau_dm |R(0,0)| (or 5e-14), as
//         dlarfg_mia does (src/dlarfg.c:126-133, src/dgeqr2.c:176-177);
//   pass 3   V2 = P_bottom M (matrix cores), written into the panel and into
//         the dense V the trailing update reads.
// Columns the truncation rejects keep their original values; the caller's
// trailing update then starts at j + fjb_cmp and applies the kept reflectors
// to them, as dlarf_ did inside the reference's panel.
//
// Any degenerate case (a pivot that does not resolve, a reflector close to
// tau = 2 or tau = 0, non-finite values) sets DevState::flags bit 1 before
// anything is written, and the caller runs the exact panel instead.
//
// The one-workgroup kernels hold a 64 x 64 matrix as 4 x 4 register tiles
// (thread t: rows 4*(t/16).., columns 4*(t%16)..).  Each elimination step
// publishes the pivot row or column to LDS, every thread reads what it needs
// once and applies a masked, branch-free update: one barrier per step.
#include "qrdm_internal.h"

#include <math.h>

namespace qrdm {

constexpr int PR_ROWS = 64;  // rows per workgroup of passes 2 and 3
constexpr double kEpsP = 1.1102230246251565e-16;
constexpr int LDP = KMAX + 1;  // padded leading dimension of LDS matrices

typedef double dbl4p __attribute__((ext_vector_type(4)));

#define TILE_LOOP                \
  _Pragma("unroll") for (int a = 0; a < 4; a++) _Pragma("unroll") for (int b = 0; b < 4; b++)

struct Tile {
  int tid, r0, c0;
  __device__ Tile() : tid(threadIdx.x), r0(4 * (threadIdx.x >> 4)), c0(4 * (threadIdx.x & 15)) {}
};

// Publishing a row (column) writes the whole 4-row (4-column) group that
// holds it: grp[q * KMAX + x] = element (4*(row/4) + q, x), so the register
// tile is only indexed with compile-time indices (a runtime row index would
// push the tile to scratch memory).  Readers use grp + (row & 3) * KMAX.
__device__ __forceinline__ void pub_row(double* grp, const double (&w)[4][4], int row,
                                        const Tile& t) {
  if (t.r0 == (row & ~3)) {
    TILE_LOOP { grp[a * KMAX + t.c0 + b] = w[a][b]; }
  }
}

__device__ __forceinline__ void pub_col(double* grp, const double (&w)[4][4], int col,
                                        const Tile& t) {
  if (t.c0 == (col & ~3)) {
    TILE_LOOP { grp[b * KMAX + t.r0 + a] = w[a][b]; }
  }
}

__device__ __forceinline__ void scale_row(double (&w)[4][4], int row, double s, const Tile& t) {
  const bool mine = t.r0 == (row & ~3);
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const double f = (mine && (row & 3) == a) ? s : 1.0;
#pragma unroll
    for (int b = 0; b < 4; b++) w[a][b] *= f;
  }
}

__device__ __forceinline__ void tile_load(double (&w)[4][4], const double (*S)[LDP], const Tile& t) {
  TILE_LOOP { w[a][b] = S[t.r0 + a][t.c0 + b]; }
}

__device__ __forceinline__ void tile_store(double (*S)[LDP], const double (&w)[4][4], const Tile& t) {
  TILE_LOOP { S[t.r0 + a][t.c0 + b] = w[a][b]; }
}

__device__ __forceinline__ void tile_identity(double (&w)[4][4], int n, const Tile& t) {
  TILE_LOOP {
    const int ra = t.r0 + a, cb = t.c0 + b;
    w[a][b] = (ra == cb && ra < n) ? 1.0 : 0.0;
  }
}

// w = X Y for 64 x 64 LDS operands (zeros outside their triangles do the masking)
__device__ __forceinline__ void tile_gemm(double (&w)[4][4], const double (*X)[LDP],
                                          const double (*Y)[LDP], const Tile& t) {
  TILE_LOOP { w[a][b] = 0.0; }
  for (int q = 0; q < KMAX; q++) {
    double xv[4], yv[4];
#pragma unroll
    for (int a = 0; a < 4; a++) xv[a] = X[t.r0 + a][q];
#pragma unroll
    for (int b = 0; b < 4; b++) yv[b] = Y[q][t.c0 + b];
    TILE_LOOP { w[a][b] = fma(xv[a], yv[b], w[a][b]); }
  }
}

// In-place upper Cholesky of the n x n tiled matrix (upper part used).
// Leaves the unscaled rows in w and the pivots d_i in dpiv; row i of the
// factor is w(i, :) / sqrt(d_i).  stopfn(i, d) returns 0 to continue; the
// return value is the number of columns factored, *why the stop code.
template <class StopFn>
__device__ __forceinline__ int tile_chol(double (&w)[4][4], int n, double* rb, double* dpiv, StopFn stopfn,
                         int* why, const Tile& t) {
  pub_row(rb, w, 0, t);
  __syncthreads();
  for (int i = 0; i < n; i++) {
    const double* buf = rb + (i & 1) * 4 * KMAX + (i & 3) * KMAX;
    const double d = buf[i];
    const int st = stopfn(i, d);
    if (st) {
      *why = st;
      return i;
    }
    if (t.tid == 0) dpiv[i] = d;
    const double rd = 1.0 / d;
    double rv[4], cv[4];
#pragma unroll
    for (int a = 0; a < 4; a++) rv[a] = buf[t.r0 + a];
#pragma unroll
    for (int b = 0; b < 4; b++) cv[b] = buf[t.c0 + b];
    TILE_LOOP {
      const int ra = t.r0 + a, cb = t.c0 + b;
      const double f = (ra > i && cb >= ra) ? rv[a] * rd : 0.0;
      w[a][b] = fma(-f, cv[b], w[a][b]);
    }
    if (i + 1 < n) pub_row(rb + ((i + 1) & 1) * 4 * KMAX, w, i + 1, t);
    __syncthreads();
  }
  *why = 0;
  return n;
}

// w = R^-1 for the n x n upper triangular R in LDS (R zero outside 0..n-1):
// rows of R X = I from the bottom up, each final row scaled and published,
// then removed from the rows above it.
__device__ __forceinline__ void tile_triinv_upper(double (&w)[4][4], const double (*R)[LDP], int n, double* rb,
                                  const Tile& t) {
  tile_identity(w, n, t);
  if (n <= 0) return;
  scale_row(w, n - 1, 1.0 / R[n - 1][n - 1], t);
  pub_row(rb + ((n - 1) & 1) * 4 * KMAX, w, n - 1, t);
  __syncthreads();
  for (int i = n - 1; i >= 0; i--) {
    const double* buf = rb + (i & 1) * 4 * KMAX + (i & 3) * KMAX;
    double rv[4], xv[4];
#pragma unroll
    for (int a = 0; a < 4; a++) rv[a] = R[t.r0 + a][i];
#pragma unroll
    for (int b = 0; b < 4; b++) xv[b] = buf[t.c0 + b];
    TILE_LOOP {
      const int ra = t.r0 + a;
      const double f = (ra < i) ? rv[a] : 0.0;
      w[a][b] = fma(-f, xv[b], w[a][b]);
    }
    if (i >= 1) {
      scale_row(w, i - 1, 1.0 / R[i - 1][i - 1], t);
      pub_row(rb + ((i - 1) & 1) * 4 * KMAX, w, i - 1, t);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_chol1(const double* __restrict__ Gc,
                                               const int* __restrict__ rank, int fjb,
                                               double tau_dm, double* R1, double* R1inv,
                                               DevState* st) {
  __shared__ double Rs[KMAX][LDP];
  __shared__ double rb[8 * KMAX];
  __shared__ double dpiv[KMAX], diag[KMAX];
  __shared__ int srk[KMAX];
  const Tile t;
  if (t.tid < KMAX) {
    const int rk = t.tid < fjb ? rank[t.tid] : 0;
    srk[t.tid] = rk;
    diag[t.tid] = t.tid < fjb ? Gc[rk * (KMAX + 1)] : 0.0;
  }
  __syncthreads();
  double w[4][4];
  TILE_LOOP {
    const int ra = t.r0 + a, cb = t.c0 + b;
    w[a][b] = (ra < fjb && cb < fjb) ? Gc[srk[ra] + srk[cb] * KMAX] : 0.0;
  }
  double maxd = 0.0;
  for (int q = 0; q < KMAX; q++) maxd = fmax(maxd, diag[q]);
  const double g00 = diag[0];
  const double tolp = 1e4 * 64.0 * kEpsP * fjb * maxd;  // unresolved pivots
  const double thres = (tau_dm > 0.0 && fjb > 1) ? tau_dm * sqrt(g00) : 5e-14;
  const double thr2 = 0.25 * thres * thres;
  const bool scale_ok = g00 > 1e-280 && g00 < 1e280 && maxd < 1e280;
  int why = 0;
  const int k1 = tile_chol(
      w, fjb, rb, dpiv,
      [&](int i, double d) -> int {
        if (!scale_ok) return 2;
        if (i >= 1 && d < thr2 && thr2 > tolp) return 1;  // truncation at <= i is certain
        if (!(d > tolp)) return 2;
        return 0;
      },
      &why, t);
  if (why == 2) {
    if (t.tid == 0) st->flags |= 2;
    return;
  }
  __syncthreads();
  TILE_LOOP {
    const int ra = t.r0 + a, cb = t.c0 + b;
    const double v = (ra <= cb && cb < k1) ? w[a][b] / sqrt(dpiv[ra]) : 0.0;
    Rs[ra][cb] = v;
    R1[ra + cb * KMAX] = v;
  }
  __syncthreads();
  tile_triinv_upper(w, Rs, k1, rb, t);
  TILE_LOOP { R1inv[(t.r0 + a) + (t.c0 + b) * KMAX] = w[a][b]; }
  if (t.tid == 0) st->k1 = k1;
}

// ---------------------------------------------------------------------------
// 64 x 64 x 64 product on the matrix cores from LDS operands: wave w computes
// rows 16w..16w+15.  X is read as X[row][q] (or X[q][row] when xt), Y[q][col].
// acc[u][v] holds row 16w + lane/16 + 4v, column 16u + lane%16.
__device__ __forceinline__ void mfma_64(dbl4p (&acc)[4], const double (*X)[LDP], bool xt,
                                        const double (*Y)[LDP]) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int u = 0; u < 4; u++) acc[u] = (dbl4p){0.0, 0.0, 0.0, 0.0};
  const int ri = 16 * wv + (lane & 15), kq = lane >> 4;
#pragma unroll 4
  for (int k0 = 0; k0 < KMAX; k0 += 4) {
    const double av = xt ? X[k0 + kq][ri] : X[ri][k0 + kq];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const double bv = Y[k0 + kq][16 * u + (lane & 15)];
      acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[u], 0, 0, 0);
    }
  }
}

__device__ __forceinline__ void mfma_store(double (*S)[LDP], const dbl4p (&acc)[4]) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++) S[16 * wv + (lane >> 4) + 4 * v][16 * u + (lane & 15)] = acc[u][v];
}

// Pass 2: Q = P R1^-1 on a 64-row slab, partial Gram Q^T Q.
__global__ __launch_bounds__(256) void k_qgram(const double* __restrict__ A, int lda, int j,
                                               int m, const double* __restrict__ R1inv,
                                               const DevState* st, double* partial) {
  __shared__ double Ps[PR_ROWS][LDP];  // P slab (row, col), then Q
  __shared__ double Ri[KMAX][LDP];
  if (st->flags & 2) return;
  const int k1 = st->k1;
  const int tid = threadIdx.x;
  const int rs = j + blockIdx.x * PR_ROWS;
  for (int e = tid; e < KMAX * PR_ROWS; e += blockDim.x) {
    const int c = e / PR_ROWS, r = e % PR_ROWS;
    Ps[r][c] = (c < k1 && rs + r < m) ? A[(size_t)(j + c) * lda + rs + r] : 0.0;
  }
  for (int e = tid; e < KMAX * KMAX; e += blockDim.x) Ri[e & 63][e >> 6] = R1inv[e];
  __syncthreads();
  dbl4p acc[4];
  mfma_64(acc, Ps, false, Ri);
  __syncthreads();
  mfma_store(Ps, acc);
  __syncthreads();
  mfma_64(acc, Ps, true, Ps);  // Q^T Q
  const int lane = tid & 63, wv = tid >> 6;
  double* out = partial + (size_t)blockIdx.x * KMAX * KMAX;
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++)
      out[(16 * wv + (lane >> 4) + 4 * v) + (16 * u + (lane & 15)) * KMAX] = acc[u][v];
}

// ---------------------------------------------------------------------------
// Reconstruction: see the file header.  Writes the top k x k block of the
// panel (R above and on the diagonal, V1 below), tau, T and M; or only the
// fallback flag.  st->fjb_cmp = k.
__global__ __launch_bounds__(256) void k_recon(const double* __restrict__ G2,
                                               const double* __restrict__ R1, double* A,
                                               int lda, int j, int fjb, double tau_dm,
                                               double* tau, double* T_out, double* M_out,
                                               DevState* st) {
  __shared__ double SA[KMAX][LDP];  // R2, then R^-1
  __shared__ double SB[KMAX][LDP];  // R1, then R
  __shared__ double SC[KMAX][LDP];  // P_top, then L
  __shared__ double SD[KMAX][LDP];  // U, then U^-1
  __shared__ double rb[8 * KMAX], cl[8 * KMAX];
  __shared__ double dsg[KMAX], tl[KMAX], dpiv[KMAX];
  if (st->flags & 2) return;
  const Tile t;
  const int k1 = st->k1;
  double w[4][4];
  // 1. R2 = chol(G2); G2 is close to I after the first pass
  TILE_LOOP {
    const int ra = t.r0 + a, cb = t.c0 + b;
    w[a][b] = (ra < k1 && cb < k1) ? G2[ra + cb * KMAX] : 0.0;
  }
  int why = 0;
  tile_chol(
      w, k1, rb, dpiv, [&](int, double d) -> int { return (d > 0.25 && d < 4.0) ? 0 : 2; },
      &why, t);
  if (why) {
    if (t.tid == 0) st->flags |= 2;
    return;
  }
  __syncthreads();
  TILE_LOOP {
    const int ra = t.r0 + a, cb = t.c0 + b;
    SA[ra][cb] = (ra <= cb && cb < k1) ? w[a][b] / sqrt(dpiv[ra]) : 0.0;
    SB[ra][cb] = R1[ra + cb * KMAX];
    SC[ra][cb] = (ra < k1 && cb < k1) ? A[(size_t)(j + cb) * lda + j + ra] : 0.0;
  }
  __syncthreads();
  // 2. R = R2 R1
  tile_gemm(w, SA, SB, t);
  __syncthreads();
  tile_store(SB, w, t);
  __syncthreads();
  // 3. R^-1 -> SA
  tile_triinv_upper(w, SB, k1, rb, t);
  tile_store(SA, w, t);
  __syncthreads();
  // 4. Q11 = P_top R^-1
  tile_gemm(w, SC, SA, t);
  // 5. LU of Q11 D - I (unsigned running matrix), dlarfg's signs, truncation
  const double thres = (tau_dm > 0.0 && fjb > 1) ? tau_dm * SB[0][0] : 5e-14;
  pub_row(rb, w, 0, t);
  pub_col(cl, w, 0, t);
  __syncthreads();
  int k = k1;
  for (int i = 0; i < k1; i++) {
    const double* rowb = rb + (i & 1) * 4 * KMAX + (i & 3) * KMAX;
    const double* colb = cl + (i & 1) * 4 * KMAX + (i & 3) * KMAX;
    const double al = rowb[i];
    const double aa = fabs(al);
    const double om = 1.0 - aa;  // 2 - tau
    const double xnorm = SB[i][i] * sqrt(fmax(om, 0.0) * (1.0 + aa));
    int stop = 0;
    if (!(al == al) || !(xnorm == xnorm)) stop = 2;
    else if (i >= 1 && xnorm < thres) stop = 1;
    else if (om < 1e-6) stop = 2;  // tau near 2: dlarfg's tau = 0 branch or lost digits
    if (stop == 1) {
      k = i;
      break;
    }
    if (stop == 2) {
      if (t.tid == 0) st->flags |= 2;
      return;
    }
    const double di = (al >= 0.0) ? -1.0 : 1.0;  // beta = -d_sign(norm, alpha)
    const double sc = di / (-1.0 - aa);           // L column = d_i * column / pivot
    if (t.tid == 0) {
      dsg[i] = di;
      tl[i] = 1.0 + aa;
    }
    double rv[4], cv[4];
#pragma unroll
    for (int a = 0; a < 4; a++) rv[a] = colb[t.r0 + a] * sc;
#pragma unroll
    for (int b = 0; b < 4; b++) cv[b] = rowb[t.c0 + b];
    TILE_LOOP {
      const int ra = t.r0 + a, cb = t.c0 + b;
      const double upd = fma(-rv[a], cv[b], w[a][b]);
      w[a][b] = (ra > i) ? ((cb == i) ? rv[a] : ((cb > i) ? upd : w[a][b])) : w[a][b];
    }
    if (i + 1 < k1) {
      pub_row(rb + ((i + 1) & 1) * 4 * KMAX, w, i + 1, t);
      pub_col(cl + ((i + 1) & 1) * 4 * KMAX, w, i + 1, t);
    }
    __syncthreads();
  }
  __syncthreads();
  // 6. top block (R rows scaled by the signs, V1 below), tau; L -> SC; U -> SD
  double u[4][4];
  TILE_LOOP {
    const int ra = t.r0 + a, cb = t.c0 + b;
    const bool in = ra < k && cb < k;
    if (in) A[(size_t)(j + cb) * lda + j + ra] = (ra <= cb) ? dsg[ra] * SB[ra][cb] : w[a][b];
    SC[ra][cb] = (in && ra > cb) ? w[a][b] : 0.0;
    u[a][b] = in ? (ra < cb ? dsg[cb] * w[a][b] : (ra == cb ? -tl[cb] : 0.0)) : 0.0;
    SD[ra][cb] = u[a][b];
  }
  if (t.tid < k) tau[t.tid] = tl[t.tid];
  // 7. T V1^T = -U, right-looking over columns (V1 unit lower)
  TILE_LOOP { w[a][b] = -u[a][b]; }
  pub_col(cl, w, 0, t);
  __syncthreads();
  for (int c = 0; c < k; c++) {
    const double* buf = cl + (c & 1) * 4 * KMAX + (c & 3) * KMAX;
    double xv[4], lv[4];
#pragma unroll
    for (int a = 0; a < 4; a++) xv[a] = buf[t.r0 + a];
#pragma unroll
    for (int b = 0; b < 4; b++) lv[b] = (t.c0 + b > c) ? SC[t.c0 + b][c] : 0.0;
    TILE_LOOP { w[a][b] = fma(-xv[a], lv[b], w[a][b]); }
    if (c + 1 < k) pub_col(cl + ((c + 1) & 1) * 4 * KMAX, w, c + 1, t);
    __syncthreads();
  }
  TILE_LOOP {
    const int ra = t.r0 + a, cb = t.c0 + b;
    T_out[ra + cb * KMAX] = (ra < k && cb < k) ? w[a][b] : 0.0;
  }
  // 8. U^-1 -> SD
  tile_triinv_upper(w, SD, k, rb, t);
  __syncthreads();
  tile_store(SD, w, t);
  // 9. M = R^-1 D U^-1 (D folded into the rows of U^-1)
  if (t.tid < KMAX) dpiv[t.tid] = t.tid < k ? dsg[t.tid] : 0.0;
  __syncthreads();
  TILE_LOOP { SD[t.r0 + a][t.c0 + b] *= dpiv[t.r0 + a]; }
  __syncthreads();
  tile_gemm(w, SA, SD, t);
  TILE_LOOP {
    const int ra = t.r0 + a, cb = t.c0 + b;
    M_out[ra + cb * KMAX] = (ra < k && cb < k) ? w[a][b] : 0.0;
  }
  if (t.tid == 0) st->fjb_cmp = k;
}

// ---------------------------------------------------------------------------
// Pass 3: V2 = P_bottom M on a 64-row slab; writes V2 into the panel below the
// top block and the dense zero-padded V (rows [j, m), KMAX columns) with V1's
// unit lower block on top.
__global__ __launch_bounds__(256) void k_v2(double* A, int lda, int j, int m,
                                            const double* __restrict__ M, const DevState* st,
                                            double* V, int ldv) {
  __shared__ double Ps[PR_ROWS][LDP];
  __shared__ double Ms[KMAX][LDP];
  if (st->flags & 2) return;
  const int k = st->fjb_cmp;
  const int tid = threadIdx.x;
  const int rs = j + blockIdx.x * PR_ROWS;
  for (int e = tid; e < KMAX * PR_ROWS; e += blockDim.x) {
    const int c = e / PR_ROWS, r = e % PR_ROWS;
    Ps[r][c] = (c < k && rs + r < m) ? A[(size_t)(j + c) * lda + rs + r] : 0.0;
  }
  for (int e = tid; e < KMAX * KMAX; e += blockDim.x) Ms[e & 63][e >> 6] = M[e];
  __syncthreads();
  dbl4p acc[4];
  mfma_64(acc, Ps, false, Ms);
  __syncthreads();
  // rows of the top block keep V1 (unit lower, written by k_recon): overlay
  {
    const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = 16 * wv + (lane >> 4) + 4 * v, c = 16 * u + (lane & 15);
        const int pr = rs - j + r;
        double x = acc[u][v];
        if (c >= k) x = 0.0;
        else if (pr < k) x = (pr < c) ? 0.0 : (pr == c ? 1.0 : Ps[r][c]);
        acc[u][v] = x;
      }
  }
  __syncthreads();
  mfma_store(Ps, acc);
  __syncthreads();
  for (int e = tid; e < KMAX * PR_ROWS; e += blockDim.x) {
    const int c = e / PR_ROWS, r = e % PR_ROWS;
    const int row = rs + r;
    if (row < m) {
      const double x = Ps[r][c];
      V[(size_t)c * ldv + row] = x;
      if (c < k && row - j >= k) A[(size_t)(j + c) * lda + row] = x;
    }
  }
}

void launch_panel_fast(double* A, int lda, int j, int m, int fjb, double tau_dm, double* tau,
                       const double* Gc, const int* rank, double* gram_part, double* R1,
                       double* R1inv, double* G2, double* T, double* M, double* V, int ldv,
                       DevState* st, hipStream_t s) {
  hipLaunchKernelGGL(k_chol1, dim3(1), dim3(256), 0, s, Gc, rank, fjb, tau_dm, R1, R1inv, st);
  const int nslab = (m - j + PR_ROWS - 1) / PR_ROWS;
  hipLaunchKernelGGL(k_qgram, dim3(nslab), dim3(256), 0, s, A, lda, j, m, R1inv, st, gram_part);
  launch_gram_reduce(gram_part, nslab, G2, s);
  hipLaunchKernelGGL(k_recon, dim3(1), dim3(256), 0, s, G2, R1, A, lda, j, fjb, tau_dm, tau, T,
                     M, st);
  hipLaunchKernelGGL(k_v2, dim3(nslab), dim3(256), 0, s, A, lda, j, m, M, st, V, ldv);
}

}  // namespace qrdm
