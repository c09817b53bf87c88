// panel.hip -- the DM panel (dgeqr2_mia, src/dgeqr2.c:144-191) for tall
// panels, computed without a per-column grid-wide reduction.
//
// Householder QR of the m_sub x fjb panel P is determined by P's Gram matrix
// and its top fjb x fjb block, so it is computed as:
//   G   = P^T P: the DM step's candidate Gram (same columns, same rows j..m),
//         gathered in panel order -- no pass over P;
//   chol1_body  R1 = chol(G) and R1^-1 (one workgroup, in k_perm_chol1
//         beside the column moves); the factorisation stops at a column
//         whose pivot proves the DM truncation test fires there;
//   k_qgram  G2 = (P R1^-1)^T (P R1^-1): CholeskyQR2 re-orthogonalisation
//         (matrix cores, 64-row slabs);
//   recon_body  (the last workgroup of G2's reduce, k_reduce_recon)
//         R = R2 R1 with R2^T R2 = G2; Q11 = P_top R^-1; LU without
//         pivoting of Q11 D - I with the sign D chosen per column as
//         dlarfg's beta = -sign(alpha) ||.|| does (Householder
//         reconstruction, Ballard et al., "Reconstructing Householder
//         vectors from tall-skinny QR", 2014): L = V1, U = -T V1^T,
//         diag(U) = -tau; T from T V1^T = -U; M = R^-1 D U^-1.  The DM
//         truncation test of column i uses ||x_i|| = R_ii sqrt(1 - a_i^2),
//         a_i = the LU's unsigned pivot, against thres = tau_dm |R(0,0)|
//         (or 5e-14), as dlarfg_mia does (src/dlarfg.c:126-133,
//         src/dgeqr2.c:176-177);
//   k_v2     V2 = P_bottom M (matrix cores), written into the panel and into
//         the dense V the trailing update reads, with V^T V per slab;
//   k_reduce_t  tau and T from V^T V (the compact WY identity), so that the
//         block reflector is orthogonal for the V stored.
// Columns the truncation rejects keep their original values; the caller's
// trailing update then starts at j + fjb_cmp and applies the kept reflectors
// to them, as dlarf_ did inside the reference's panel.
//
// Any degenerate case (a pivot that does not resolve, a reflector close to
// tau = 2 or tau = 0, non-finite values) sets DevState::flags bit 1 and the
// halt before anything is written, and the step continues with the exact
// panel.
//
// The one-workgroup stages keep 64 x 64 matrices in LDS and work in
// 16-column blocks: wave 0 factors a diagonal block with one lane per row,
// taking the pivot and the pivot row from lane i by a DPP row broadcast
// (register to register, all in vector registers; no barrier
// inside the block; each column's reciprocal or inverse square root is a
// hardware estimate refined by Newton steps, and the column's results are
// computed before its stop test, which only decides what is kept); all
// threads then solve the block row or column (one lane per row or column,
// independent), and the matrix cores update the trailing matrix.  Three
// barriers per block instead of one per column.
#include "qrdm_internal.h"

#include <math.h>

namespace qrdm {

constexpr int PR_ROWS = 64;  // rows per workgroup of passes 2 and 3
constexpr double kEpsP = 1.1102230246251565e-16;
constexpr int LDP = KMAX + 1;  // padded leading dimension of LDS matrices
constexpr int BLK = 16;        // block width of the one-workgroup factorisations
// Threads of every workgroup that runs this file's one-workgroup stages (the
// launches here and the batched kernel's four waves): element loops over a
// 64 x 64 matrix stride by it, so that they unroll and their loads overlap.
constexpr int NT1 = 256;

typedef double dbl4p __attribute__((ext_vector_type(4)));

// 1/sqrt(d) for a positive normal d: the hardware estimate and two Newton
// steps (about one ulp; short dependency chain for the factorisations' pivots).
__device__ __forceinline__ double rsqrt_nr(double d) {
  const double h = 0.5 * d;
  double y = __builtin_amdgcn_rsq(d);
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  return y;
}

// Workgroup barrier that orders LDS only: this file's one-workgroup stages
// share data through LDS, never through global memory, so their global
// stores may still be in flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// 1/x for x in [1, 2]: the hardware estimate and two Newton steps.
__device__ __forceinline__ double rcp_nr12(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.0), y);
  y = fma(y, fma(-x, y, 1.0), y);
  return y;
}

// S[r][c] -= sum_{t < 16} A(r, t) B(t, c) for r, c in [lo, nlim), 16 x 16
// tiles on the matrix cores.  sym: A(r, t) = S[bs + t][r] and only tiles on
// or above the diagonal (the symmetric Cholesky update); otherwise A(r, t) =
// S[r][bs + t].  B(t, c) = S[bs + t][c].  The tiles written are disjoint
// from the rows bs..bs+15 and columns bs..bs+15 read.  The blocked
// factorisations look one block ahead: the calling wave takes only the next
// diagonal tile (lo, lo) with diag_tile, and waves 1-3 share the other
// tiles (wave w takes the tiles w - 1, w + 2, ... of them) without it.
__device__ __forceinline__ void mfma_trailing(double (*S)[LDP], int bs, int lo, int nlim,
                                              bool sym, bool diag_tile) {
  const int lane = local_tid() & 63, wave = local_tid() >> 6;
  const int g = lane >> 4, li = lane & 15;
  const int nt = (KMAX - lo) >> 4;
  int e = 0;
  for (int tr = 0; tr < nt; tr++) {
    for (int tc = sym ? tr : 0; tc < nt; tc++) {
      const int r0 = lo + 16 * tr, c0 = lo + 16 * tc;
      const bool first = tr == 0 && tc == 0;
      bool mine;
      if (diag_tile) {
        mine = first;
      } else {
        mine = !first && (e % 3) == wave - 1;
        if (!first) e++;
      }
      if (!mine || r0 >= nlim || c0 >= nlim) continue;
      dbl4p acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k0 = 0; k0 < BLK; k0 += 4) {
        const double a = sym ? S[bs + k0 + g][r0 + li] : S[r0 + li][bs + k0 + g];
        const double b = S[bs + k0 + g][c0 + li];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
      }
#pragma unroll
      for (int v = 0; v < 4; v++) S[r0 + g + 4 * v][c0 + li] -= acc[v];
    }
  }
}

// In-place upper Cholesky (R^T R = S, upper part of S read) of the leading
// nlim x nlim block; 256 threads.  stopfn(i, d) is called with each pivot d
// before it is used and returns 0 to continue, else a stop code: the
// factorisation then ends before column i.  Returns the number of columns
// factored (rows < that of S then hold R) and the stop code in *why.
// rinv[i] = 1 / R(i, i).
template <class StopFn>
__device__ __forceinline__ int chol_blocked(double (*S)[LDP], int nlim, StopFn stopfn, double* rinv,
                                            int* s_done, int* s_why, int* why) {
  const int tid = local_tid(), wave = tid >> 6;
  // Wave 0 factors the diagonal block at bs: lane r (= lane % 16) holds row
  // bs + r of the symmetric block; both triangles are updated, so R(i, r) =
  // S(r, i) rs is the lane's own value and row i comes from lane i by a row
  // broadcast.  Branch-free pivot steps: a pivot that is not taken updates
  // nothing (zero multipliers; entries past the stop are not read), and lane
  // r stores 1 / R(r, r) after the block.  (The row index is re-read per
  // block: one fixed across the blocks lets the compiler hoist all 32 row
  // masks out of the loop and spill them.)
  auto diag = [&](int bs) {
    const int lane = local_tid() & 63, r = lane & 15;
    double w[BLK];
#pragma unroll
    for (int c = 0; c < BLK; c++) w[c] = S[bs + r][bs + c];
    double rsr = 0.0;  // 1 / sqrt(pivot) of this lane's row once eliminated
    bool act = true;
    int done = BLK, code = 0;
#pragma unroll
    for (int i = 0; i < BLK; i++) {
      const int gi = bs + i;
      const double d = row_bcast(w[i], i);
      const int stc = (gi < nlim) ? stopfn(gi, d) : 0;
      const double rs = rsqrt_nr(d);  // committed only when the pivot is accepted
      const bool go = act && gi < nlim && stc == 0;
      done = (act && !go) ? i : done;
      code = (act && !go) ? stc : code;
      act = go;
      rsr = (r == i) ? rs : rsr;
      // S(r, c) -= R(i, r) R(i, c) = S(r, i) rs^2 S(i, c) for rows r > i
      const double a = (go && r > i) ? w[i] * rs * rs : 0.0;
#pragma unroll
      for (int c = i + 1; c < BLK; c++) w[c] = fma(-a, row_bcast(w[c], i), w[c]);
    }
    // rows < done: R(r, c) = S(r, c) / sqrt(pivot) on and above the diagonal
    if (lane < BLK) {
#pragma unroll
      for (int c = 0; c < BLK; c++)
        S[bs + r][bs + c] = (r < done) ? ((c < r) ? 0.0 : w[c] * rsr) : w[c];
      if (r < done) rinv[bs + r] = rsr;
    }
    if (lane == 0) {
      *s_done = done;
      *s_why = code;
    }
  };
  for (int bs = 0; bs < nlim; bs += BLK) {
    // look-ahead: the previous block's trailing update, wave 0 on this
    // diagonal tile (then its factorisation), waves 1-3 on the other tiles
    if (wave == 0) {
      if (bs > 0) {
        mfma_trailing(S, bs - BLK, bs, nlim, true, true);
        wave_lds_sync();
      }
      diag(bs);
    } else if (bs > 0) {
      mfma_trailing(S, bs - BLK, bs, nlim, true, false);
    }
    lds_barrier();
    const int done = *s_done, code = *s_why;
    if (done < BLK || code) {
      *why = code;
      return bs + done;
    }
    const int lo = bs + BLK;
    if (lo >= nlim) break;
    // block row: R(bs.., c) = R_bb^-T S(bs.., c), one thread per column
    for (int c = lo + tid; c < nlim; c += NT1) {
      double x[BLK];
#pragma unroll
      for (int i = 0; i < BLK; i++) x[i] = S[bs + i][c];
#pragma unroll
      for (int i = 0; i < BLK; i++) {
        double sacc = x[i];
#pragma unroll
        for (int l = 0; l < i; l++) sacc = fma(-S[bs + l][bs + i], x[l], sacc);
        x[i] = sacc * rinv[bs + i];
      }
#pragma unroll
      for (int i = 0; i < BLK; i++) S[bs + i][c] = x[i];
    }
    lds_barrier();
  }
  *why = 0;
  return nlim;
}

// In-place LU without pivoting of the Householder reconstruction's running
// matrix W (the unsigned Q11, see k_recon) on the leading nlim x nlim block;
// 256 threads.  Column i is eliminated with the pivot d_i a_i - 1 of
// Q11 D - I, d_i = -sign(a_i) (dlarfg's beta = -sign(alpha) ||.||), a_i the
// running diagonal: multipliers L(r, i) = d_i W(r, i) / (-1 - |a_i|); the
// identity part never enters the updates, so W keeps the unsigned U rows.
// Before column i the DM truncation test runs on ||x_i|| = R(i,i)
// sqrt(1 - a_i^2) (stop code 1 for i >= 1 below thres) and degenerate
// pivots stop with code 2.  Returns the number of columns eliminated;
// dsg[i] = d_i, tl[i] = 1 + |a_i| (= tau_i), scv[i] = the multiplier scale.
__device__ __forceinline__ int lu_signed_blocked(double (*W)[LDP], int nlim, const double (*SR)[LDP],
                                                 double thres, double* dsg, double* tl,
                                                 double* scv, int* s_done, int* s_why,
                                                 int* why) {
  const int tid = local_tid(), wave = tid >> 6;
  // Wave 0 factors the diagonal block at bs: lane r (= lane % 16) holds row
  // bs + r; the pivot row comes from lane i by a row broadcast.  Branch-free
  // pivot steps: the stop code is a select, and a pivot that is not taken
  // updates nothing (its multipliers are zero; entries past the stop are not
  // read).  Lane r keeps d, tau and the scale of its own column and stores
  // them after the block.  (The row index is re-read per block: one fixed
  // across the blocks lets the compiler hoist all 32 row masks out of the
  // loop and spill them.)
  auto diag = [&](int bs) {
    const int lane = local_tid() & 63, r = lane & 15;
    double w[BLK];
#pragma unroll
    for (int c = 0; c < BLK; c++) w[c] = W[bs + r][bs + c];
    // ||x_i|| < thres  <=>  (1 - a^2) < (thres / R(i, i))^2: lane r holds the
    // bound of column bs + r (R's diagonal is positive)
    double rt2;
    {
      const double rt = thres / SR[bs + r][bs + r];
      rt2 = rt * rt;
    }
    bool act = true;
    int done = BLK, code = 0;
    double d_r = 0.0, tl_r = 0.0, sc_r = 0.0;
#pragma unroll
    for (int i = 0; i < BLK; i++) {
      const int gi = bs + i;
      const double al = row_bcast(w[i], i);
      const double aa = fabs(al);
      const double om = 1.0 - aa;  // 2 - tau
      const double x2 = fmax(om, 0.0) * (1.0 + aa);  // (||x_i|| / R(i, i))^2
      const double di = (al >= 0.0) ? -1.0 : 1.0;
      const double sc = -di * rcp_nr12(1.0 + aa);  // d_i / (-1 - |a_i|)
      const double rt2i = row_bcast(rt2, i);
      // non-finite pivot, the truncation test, tau near 2 (dlarfg's tau = 0
      // branch or lost digits)
      const int stc = (gi >= nlim) ? 0
                                   : (!(al == al) ? 2 : ((gi >= 1 && x2 < rt2i) ? 1
                                                                               : ((om < 1e-6) ? 2 : 0)));
      const bool go = act && gi < nlim && stc == 0;
      done = (act && !go) ? i : done;
      code = (act && !go) ? stc : code;
      act = go;
      d_r = (r == i) ? di : d_r;
      tl_r = (r == i) ? 1.0 + aa : tl_r;
      sc_r = (r == i) ? sc : sc_r;
      const double lv = (go && r > i) ? w[i] * sc : 0.0;  // L(r, i)
#pragma unroll
      for (int c = i + 1; c < BLK; c++) w[c] = fma(-lv, row_bcast(w[c], i), w[c]);
      w[i] = (go && r > i) ? lv : w[i];
    }
    if (lane < BLK) {
#pragma unroll
      for (int c = 0; c < BLK; c++) W[bs + r][bs + c] = w[c];
      if (r < done) {
        dsg[bs + r] = d_r;
        tl[bs + r] = tl_r;
        scv[bs + r] = sc_r;
      }
    }
    if (lane == 0) {
      *s_done = done;
      *s_why = code;
    }
  };
  for (int bs = 0; bs < nlim; bs += BLK) {
    // look-ahead: the previous block's trailing update, wave 0 on this
    // diagonal tile (then its factorisation), waves 1-3 on the other tiles
    if (wave == 0) {
      if (bs > 0) {
        mfma_trailing(W, bs - BLK, bs, nlim, false, true);
        wave_lds_sync();
      }
      diag(bs);
    } else if (bs > 0) {
      mfma_trailing(W, bs - BLK, bs, nlim, false, false);
    }
    lds_barrier();
    const int done = *s_done, code = *s_why;
    if (done < BLK || code) {
      *why = code;
      return bs + done;
    }
    const int lo = bs + BLK;
    if (lo >= nlim) break;
    if (tid < 64) {
      // block column below: L(r, bs..) by forward substitution with U_bb
      const int r = lo + tid;
      if (r < nlim) {
        double x[BLK];
#pragma unroll
        for (int i = 0; i < BLK; i++) x[i] = W[r][bs + i];
#pragma unroll
        for (int i = 0; i < BLK; i++) {
          double sacc = x[i];
#pragma unroll
          for (int l = 0; l < i; l++) sacc = fma(-x[l], W[bs + l][bs + i], sacc);
          x[i] = sacc * scv[bs + i];
        }
#pragma unroll
        for (int i = 0; i < BLK; i++) W[r][bs + i] = x[i];
      }
    } else if (tid < 128) {
      // block row to the right: U(bs.., c) = L_bb^-1 W(bs.., c)
      const int c = lo + tid - 64;
      if (c < nlim) {
        double y[BLK];
#pragma unroll
        for (int i = 0; i < BLK; i++) y[i] = W[bs + i][c];
#pragma unroll
        for (int i = 0; i < BLK; i++) {
          double sacc = y[i];
#pragma unroll
          for (int l = 0; l < i; l++) sacc = fma(-W[bs + i][bs + l], y[l], sacc);
          y[i] = sacc;
        }
#pragma unroll
        for (int i = 0; i < BLK; i++) W[bs + i][c] = y[i];
      }
    }
    lds_barrier();
  }
  *why = 0;
  return nlim;
}

// X := L^-1 and U^-1 of the pair held in one 64 x 64 matrix W (L unit lower
// triangular strictly below the diagonal, U upper triangular on and above
// it, non-zero diagonal), X in the same layout: L^-1 strictly below the
// diagonal (its unit diagonal implied), U^-1 on and above; 256 threads, ends
// with a barrier.  BOTH = false: U^-1 only (W's lower part is not read, X's
// is set to zero).  Both go by 16 x 16 blocks, L^-1 as (L^T)^-1 read and
// stored transposed: the diagonal blocks by back substitution (wave w takes
// block w of both, one lane per column: lanes 0-15 U, 16-31 L^T), then the
// blocks at block distance d = 1, 2, 3 on the matrix cores, one block per
// wave: X_ij = -X_ii Y with Y = sum_{k = i+1..j} R_ik X_kj (R = U or L^T).
// Ys: one 16 x 16 scratch tile per wave.
template <bool BOTH = true>
__device__ __forceinline__ void lu_inverse(const double (*W)[LDP], double (*X)[LDP], double* rinv,
                                           double (*Ys)[BLK][BLK + 1]) {
  const int tid = local_tid(), lane = tid & 63, wave = tid >> 6;
  // R_m(i, l): m = 1 U, m = 0 L^T (unit diagonal); zero below the diagonal
  auto Rm = [&](int m, int i, int l) -> double {
    return m ? ((i <= l) ? W[i][l] : 0.0) : ((i < l) ? W[l][i] : (i == l ? 1.0 : 0.0));
  };
  // X_m(i, j) of the inverse of R_m, from X's storage
  auto Xm = [&](int m, int i, int j) -> double {
    return m ? ((i <= j) ? X[i][j] : 0.0) : ((i < j) ? X[j][i] : (i == j ? 1.0 : 0.0));
  };
  if (tid < KMAX) rinv[tid] = 1.0 / W[tid][tid];
  if (!BOTH) {
    for (int e = tid; e < KMAX * KMAX; e += NT1) {
      const int r = e & 63, c = e >> 6;
      if (r > c) X[r][c] = 0.0;
    }
  }
  lds_barrier();
  if (lane < (BOTH ? 32 : 16)) {
    const int m = (lane < 16) ? 1 : 0, c = lane & 15, base = BLK * wave;
    double x[BLK];
#pragma unroll
    for (int i = BLK - 1; i >= 0; i--) {
      double sacc = (i == c) ? 1.0 : 0.0;
#pragma unroll
      for (int l = i + 1; l < BLK; l++) sacc = fma(-Rm(m, base + i, base + l), x[l], sacc);
      x[i] = m ? sacc * rinv[base + i] : sacc;
    }
#pragma unroll
    for (int i = 0; i < BLK; i++) {
      if (m && i <= c) X[base + i][base + c] = x[i];
      if (!m && i < c) X[base + c][base + i] = x[i];
    }
  }
  lds_barrier();
  const int g = lane >> 4, li = lane & 15;
  for (int d = 1; d < 4; d++) {
    const int ntgt = (BOTH ? 2 : 1) * (4 - d);
    for (int t = wave; t < ntgt; t += 4) {
      const int m = BOTH ? (t & 1) : 1, bi = BOTH ? (t >> 1) : t, bj = bi + d;
      dbl4p acc = {0.0, 0.0, 0.0, 0.0};
      for (int kb = bi + 1; kb <= bj; kb++) {
#pragma unroll
        for (int k0 = 0; k0 < BLK; k0 += 4) {
          const double a = Rm(m, BLK * bi + li, BLK * kb + k0 + g);
          const double b = Xm(m, BLK * kb + k0 + g, BLK * bj + li);
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
      }
#pragma unroll
      for (int v = 0; v < 4; v++) Ys[wave][g + 4 * v][li] = acc[v];
      wave_lds_sync();
      dbl4p acc2 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k0 = 0; k0 < BLK; k0 += 4) {
        const double a = Xm(m, BLK * bi + li, BLK * bi + k0 + g);
        acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Ys[wave][k0 + g][li], acc2, 0, 0, 0);
      }
      wave_lds_sync();  // Ys is reused by the wave's next block
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = BLK * bi + g + 4 * v, c = BLK * bj + li;
        if (m)
          X[r][c] = -acc2[v];
        else
          X[c][r] = -acc2[v];
      }
    }
    lds_barrier();
  }
}

// ---------------------------------------------------------------------------
// R1 = chol(G) of the panel's Gram, held in S (identity beyond fjb; diag[]:
// its diagonal), stopped at the first pivot that proves the DM truncation
// fires there, and R1^-1.  Returns k1 >= 0 with S = R1 (on and above the
// diagonal, zero below) identity-padded beyond k1 and X = R1^-1 (upper,
// identity beyond k1); or -1 when a pivot does not resolve (the exact panel
// takes the step over).  Y: lu_inverse's scratch.  The caller's S and diag
// writes are ordered by a barrier before the call.
__device__ __forceinline__ int chol1_core(double (*S)[LDP], double (*X)[LDP], double (*Y)[BLK][BLK + 1],
                          double* rinv, const double* diag, int fjb, double tau_dm, int* s_done,
                          int* s_why) {
  const int tid = local_tid();
  // the largest diagonal entry (>= 0), a max over each wave's 64 lanes
  double maxd = fmax(diag[tid & 63], 0.0);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) maxd = fmax(maxd, __shfl_xor(maxd, o, 64));
  const double g00 = diag[0];
  const double tolp = 1e4 * 64.0 * kEpsP * fjb * maxd;  // unresolved pivots
  const double thres = (tau_dm > 0.0 && fjb > 1) ? tau_dm * sqrt(g00) : 5e-14;
  const double thr2 = 0.25 * thres * thres;
  const bool scale_ok = g00 > 1e-280 && g00 < 1e280 && maxd < 1e280;
  lds_barrier();
  int why = 0;
  const int k1 = chol_blocked(
      S, fjb,
      [&](int i, double d) -> int {
        // 1: the truncation at <= i is certain; 2: out of range or unresolved
        return !scale_ok ? 2 : ((i >= 1 && d < thr2 && thr2 > tolp) ? 1 : (!(d > tolp) ? 2 : 0));
      },
      rinv, s_done, s_why, &why);
  if (why == 2) return -1;
  // R1 (zero below the diagonal), identity-padded beyond k1 for the inverse
  for (int e = tid; e < KMAX * KMAX; e += NT1) {
    const int r = e >> 6, c = e & 63;
    const bool in = r < k1 && c < k1;
    const double v = (in && c >= r) ? S[r][c] : 0.0;
    S[r][c] = in ? v : (r == c ? 1.0 : 0.0);
  }
  lds_barrier();
  lu_inverse<false>(S, X, rinv, Y);  // X = R1^-1 (upper part)
  return k1;
}

// chol1_core on the panel's Gram gathered from the candidate Gram Gc
// (column-major, leading dimension KMAX) through the rank map; R1 and R1^-1
// out (column-major, zero outside k1 x k1), or the fallback flag.
__device__ void chol1_body(const double* __restrict__ Gc, const int* __restrict__ rank,
                           double tau_dm, double* R1, double* R1inv, DevState* st) {
  __shared__ double S[KMAX][LDP];
  __shared__ double X[KMAX][LDP];
  __shared__ double Y[4][BLK][BLK + 1];
  __shared__ double rinv[KMAX], diag[KMAX];
  __shared__ int srk[KMAX];
  __shared__ int s_done, s_why;
  if (!step_runs(st, RUN_FAST)) return;
  const int fjb = st->fjb;
  const int tid = threadIdx.x;
  // the whole candidate Gram and the rank map in one round trip (X holds the
  // Gram until the inverse overwrites it), then the gather from LDS
  {
    const int rk = (tid < fjb) ? rank[tid] : 0;
    double gv[KMAX * KMAX / NT1];
#pragma unroll
    for (int q = 0; q < KMAX * KMAX / NT1; q++) gv[q] = Gc[tid + NT1 * q];
    if (tid < KMAX) srk[tid] = rk;
#pragma unroll
    for (int q = 0; q < KMAX * KMAX / NT1; q++) {
      const int e = tid + NT1 * q;
      X[e & 63][e >> 6] = gv[q];  // X[r][c] = Gc(r, c)
    }
  }
  lds_barrier();
  if (tid < KMAX) diag[tid] = tid < fjb ? X[srk[tid]][srk[tid]] : 0.0;
#pragma unroll
  for (int q = 0; q < KMAX * KMAX / NT1; q++) {
    const int e = tid + NT1 * q, r = e >> 6, c = e & 63;
    S[r][c] = (r < fjb && c < fjb) ? X[srk[r]][srk[c]] : (r == c ? 1.0 : 0.0);
  }
  lds_barrier();
  const int k1 = chol1_core(S, X, Y, rinv, diag, fjb, tau_dm, &s_done, &s_why);
  if (k1 < 0) {
    if (tid == 0) {
      st->flags |= 2;  // the exact panel takes the step over
      st->halt = 2;
    }
    return;
  }
  for (int e = tid; e < KMAX * KMAX; e += NT1) {
    const int r = e & 63, c = e >> 6;  // consecutive lanes: consecutive rows (coalesced)
    const bool in = r < k1 && c < k1;
    R1[r + c * KMAX] = (in && c >= r) ? S[r][c] : 0.0;
    R1inv[r + c * KMAX] = in ? X[r][c] : 0.0;
  }
  if (tid == 0) st->k1 = k1;
}

// The step's column moves (workgroups 1..) and, for the Gram-based panel, the
// first Cholesky (workgroup 0, dispatched first): the Cholesky reads only the
// candidate Gram, so the two run side by side.
__global__ __launch_bounds__(256) void k_perm_chol1(double* A, int lda, int m, double* work,
                                                    int* jp, const int* __restrict__ perm,
                                                    const double* __restrict__ Gc,
                                                    double tau_dm, double* R1, double* R1inv,
                                                    DevState* st, StepBufs bufs) {
  if (blockIdx.x == 0)
    chol1_body(Gc, perm + 4 * KMAX, tau_dm, R1, R1inv, st);
  else
    apply_perm_rows(A, lda, m, work, jp, perm, st, blockIdx.x - 1, gridDim.x - 1, bufs);
}

void launch_perm_chol1(double* A, int lda, int m, double* work, int* jp, const int* perm,
                       const double* Gc, double tau_dm, double* R1, double* R1inv, DevState* st,
                       const StepBufs& bufs, hipStream_t s) {
  hipLaunchKernelGGL(k_perm_chol1, dim3(1 + apply_perm_blocks(m)), dim3(256), 0, s, A, lda, m,
                     work, jp, perm, Gc, tau_dm, R1, R1inv, st, bufs);
}

// ---------------------------------------------------------------------------
// 64 x 64 x 64 product on the matrix cores from LDS operands: wave w computes
// rows 16w..16w+15.  X is read as X[row][q] (or X[q][row] when xt), Y[q][col].
// acc[u][v] holds row 16w + lane/16 + 4v, column 16u + lane%16.
__device__ __forceinline__ void mfma_64(dbl4p (&acc)[4], const double (*X)[LDP], bool xt,
                                        const double (*Y)[LDP]) {
  const int lane = local_tid() & 63, wv = local_tid() >> 6;
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
  const int lane = local_tid() & 63, wv = local_tid() >> 6;
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++) S[16 * wv + (lane >> 4) + 4 * v][16 * u + (lane & 15)] = acc[u][v];
}

// 64 x 64 products with triangular factors on the matrix cores, as jobs of
// 16 x 16 output tiles (ti, tj) over 16-deep k blocks k0..k1, balanced over
// the four waves (five k blocks each for the triangular kinds):
//   TRI_UU  upper x upper (the result is upper): tiles ti <= tj, k blocks ti..tj;
//   TRI_LU  lower x upper, the upper half of the result: ti <= tj, k blocks 0..ti;
//   TRI_FU  full x upper: every tile, k blocks 0..tj (wave w: tile row w).
enum { TRI_UU = 0, TRI_LU = 1, TRI_FU = 2 };

struct TileJob {
  int ti, tj, k0, k1;
};

__device__ __forceinline__ int tri_jobs(int kind, int wave, TileJob (&jb)[4]) {
  if (kind == TRI_FU) {
#pragma unroll
    for (int t = 0; t < 4; t++) jb[t] = TileJob{wave, t, 0, t};
    return 4;
  }
  if (kind == TRI_UU) {
    switch (wave) {
      case 0: jb[0] = TileJob{0, 3, 0, 3}; jb[1] = TileJob{0, 0, 0, 0}; return 2;
      case 1: jb[0] = TileJob{0, 2, 0, 2}; jb[1] = TileJob{0, 1, 0, 1}; return 2;
      case 2: jb[0] = TileJob{1, 3, 1, 3}; jb[1] = TileJob{1, 2, 1, 2}; return 2;
      default:
        jb[0] = TileJob{2, 3, 2, 3}; jb[1] = TileJob{1, 1, 1, 1};
        jb[2] = TileJob{2, 2, 2, 2}; jb[3] = TileJob{3, 3, 3, 3}; return 4;
    }
  }
  switch (wave) {  // TRI_LU
    case 0: jb[0] = TileJob{3, 3, 0, 3}; jb[1] = TileJob{0, 0, 0, 0}; return 2;
    case 1: jb[0] = TileJob{2, 3, 0, 2}; jb[1] = TileJob{1, 1, 0, 1}; return 2;
    case 2: jb[0] = TileJob{2, 2, 0, 2}; jb[1] = TileJob{1, 2, 0, 1}; return 2;
    default:
      jb[0] = TileJob{1, 3, 0, 1}; jb[1] = TileJob{0, 1, 0, 0};
      jb[2] = TileJob{0, 2, 0, 0}; jb[3] = TileJob{0, 3, 0, 0}; return 4;
  }
}

// acc[t] = the product's tile jb[t] (fa(row, q), fb(q, col) element access);
// acc[t][v] holds row 16 ti + lane/16 + 4v, column 16 tj + lane%16.
template <class FA, class FB>
__device__ __forceinline__ int tri_product(int kind, FA fa, FB fb, TileJob (&jb)[4],
                                           dbl4p (&acc)[4]) {
  const int lane = local_tid() & 63, wave = local_tid() >> 6, g = lane >> 4, li = lane & 15;
  const int nj = tri_jobs(kind, wave, jb);
#pragma unroll
  for (int t = 0; t < 4; t++) {
    acc[t] = (dbl4p){0.0, 0.0, 0.0, 0.0};
    if (t < nj) {
      for (int kb = jb[t].k0; kb <= jb[t].k1; kb++) {
#pragma unroll
        for (int k4 = 0; k4 < BLK; k4 += 4) {
          const int q = BLK * kb + k4 + g;
          acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa(BLK * jb[t].ti + li, q),
                                                        fb(q, BLK * jb[t].tj + li), acc[t], 0, 0,
                                                        0);
        }
      }
    }
  }
  return nj;
}

// Ps[r][c] = A(rs + r, j + c) for c < ncol and rows < m, else 0 (a 64-row
// slab of the panel); 256 threads, all loads in flight before the LDS writes.
__device__ __forceinline__ void load_slab(double (*Ps)[LDP], const double* __restrict__ A,
                                          int lda, int j, int rs, int m, int ncol) {
  const int tid = local_tid();
  double v[KMAX * PR_ROWS / 256];
#pragma unroll
  for (int q = 0; q < KMAX * PR_ROWS / 256; q++) {
    const int e = tid + 256 * q, c = e / PR_ROWS, r = e % PR_ROWS;
    v[q] = (c < ncol && rs + r < m) ? A[(size_t)(j + c) * lda + rs + r] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < KMAX * PR_ROWS / 256; q++) {
    const int e = tid + 256 * q;
    Ps[e % PR_ROWS][e / PR_ROWS] = v[q];
  }
}

// S[r][c] = M[r + c * KMAX] for a column-major 64 x 64 matrix in memory.
__device__ __forceinline__ void load_mat(double (*S)[LDP], const double* __restrict__ M) {
  const int tid = local_tid();
  double v[KMAX * KMAX / 256];
#pragma unroll
  for (int q = 0; q < KMAX * KMAX / 256; q++) v[q] = M[tid + 256 * q];
#pragma unroll
  for (int q = 0; q < KMAX * KMAX / 256; q++) {
    const int e = tid + 256 * q;
    S[e & 63][e >> 6] = v[q];
  }
}

// The sum over the four waves (in wave order) of their upper-tile Grams gt,
// both triangles, column-major into out (64 x 64).  Gp: LDS for the waves'
// tiles.  Ends after the writes (no barrier).
__device__ __forceinline__ void gram_sum_store(const dbl4p (&gt)[10], double (*Gp)[10][4][64],
                                               double* out) {
  const int tid = local_tid(), lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int p = 0; p < 10; p++)
#pragma unroll
    for (int v = 0; v < 4; v++) Gp[wv][p][v][lane] = gt[p][v];
  lds_barrier();
  for (int idx = tid; idx < 10 * 4 * 64; idx += 256) {
    const int p = idx >> 8, v = (idx >> 6) & 3, ln = idx & 63;
    const int u = p < 4 ? 0 : (p < 7 ? 1 : (p < 9 ? 2 : 3));
    const int u2 = u + p - (u == 0 ? 0 : (u == 1 ? 4 : (u == 2 ? 7 : 9)));
    const double x = ((Gp[0][p][v][ln] + Gp[1][p][v][ln]) + Gp[2][p][v][ln]) + Gp[3][p][v][ln];
    const int r = 16 * u + (ln >> 4) + 4 * v, c = 16 * u2 + (ln & 15);
    out[r + (size_t)c * KMAX] = x;
    if (u != u2) out[c + (size_t)r * KMAX] = x;
  }
}

// Pass 2: Q = P R1^-1 on a 64-row slab, partial Gram Q^T Q.  Wave w takes
// rows 16w .. 16w + 15 of the slab, read straight into the matrix cores' A
// operand; the product's output layout is the operand layout of Q^T Q over
// those rows (gram_tiles), so nothing but R1^-1 goes through LDS.  Slabs
// past m (the grid is sized for the lowest j) write zero partials.
__global__ __launch_bounds__(256) void k_qgram(const double* __restrict__ A, int lda, int m,
                                               const double* __restrict__ R1inv,
                                               const DevState* st, double* partial) {
  __shared__ double Ri[KMAX][LDP];
  __shared__ double Gp[4][10][4][64];
  if (!step_runs(st, RUN_FAST)) return;
  const int j = st->j;
  const int k1 = st->k1;
  const int tid = local_tid(), lane = tid & 63, wv = tid >> 6, g = lane >> 4, li = lane & 15;
  const int rb = j + blockIdx.x * PR_ROWS + 16 * wv;
  load_mat(Ri, R1inv);
  double pa[KMAX / 4];
#pragma unroll
  for (int t = 0; t < KMAX / 4; t++)
    pa[t] = (rb + li < m && 4 * t + g < k1) ? A[(size_t)(j + 4 * t + g) * lda + rb + li] : 0.0;
  lds_barrier();
  dbl4p qb[4];
#pragma unroll
  for (int u = 0; u < 4; u++) qb[u] = (dbl4p){0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < KMAX / 4; t++)
#pragma unroll
    for (int u = 0; u < 4; u++)
      qb[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[t], Ri[4 * t + g][16 * u + li], qb[u], 0, 0, 0);
  dbl4p gt[10];
#pragma unroll
  for (int p = 0; p < 10; p++) gt[p] = (dbl4p){0.0, 0.0, 0.0, 0.0};
  gram_tiles(qb, gt);
  gram_sum_store(gt, Gp, partial + (size_t)blockIdx.x * KMAX * KMAX);
}

// ---------------------------------------------------------------------------
// Reconstruction (see the file header) on LDS matrices: SX = G2 (identity
// beyond k1), SR = R1 (zero beyond k1).  r1inv_into(SX) writes R1^-1 (upper,
// zero outside k1 x k1) into SX, reading SA where SA holds it (it then ends
// with a barrier; SA is overwritten after the call); top_into(SW) writes the
// panel's top block A(j + r, j + c), r, c < k1, zero elsewhere.  Writes the
// top k x k block of the panel at column j (R above and on the diagonal, V1
// below), tau[0..k) and M (column-major, leading dimension KMAX) to M_out,
// and returns k; or returns -1, with nothing written, when the Gram-based
// form falls back.  Y may share SW's storage.  The caller's LDS writes are
// ordered by a barrier before the call.
template <class R1InvFn, class TopFn>
__device__ __forceinline__ int recon_core(double (*SR)[LDP], double (*SA)[LDP], double (*SW)[LDP],
                          double (*SX)[LDP], double (*Y)[BLK][BLK + 1], double* rinv,
                          double* dsg, double* tl, double* scv, int* s_done, int* s_why,
                          R1InvFn r1inv_into, TopFn top_into, double* A, int lda, int j,
                          int fjb, int k1, double tau_dm, double* tau, double* M_out) {
  const int tid = local_tid();
  // 1-3. R = R2 R1 with R2^T R2 = G2 and R^-1 (into SR and SA).  After the
  // first pass G2 = I + E with E tiny: then R2 = I + F, F upper triangular
  // with F + F^T + F^T F = E, is the fixed point of F = triu_half(E - F^T F)
  // (triu_half: the upper triangle with the diagonal halved), reached from
  // F = triu_half(E) with an error of order |E|^(k+2) after k more passes,
  // and R2^-1 = I - F + F^2 - ...; each pass is one 64 x 64 x 64 product on
  // the matrix cores.  Larger E: blocked Cholesky and inverse.
  double emax;
  {
    double em = 0.0;
    for (int e = tid; e < KMAX * KMAX; e += NT1) {
      const int r = e >> 6, c = e & 63;
      em = fmax(em, fabs(SX[r][c] - (r == c ? 1.0 : 0.0)));
    }
    em = fmax(em, __shfl_xor(em, 32, 64));
    em = fmax(em, __shfl_xor(em, 16, 64));
    em = fmax(em, __shfl_xor(em, 8, 64));
    em = fmax(em, __shfl_xor(em, 4, 64));
    em = fmax(em, __shfl_xor(em, 2, 64));
    em = fmax(em, __shfl_xor(em, 1, 64));
    if ((tid & 63) == 0) rinv[tid >> 6] = em;
    lds_barrier();
    emax = fmax(fmax(rinv[0], rinv[1]), fmax(rinv[2], rinv[3]));
    lds_barrier();
  }
  dbl4p acc[4];
  int why = 0;
  if (emax < 1e-3) {
    // passes: |E|^(npass + 2) <= 2^-60 (relative to the identity)
    int npass = 0;
    for (double err = emax * emax; err > 8.7e-19 && npass < 6; err *= emax) npass++;
    // SW = E, SX = F = triu_half(E)
    for (int e = tid; e < KMAX * KMAX; e += NT1) {
      const int r = e >> 6, c = e & 63;
      const double ev = SX[r][c] - (r == c ? 1.0 : 0.0);
      SW[r][c] = ev;
      SX[r][c] = (r < c) ? ev : (r == c ? 0.5 * ev : 0.0);
    }
    lds_barrier();
    TileJob jb[4];
    const int lane = tid & 63, g = lane >> 4, li = lane & 15;
    for (int ps = 0; ps < npass; ps++) {
      // F := triu_half(E - F^T F): the upper half of F^T F (F upper)
      const int nj = tri_product(
          TRI_LU, [&](int r, int q) { return SX[q][r]; }, [&](int q, int c) { return SX[q][c]; },
          jb, acc);
      lds_barrier();
#pragma unroll
      for (int t = 0; t < 4; t++)  // static indices into jb and acc
#pragma unroll
        for (int v = 0; v < 4; v++) if (t < nj) {
          const int r = BLK * jb[t].ti + g + 4 * v, c = BLK * jb[t].tj + li;
          const double x = SW[r][c] - acc[t][v];
          SX[r][c] = (r < c) ? x : (r == c ? 0.5 * x : 0.0);
        }
      lds_barrier();
    }
    // R = R1 + F R1 (both upper)
    {
      const int nj = tri_product(
          TRI_UU, [&](int r, int q) { return SX[r][q]; }, [&](int q, int c) { return SR[q][c]; },
          jb, acc);
      lds_barrier();
#pragma unroll
      for (int t = 0; t < 4; t++)  // static indices into jb and acc
#pragma unroll
        for (int v = 0; v < 4; v++) if (t < nj) SR[BLK * jb[t].ti + g + 4 * v][BLK * jb[t].tj + li] += acc[t][v];
    }
    // R2^-1 = I - F (I - F (...)) into SW (upper)
    for (int e = tid; e < KMAX * KMAX; e += NT1) {
      const int r = e >> 6, c = e & 63;
      SW[r][c] = (r == c ? 1.0 : 0.0) - SX[r][c];
    }
    lds_barrier();
    for (int ps = 0; ps < npass; ps++) {
      const int nj = tri_product(
          TRI_UU, [&](int r, int q) { return SX[r][q]; }, [&](int q, int c) { return SW[q][c]; },
          jb, acc);
      lds_barrier();
#pragma unroll
      for (int t = 0; t < 4; t++)  // static indices into jb and acc
#pragma unroll
        for (int v = 0; v < 4; v++) if (t < nj) {
          const int r = BLK * jb[t].ti + g + 4 * v, c = BLK * jb[t].tj + li;
          SW[r][c] = (r == c ? 1.0 : 0.0) - acc[t][v];
        }
      lds_barrier();
    }
    // R^-1 = R1^-1 R2^-1 (both upper) into SA, zero below the diagonal blocks
    r1inv_into(SX);
    for (int e = tid; e < KMAX * KMAX; e += NT1) {
      const int r = e & 63, c = e >> 6;
      if ((r >> 4) > (c >> 4)) SA[r][c] = 0.0;
    }
    lds_barrier();
    {
      const int nj = tri_product(
          TRI_UU, [&](int r, int q) { return SX[r][q]; }, [&](int q, int c) { return SW[q][c]; },
          jb, acc);
#pragma unroll
      for (int t = 0; t < 4; t++)  // static indices into jb and acc
#pragma unroll
        for (int v = 0; v < 4; v++) if (t < nj) SA[BLK * jb[t].ti + g + 4 * v][BLK * jb[t].tj + li] = acc[t][v];
    }
    lds_barrier();
  } else {
    chol_blocked(
        SX, k1, [&](int, double d) -> int { return (d > 0.25 && d < 4.0) ? 0 : 2; }, rinv,
        s_done, s_why, &why);
    if (why) return -1;
    for (int e = tid; e < KMAX * KMAX; e += NT1) {
      const int r = e >> 6, c = e & 63;
      if (!(r < k1 && c < k1 && c >= r)) SX[r][c] = 0.0;
    }
    lds_barrier();
    mfma_64(acc, SX, false, SR);  // R = R2 R1, identity-padded for the inverse
    lds_barrier();
    mfma_store(SR, acc);
    lds_barrier();
    if (tid >= k1 && tid < KMAX) SR[tid][tid] = 1.0;
    lds_barrier();
    lu_inverse<false>(SR, SA, rinv, Y);  // SA = R^-1 (upper part)
  }
  // 4. Q11 = P_top R^-1 -> SW
  top_into(SW);
  lds_barrier();
  {
    TileJob jq[4];
    const int lane = tid & 63, g = lane >> 4, li = lane & 15;
    tri_product(
        TRI_FU, [&](int r, int q) { return SW[r][q]; }, [&](int q, int c) { return SA[q][c]; }, jq,
        acc);
    lds_barrier();
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) SW[BLK * jq[t].ti + g + 4 * v][BLK * jq[t].tj + li] = acc[t][v];
  }
  lds_barrier();
  // 5. LU of Q11 D - I with dlarfg's signs and the truncation test
  const double thres = (tau_dm > 0.0 && fjb > 1) ? tau_dm * SR[0][0] : 5e-14;
  const int k = lu_signed_blocked(SW, k1, SR, thres, dsg, tl, scv, s_done, s_why, &why);
  if (why == 2) return -1;
  // 6. top block (R rows scaled by the signs, V1 below), tau; SX := the
  // signed U (on and above the diagonal) and L (below), identity beyond k
  for (int e = tid; e < KMAX * KMAX; e += NT1) {
    const int r = e & 63, c = e >> 6;  // consecutive threads: consecutive rows
    const bool in = r < k && c < k;
    const double w = SW[r][c];
    if (in) A[(size_t)(j + c) * lda + j + r] = (r <= c) ? dsg[r] * SR[r][c] : w;
    SX[r][c] = in ? (r < c ? dsg[c] * w : (r == c ? -tl[c] : w)) : (r == c ? 1.0 : 0.0);
  }
  if (tid < k) tau[tid] = tl[tid];
  lds_barrier();
  // 7. SR := U^-1 (on and above the diagonal; zero below)
  lu_inverse<false>(SX, SR, rinv, Y);
  // 8. M = R^-1 D U^-1, upper triangular, computed transposed (lower) so that
  // consecutive lanes store consecutive rows: the TRI_UU jobs with the tile
  // indices swapped; the tiles below the diagonal blocks are zero.  (T and
  // tau follow from the V that k_v2 stores, k_reduce_t.)
  {
    const int lane = tid & 63, wv = tid >> 6, li = lane & 15, g = lane >> 4;
    TileJob jt[4];
    const int nj = tri_jobs(TRI_UU, wv, jt);
    dbl4p am[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      am[t] = (dbl4p){0.0, 0.0, 0.0, 0.0};
      if (t < nj) {
        const int cc = BLK * jt[t].tj + li;  // column of M (row of the transposed product)
        const int rr = BLK * jt[t].ti + li;  // row of M
        for (int kb = jt[t].k0; kb <= jt[t].k1; kb++) {
#pragma unroll
          for (int k4 = 0; k4 < BLK; k4 += 4) {
            const int q = BLK * kb + k4 + g;
            // M^T(c, r) = sum_q (D U^-1)(q, c) R^-1(r, q)
            const double a_m = (q <= cc && q < k) ? dsg[q] * SR[q][cc] : 0.0;
            am[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_m, SA[rr][q], am[t], 0, 0, 0);
          }
        }
        // tile rows = columns c of M, tile columns = rows r
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int c = BLK * jt[t].tj + g + 4 * v, r = BLK * jt[t].ti + li;
          M_out[r + c * KMAX] = (r < k && c < k) ? am[t][v] : 0.0;
        }
      }
    }
    for (int e = tid; e < KMAX * KMAX; e += NT1) {
      const int r = e & 63, c = e >> 6;
      if ((r >> 4) > (c >> 4)) M_out[r + c * KMAX] = 0.0;
    }
  }
  return k;
}

// What the reconstruction reads besides G2, all written by earlier kernels
// of the step: read by every workgroup of the reduce before its ticket, so
// that it arrives with the partials (only the last workgroup uses it).
// Column-major 64 x 64 in registers, thread tid holding e = tid + 256 q.
struct ReconIn {
  int j, fjb, k1;
  double gr[KMAX * KMAX / NT1];    // R1
  double pri[KMAX * KMAX / NT1];   // R1^-1
  double ptop[KMAX * KMAX / NT1];  // the panel's top block
};
__device__ __forceinline__ void recon_inputs(ReconIn& in, const DevState* st,
                                             const double* __restrict__ R1,
                                             const double* __restrict__ R1inv,
                                             const double* A, int lda) {
  const int tid = local_tid();
  in.j = st->j;
  in.fjb = st->fjb;
  in.k1 = st->k1;
  // e walks rows fastest: coalesced column-major reads
#pragma unroll
  for (int q = 0; q < KMAX * KMAX / NT1; q++) {
    const int e = tid + NT1 * q, r = e & 63, c = e >> 6;
    const bool in_blk = r < in.k1 && c < in.k1;
    in.gr[q] = R1[r + c * KMAX];
    in.pri[q] = R1inv[r + c * KMAX];
    in.ptop[q] = in_blk ? A[(size_t)(in.j + c) * lda + in.j + r] : 0.0;
  }
}

// The reconstruction of the DM loop's Gram-based panel: G2 from memory, the
// rest from rin; writes st->fjb_cmp and st->c0, or the fallback flag.
__device__ __forceinline__ void recon_body(const double* __restrict__ G2, const ReconIn& rin,
                                           double* A, int lda, double tau_dm, double* tau,
                                           DevState* st, double* M_out) {
  __shared__ double SR[KMAX][LDP];  // R1, then R, then L^T, then U^-1 (rows scaled by D)
  __shared__ double SA[KMAX][LDP];  // R^-1
  __shared__ double SW[KMAX][LDP];  // E, R2^-1, then Q11 and its LU, then L^-T
  __shared__ double SX[KMAX][LDP];  // G2, F or R2, R1^-1, then U
  __shared__ double Y[4][BLK][BLK + 1];
  __shared__ double rinv[KMAX], dsg[KMAX], tl[KMAX], scv[KMAX];
  __shared__ int s_done, s_why;
  const int j = rin.j, fjb = rin.fjb, k1 = rin.k1;
  const int tid = threadIdx.x;
  // G2 from the launch's reduce (the one global read left); R1 from the
  // registers read before the ticket
  {
    double gx[KMAX * KMAX / NT1];
#pragma unroll
    for (int q = 0; q < KMAX * KMAX / NT1; q++) {
      const int e = tid + NT1 * q, r = e & 63, c = e >> 6;
      const bool in = r < k1 && c < k1;
      gx[q] = in ? ld_dev(G2 + r + c * KMAX) : (r == c ? 1.0 : 0.0);
    }
#pragma unroll
    for (int q = 0; q < KMAX * KMAX / NT1; q++) {
      const int e = tid + NT1 * q, r = e & 63, c = e >> 6;
      SX[r][c] = gx[q];
      SR[r][c] = rin.gr[q];
    }
  }
  lds_barrier();
  auto from_regs = [&](const double(&v)[KMAX * KMAX / NT1], double(*S)[LDP]) {
#pragma unroll
    for (int q = 0; q < KMAX * KMAX / NT1; q++) {
      const int e = tid + NT1 * q;
      S[e & 63][e >> 6] = v[q];
    }
  };
  const int k = recon_core(
      SR, SA, SW, SX, Y, rinv, dsg, tl, scv, &s_done, &s_why,
      [&](double(*S)[LDP]) { from_regs(rin.pri, S); },
      [&](double(*S)[LDP]) { from_regs(rin.ptop, S); }, A,
      lda, j, fjb, k1, tau_dm, tau + j, M_out);
  if (k < 0) {
    if (tid == 0) {
      st->flags |= 2;  // the exact panel takes the step over
      st->halt = 2;
    }
    return;
  }
  if (tid == 0) {
    st->fjb_cmp = k;
    // the trailing update also applies the kept reflectors to the rejected
    // in-block columns j+k..j+fjb-1 (dlarf_ did that inside the reference's
    // panel)
    st->c0 = j + k;
  }
}

// G2 = the sum of pass 2's partial Grams (as k_gram_reduce), then the last
// workgroup to finish (atomic ticket, reset for the next launch) runs the
// reconstruction on it.
__global__ __launch_bounds__(256) void k_reduce_recon(const double* __restrict__ partial,
                                                      int nslab, double* G2,
                                                      const double* __restrict__ R1,
                                                      const double* __restrict__ R1inv,
                                                      double* A, int lda, double tau_dm,
                                                      double* tau, double* T_out, double* M_out,
                                                      unsigned* ticket, DevState* st) {
  __shared__ double part[4][64];
  __shared__ int s_last;
  if (!step_runs(st, RUN_FAST)) return;
  ReconIn rin;
  recon_inputs(rin, st, R1, R1inv, A, lda);
  const int e = blockIdx.x * 64 + (threadIdx.x & 63);
  const int q = threadIdx.x >> 6;
  double sacc = 0.0;
  for (int b0 = q; b0 < nslab; b0 += 32) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int b = b0 + 4 * u;
      v[u] = partial[(size_t)(b < nslab ? b : q) * KMAX * KMAX + e];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) sacc += (b0 + 4 * u < nslab) ? v[u] : 0.0;
  }
  part[q][threadIdx.x & 63] = sacc;
  __syncthreads();
  if (q == 0)
    st_dev(G2 + e, (part[0][threadIdx.x] + part[1][threadIdx.x]) +
                       (part[2][threadIdx.x] + part[3][threadIdx.x]));
  handoff_wait();
  __syncthreads();
  if (threadIdx.x == 0) s_last = last_to_arrive(ticket, gridDim.x) ? 1 : 0;
  __syncthreads();
  if (!s_last) return;
  recon_body(G2, rin, A, lda, tau_dm, tau, st, M_out);
}

// ---------------------------------------------------------------------------
// Pass 3: V2 = P_bottom M on a 64-row slab; writes V2 into the panel below the
// top block and the dense zero-padded V (rows [j, m), KMAX columns) with V1's
// unit lower block on top, and the slab's V^T V.  As in pass 2, wave w takes
// rows 16w .. 16w + 15 straight into the A operand (M through LDS), and V^T V
// accumulates from the product's layout.
__global__ __launch_bounds__(256) void k_v2(double* A, int lda, int m,
                                            const double* __restrict__ M, const DevState* st,
                                            StepBufs bufs, double* gram_part) {
  __shared__ double Ms[KMAX][LDP];
  __shared__ double Gp[4][10][4][64];
  if (!step_runs(st, RUN_FAST)) return;
  double* V = v_cur(bufs, st);
  const int ldv = bufs.ldv;
  const int j = st->j;
  const int k = st->fjb_cmp;
  const int tid = local_tid(), lane = tid & 63, wv = tid >> 6, g = lane >> 4, li = lane & 15;
  const int rs = j + blockIdx.x * PR_ROWS;
  if (rs >= m) return;
  if (blockIdx.x == 0) {
    // V is zero above the panel: the trailing update's slabs start at the
    // 32-row boundary at or below j and rely on it
    const int z0 = j >= KMAX ? j - KMAX : 0;
    for (int e = tid; e < KMAX * KMAX; e += NT1) {
      const int r = z0 + (e & 63), c = e >> 6;
      if (r < j) V[(size_t)c * ldv + r] = 0.0;
    }
  }
  const int rb = rs + 16 * wv;
  load_mat(Ms, M);
  double pa[KMAX / 4];
#pragma unroll
  for (int t = 0; t < KMAX / 4; t++)
    pa[t] = (rb + li < m && 4 * t + g < k) ? A[(size_t)(j + 4 * t + g) * lda + rb + li] : 0.0;
  lds_barrier();
  dbl4p vb[4];
#pragma unroll
  for (int u = 0; u < 4; u++) vb[u] = (dbl4p){0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < KMAX / 4; t++)
#pragma unroll
    for (int u = 0; u < 4; u++)
      vb[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[t], Ms[4 * t + g][16 * u + li], vb[u], 0, 0, 0);
  // rows of the top block keep V1 (unit lower, written by the
  // reconstruction): overlay; then V and the panel's V2 rows out
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = rb + g + 4 * v, c = 16 * u + li, pr = r - j;
      double x = vb[u][v];
      if (c >= k || r >= m) x = 0.0;
      else if (pr < k) x = (pr < c) ? 0.0 : (pr == c ? 1.0 : A[(size_t)(j + c) * lda + r]);
      vb[u][v] = x;
      if (r < m) {
        V[(size_t)c * ldv + r] = x;
        if (c < k && pr >= k) A[(size_t)(j + c) * lda + r] = x;
      }
    }
  // the slab's part of V^T V, for T and tau
  dbl4p gt[10];
#pragma unroll
  for (int p = 0; p < 10; p++) gt[p] = (dbl4p){0.0, 0.0, 0.0, 0.0};
  gram_tiles(vb, gt);
  gram_sum_store(gt, Gp, gram_part + (size_t)blockIdx.x * KMAX * KMAX);
}

// The block reflector of the Gram-based panel from the V it stored: S = V^T V
// summed from k_v2's slab partials (fixed order), then tau_i = 2 / S_ii and
// T = (striu(S) + diag(S) / 2)^-1, so that H_1 ... H_k = I - V T V^T holds for
// the stored V to rounding (the compact WY identity; Joffrain, Low,
// Quintana-Orti, van de Geijn, van Zee, "Accumulating Householder
// transformations, revisited", ACM TOMS 32(2), 2006).  The reconstruction's own
// tau and T agree with these only to about eps times the square of the
// panel's condition number, and a block reflector that far from orthogonal
// would carry that error into the trailing matrix.  The inverse runs in the
// last workgroup to finish (atomic ticket).
__global__ __launch_bounds__(256) void k_reduce_t(const double* __restrict__ partial, int m,
                                                  double* Sg, double* tau, double* T_out,
                                                  unsigned* ticket, const DevState* st) {
  __shared__ double part[4][64];
  __shared__ double SX[KMAX][LDP];
  __shared__ double SY[KMAX][LDP];
  __shared__ double Y[4][BLK][BLK + 1];
  __shared__ double rinv[KMAX];
  __shared__ int s_last;
  if (!step_runs(st, RUN_FAST)) return;
  const int j = st->j, k = st->fjb_cmp;
  const int nslab = (m - j + PR_ROWS - 1) / PR_ROWS;
  const int e = blockIdx.x * 64 + (threadIdx.x & 63);
  const int q = threadIdx.x >> 6;
  double sacc = 0.0;
  for (int b0 = q; b0 < nslab; b0 += 32) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int b = b0 + 4 * u;
      v[u] = partial[(size_t)(b < nslab ? b : q) * KMAX * KMAX + e];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) sacc += (b0 + 4 * u < nslab) ? v[u] : 0.0;
  }
  part[q][threadIdx.x & 63] = sacc;
  __syncthreads();
  if (q == 0)
    st_dev(Sg + e, (part[0][threadIdx.x] + part[1][threadIdx.x]) +
                       (part[2][threadIdx.x] + part[3][threadIdx.x]));
  handoff_wait();
  __syncthreads();
  if (threadIdx.x == 0) s_last = last_to_arrive(ticket, gridDim.x) ? 1 : 0;
  __syncthreads();
  if (!s_last) return;
  const int tid = threadIdx.x;
  {
    // every load in flight before the LDS stores (one round trip)
    double sv[KMAX * KMAX / NT1];
#pragma unroll
    for (int q = 0; q < KMAX * KMAX / NT1; q++) {
      const int e2 = tid + NT1 * q, r = e2 & 63, c = e2 >> 6;  // consecutive rows
      sv[q] = (r < k && c < k) ? ld_dev(Sg + r + c * KMAX) : 0.0;
    }
#pragma unroll
    for (int q = 0; q < KMAX * KMAX / NT1; q++) {
      const int e2 = tid + NT1 * q, r = e2 & 63, c = e2 >> 6;
      SX[r][c] = (r < k && c < k) ? ((r < c) ? sv[q] : (r == c ? 0.5 * sv[q] : 0.0))
                                  : ((r == c) ? 1.0 : 0.0);
    }
  }
  lds_barrier();
  if (tid < k) tau[j + tid] = 1.0 / SX[tid][tid];  // 2 / S_ii
  lu_inverse<false>(SX, SY, rinv, Y);  // SY = the upper triangular inverse
  for (int e2 = tid; e2 < KMAX * KMAX; e2 += NT1) {
    const int r = e2 & 63, c = e2 >> 6;
    T_out[r + c * KMAX] = (r < k && c < k && r <= c) ? SY[r][c] : 0.0;
  }
}

void launch_panel_fast(double* A, int lda, int j0, int m, double tau_dm, double* tau,
                       double* gram_part, double* R1, double* R1inv, double* G2, double* T,
                       double* M, const StepBufs& bufs, unsigned* ticket, DevState* st,
                       hipStream_t s) {
  const int nslab = (m - j0 + PR_ROWS - 1) / PR_ROWS;
  hipLaunchKernelGGL(k_qgram, dim3(nslab), dim3(256), 0, s, A, lda, m, R1inv, st, gram_part);
  hipLaunchKernelGGL(k_reduce_recon, dim3(KMAX * KMAX / 64), dim3(256), 0, s, gram_part, nslab,
                     G2, R1, R1inv, A, lda, tau_dm, tau, T, M, ticket, st);
  hipLaunchKernelGGL(k_v2, dim3(nslab), dim3(256), 0, s, A, lda, m, M, st, bufs, gram_part);
  hipLaunchKernelGGL(k_reduce_t, dim3(KMAX * KMAX / 64), dim3(256), 0, s, gram_part, m, G2, tau,
                     T, ticket + 1, st);
}

}  // namespace qrdm
