// trailing.hip -- the DM loop's trailing update as one pass over C per step.
//
// The reference applies each DM step's block reflector to the trailing
// matrix in one dlarfb call (src/dgeqrdm_work.c:762-767, src/dlarfb.c:142):
// W = T^T V^T C, then C -= V W -- three sweeps of C per step (read for W,
// read and write for the update).  Here the update of step s is deferred and
// merged with the next step's first product, so that C is read and written
// once per step:
//
//   k_fused (step s+1):  C -= V_s W_s            (step s's update, deferred)
//                        Wp[slab] = V_{s+1}^T C  (step s+1's partial products
//                                                 on the updated C)
//   k_wsum_norm (s+1):   W_{s+1} = T^T sum_slab Wp; the step's R12 rows
//                        C_top - V_top W_{s+1} (final R rows, written); the
//                        norm downdate of every trailing column from them
//                        (src/dgeqrdm_work.c:80-109); columns whose norm must
//                        be recomputed get the update at once (all rows) and
//                        their W column is cleared.
//
// What the next step needs before its V exists is brought forward: the R12
// rows and the norms (above), and the candidate columns, which the candidate
// Gram kernel updates as it reads them (their W columns are then cleared).
// Columns that move in the step's permutation carry their W column with
// them.  W is stored by absolute column (64 x n per buffer, a column's 64
// entries contiguous); V and W of two consecutive steps alternate between two
// buffers (st->it & 1).
//
// Rounding: each element sees the same products as in the one-step form,
// accumulated in another order (V W is summed into C on the matrix cores
// instead of being formed first).
#include "qrdm_internal.h"

namespace qrdm {

constexpr int FU_CH = 16;         // rows per chunk
constexpr int FU_NWG = 4;         // waves per workgroup
constexpr int FU_MAX_SPLIT = 16;  // row slabs of the partial products
// LDS layouts of the V chunks (64 banks of 4 bytes), measured free of bank
// conflicts for the accesses below: V_s by rows of 66 doubles (the product-1
// operand is read as 16-byte pairs of adjacent columns, a lane's rows 33
// pairs apart), V_{s+1} by columns of 17 doubles (read as pairs of rows).
constexpr int FU_VP_LD = 66;
constexpr int FU_VN_LD = FU_CH + 1;

// Launch shapes of the fused pass: U 16-column blocks per wave (a workgroup
// covers 64 U columns), C prefetched PF chunks ahead, OCC waves per SIMD.
template <int U>
struct FuShape;
template <>
struct FuShape<1> {
  static constexpr int PF = 2, OCC = 3;
};
template <>
struct FuShape<2> {
  static constexpr int PF = 1, OCC = 2;
};

// Geometry of one fused pass over rows [r_lo, m) and columns [c0, n):
// chunks of FU_CH rows from the 16-row boundary at or below r_lo, split into
// ns slabs of whole chunks, and column groups of gc columns.
struct FuGeom {
  int base, nchunk, slab_ch, cgroups;
};

__host__ __device__ __forceinline__ FuGeom fused_geom(int r_lo, int m, int c0, int n, int ns,
                                                      int gc) {
  FuGeom g;
  g.base = r_lo & ~(FU_CH - 1);
  g.nchunk = (m - g.base + FU_CH - 1) / FU_CH;
  if (g.nchunk < 0) g.nchunk = 0;
  g.slab_ch = (g.nchunk + ns - 1) / ns;
  if (g.slab_ch < 1) g.slab_ch = 1;
  g.cgroups = (n - c0 + gc - 1) / gc;
  if (g.cgroups < 0) g.cgroups = 0;
  return g;
}

// Workgroups of a fused pass per XCD: column group g goes to XCD g % 8 with
// all its row slabs, so that the step's end (one workgroup per 64-column
// tile, dispatched round-robin to the XCDs: tile g on XCD g % 8) reads the
// slab partials of its columns from its own L2.
__host__ __device__ __forceinline__ int fused_per_xcd(int ns, int cgroups) {
  return ((cgroups + 7) / 8) * ns;
}

// One pass.  APPLY: C -= Vp Wprev (Vp, Wprev of the previous step) before
// the products; ACCUM: Wp[slab] = Vn^T C.  Rows [r_lo, m) of columns
// [c0, n) are written (APPLY only).  Wave w of a workgroup owns 16 U
// columns; lane group g = lane / 16 holds rows 4g..4g+3 of a 16-row chunk,
// so C moves in 32-byte vectors (lda % 4 == 0, 32-byte aligned A); the V
// chunks are shared through LDS.  W and the partials have leading dimension
// n (absolute columns); ns slabs.  With st (the DM loop) r_lo, c0 and the
// buffers come from st; the grid is sized for the lowest j and surplus
// workgroups return.
template <bool APPLY, bool ACCUM, int U>
__global__ __launch_bounds__(64 * FU_NWG, FuShape<U>::OCC) void k_fused(
    double* A, int lda, int r_lo, int m, int c0, int n, StepBufs bufs, double* Wp, int ns,
    int cur, const DevState* st) {
  constexpr int PF = FuShape<U>::PF, NSET = PF + 1;
  constexpr int WC = 16 * U, GC = FU_NWG * WC;
  __shared__ __align__(16) double VpS[2][FU_CH][FU_VP_LD];
  __shared__ __align__(32) double VnS[2][64][FU_VN_LD];
  if (st) {
    if (dm_stopped(st)) return;
    r_lo = st->j;
    c0 = st->c0;
    cur = st->it & 1;
  }
  const int ldv = bufs.ldv;
  const double* Vn = bufs.V + (size_t)cur * bufs.vstride;
  const double* Vp = bufs.V + (size_t)(cur ^ 1) * bufs.vstride;
  const double* Wprev = bufs.W + (size_t)(cur ^ 1) * bufs.wstride;
  const FuGeom G = fused_geom(r_lo, m, c0, n, ns, GC);
  const int per = fused_per_xcd(ns, G.cgroups);
  const int xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
  if (k >= per) return;
  const int slab = k % ns, cgrp = xcd + 8 * (k / ns);
  if (cgrp >= G.cgroups) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, li = lane & 15;
  const int ncol = n - c0;
  const int cbw = cgrp * GC + wave * WC;  // this wave's columns (relative to c0)
  const int ch0 = slab * G.slab_ch;
  const int ch1 = min(ch0 + G.slab_ch, G.nchunk);
  // the wave's columns c0 + cbw + 16u + li (past ncol: a valid column, never stored)
  double* Cc[U];
  bool cok[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int c = cbw + 16 * u + li;
    cok[u] = c < ncol;
    Cc[u] = A + (size_t)(c0 + (cok[u] ? c : (ncol > 0 ? ncol - 1 : 0))) * lda;
  }
  // the previous step's W tile: wr[s][u] = W(16g + s, column 16u + li) (lane
  // group g takes the k indices 16g..16g+15 of product 1)
  double wr[16][U];
  if (APPLY) {
#pragma unroll
    for (int s = 0; s < 16; s++)
#pragma unroll
      for (int u = 0; u < U; u++)
        wr[s][u] = cok[u] ? Wprev[(size_t)(c0 + cbw + 16 * u + li) * KMAX + 16 * g + s] : 0.0;
  }
  dbl4 acc[4][U];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int u = 0; u < U; u++) acc[t][u] = (dbl4){0.0, 0.0, 0.0, 0.0};
  // last 4-row group inside the column (rows up to round_up(m, 4) - 1 exist)
  const int rclamp = ((m + 3) & ~3) - 4;
  // V chunk loaders (a thread brings rows vq..vq+3 of column vc).  V_s: wave
  // w takes rows 8 (w & 1)..+7 of columns 32 (w >> 1)..+31, a lane pair one
  // column; V_{s+1}: four lanes one column, a wave 16 columns.  Both make the
  // LDS stores of a wave conflict-free in the layouts above.
  const int pq = 4 * (2 * (wave & 1) + (lane & 1)), pc = 32 * (wave >> 1) + (lane >> 1);
  const int nq = 4 * (tid & 3), nc = tid >> 2;
  auto load_vv = [&](const double* Vb, int c, int vq, int vc) -> dbl4 {
    const int r = G.base + FU_CH * c + vq;
    return *reinterpret_cast<const dbl4*>(Vb + (size_t)vc * ldv + r);
  };
  // rows outside [r_lo, m) enter the products as zero
  auto mask4 = [&](int c, int vq, dbl4 v) -> dbl4 {
    const int r = G.base + FU_CH * c + vq;
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = (r + q >= r_lo && r + q < m) ? v[q] : 0.0;
    return v;
  };
  auto store_vp = [&](int buf, int c, dbl4 v) {
    v = mask4(c, pq, v);
#pragma unroll
    for (int q = 0; q < 4; q++) VpS[buf][pq + q][pc] = -v[q];
  };
  auto store_vn = [&](int buf, int c, dbl4 v) {
    v = mask4(c, nq, v);
#pragma unroll
    for (int q = 0; q < 4; q++) VnS[buf][nc][nq + q] = v[q];
  };
  auto load_c = [&](int c, dbl4 (&cc)[U]) {
    const int r = min(G.base + FU_CH * c + 4 * g, rclamp);
#pragma unroll
    for (int u = 0; u < U; u++) cc[u] = *reinterpret_cast<const dbl4*>(Cc[u] + r);
  };
  // product 1's A operand rows: chunk row pi(li) = 4 (li % 4) + li / 4, so
  // that output row g + 4v of the 16 x 16 tile is chunk row 4g + v
  const int prow = 4 * (li & 3) + (li >> 2);
  // One chunk: product 1 on cur (C of chunk c) and the store of its updated
  // rows, the loads of V for chunk c + 1 and of C for chunk c + PF (into pre,
  // the set stored a chunk earlier), product 2 on cur, the next V chunk into
  // LDS, and the barrier.
  auto step = [&](int c, dbl4 (&cur)[U], dbl4 (&pre)[U]) {
    const int buf = (c - ch0) & 1;
    const bool more = c + 1 < ch1;
    const int cn = more ? c + 1 : c;
    const int r4 = G.base + FU_CH * c + 4 * g;  // first row of this lane group
    const bool interior = G.base + FU_CH * c >= r_lo && G.base + FU_CH * (c + 1) <= m;
    if (APPLY) {
      // product 1: C -= Vp W on rows 4g + v of columns 16u + li
#pragma unroll
      for (int s = 0; s < 16; s += 2) {
        const double2 a = *reinterpret_cast<const double2*>(&VpS[buf][prow][16 * g + s]);
#pragma unroll
        for (int u = 0; u < U; u++)
          cur[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, wr[s][u], cur[u], 0, 0, 0);
#pragma unroll
        for (int u = 0; u < U; u++)
          cur[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, wr[s + 1][u], cur[u], 0, 0, 0);
      }
      if (interior) {
#pragma unroll
        for (int u = 0; u < U; u++)
          if (cok[u]) *reinterpret_cast<dbl4*>(Cc[u] + r4) = cur[u];
      } else {
#pragma unroll
        for (int u = 0; u < U; u++) {
          if (!cok[u]) continue;
#pragma unroll
          for (int q = 0; q < 4; q++)
            if (r4 + q >= r_lo && r4 + q < m) Cc[u][r4 + q] = cur[u][q];
        }
      }
    }
    dbl4 vpn, vnn;
    if (APPLY) vpn = load_vv(Vp, cn, pq, pc);
    if (ACCUM) vnn = load_vv(Vn, cn, nq, nc);
    if (c + PF < ch1) load_c(c + PF, pre);
    if (ACCUM) {
      if (!interior) {
        // rows outside [r_lo, m) (R rows above, padding below) enter as zero
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const bool ok = r4 + q >= r_lo && r4 + q < m;
#pragma unroll
          for (int u = 0; u < U; u++) cur[u][q] = ok ? cur[u][q] : 0.0;
        }
      }
      // product 2: Wp += Vn^T C, k-step v = rows 4g + v
#pragma unroll
      for (int v = 0; v < 4; v++) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const double a = VnS[buf][16 * t + li][4 * g + v];
#pragma unroll
          for (int u = 0; u < U; u++)
            acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, cur[u][v], acc[t][u], 0, 0, 0);
        }
      }
    }
    if (more) {
      if (APPLY) store_vp(buf ^ 1, cn, vpn);
      if (ACCUM) store_vn(buf ^ 1, cn, vnn);
    }
    __syncthreads();
  };
  dbl4 cs[NSET][U];
  if (ch1 > ch0) {
    if (APPLY) store_vp(0, ch0, load_vv(Vp, ch0, pq, pc));
    if (ACCUM) store_vn(0, ch0, load_vv(Vn, ch0, nq, nc));
#pragma unroll
    for (int p = 0; p < PF; p++)
      if (ch0 + p < ch1) load_c(ch0 + p, cs[p]);
  }
  __syncthreads();
#pragma clang loop unroll(disable)
  for (int c = ch0; c < ch1; c += NSET) {
#pragma unroll
    for (int q = 0; q < NSET; q++)
      if (c + q < ch1) step(c + q, cs[q], cs[(q + PF) % NSET]);
  }
  if (!ACCUM) return;
  // partial tile: row 16t + g + 4w, column 16u + li of the wave's columns
  double* out = Wp + (size_t)slab * KMAX * n;
#pragma unroll
  for (int u = 0; u < U; u++) {
    if (!cok[u]) continue;
    const int col = c0 + cbw + 16 * u + li;
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int w = 0; w < 4; w++) out[(size_t)(16 * t + g + 4 * w) * n + col] = acc[t][u][w];
  }
}

// The launch shape (16-column blocks per wave) of the fused passes.
int g_fused_u = 1;

static int fused_gc(int u) { return FU_NWG * 16 * u; }
// workgroups in one round: OCC waves per SIMD, four SIMDs per CU, 256 CUs
static int fused_wgs(int u) { return 256 * (u == 1 ? FuShape<1>::OCC : FuShape<2>::OCC); }

int fused_max_slabs() { return FU_MAX_SPLIT; }

// Row slabs of a fused pass: one round of workgroups on every XCD (its
// column groups times the slabs within an eighth of the chip's slots) where
// the column groups allow it.
int fused_slabs(int r_lo, int m, int c0, int n) {
  const FuGeom g = fused_geom(r_lo, m, c0, n, 1, fused_gc(g_fused_u));
  if (g.cgroups <= 0 || g.nchunk <= 0) return 1;
  int ns = (fused_wgs(g_fused_u) / 8) / ((g.cgroups + 7) / 8);
  if (ns > FU_MAX_SPLIT) ns = FU_MAX_SPLIT;
  if (ns > g.nchunk) ns = g.nchunk;
  return ns < 1 ? 1 : ns;
}

template <int U>
static void launch_fused_u(double* A, int lda, int r_lo, int m, int c0, int n,
                           const StepBufs& bufs, double* Wp, int ns, int apply, int accum,
                           int cur, hipStream_t s, const DevState* st) {
  const FuGeom g = fused_geom(r_lo, m, c0, n, ns, fused_gc(U));
  const dim3 grid(8 * fused_per_xcd(ns, g.cgroups)), block(64 * FU_NWG);
  if (apply && accum)
    hipLaunchKernelGGL((k_fused<true, true, U>), grid, block, 0, s, A, lda, r_lo, m, c0, n, bufs,
                       Wp, ns, cur, st);
  else if (accum)
    hipLaunchKernelGGL((k_fused<false, true, U>), grid, block, 0, s, A, lda, r_lo, m, c0, n,
                       bufs, Wp, ns, cur, st);
  else
    hipLaunchKernelGGL((k_fused<true, false, U>), grid, block, 0, s, A, lda, r_lo, m, c0, n,
                       bufs, Wp, ns, cur, st);
}

void launch_fused(double* A, int lda, int r_lo, int m, int c0, int n, const StepBufs& bufs,
                  double* Wp, int ns, int apply, int accum, int cur, hipStream_t s,
                  const DevState* st) {
  if (n - c0 <= 0 || m - r_lo <= 0 || !(apply || accum)) return;
  if (g_fused_u == 2)
    launch_fused_u<2>(A, lda, r_lo, m, c0, n, bufs, Wp, ns, apply, accum, cur, s, st);
  else
    launch_fused_u<1>(A, lda, r_lo, m, c0, n, bufs, Wp, ns, apply, accum, cur, s, st);
}

// ---------------------------------------------------------------------------
// The end of a fused step, one 64-column tile of [j + kk, n) per workgroup
// (256 threads): W = T^T sum_slab Wp (the fused pass's partials, summed in
// slab order; zero for the columns below c0, which the exact panel's dlarf_
// already updated), the step's R12 rows R12 = C_top - V_top W (rows j..j+kk,
// final, written), then the norm downdate of src/dgeqrdm_work.c:80-109 from
// them: s = sum of the kk squares (as the reference's loop, :84-86), temp =
// max(0, (1 + t)(1 - t)) with t = sqrt(|s|) / work, temp2 = temp (work /
// workold)^2; temp2 <= tol3z recomputes the norm from the rows below -- such a
// column receives the whole update now (rows j + kk .. m, C -= V W on the
// matrix cores) and its W column is zero, so the next fused pass leaves it --
// otherwise work *= sqrt(temp).  The running max into st (maxnrm).  W goes to
// the step's W buffer.
__global__ __launch_bounds__(256) void k_wsum_norm(double* A, int lda, int m, int n,
                                                   const double* __restrict__ T,
                                                   const double* __restrict__ Wp, int ns,
                                                   StepBufs bufs, double* work, double* workold,
                                                   DevState* st, double tol3z) {
  __shared__ double S0[KMAX][LDP];  // T; V's top rows; V chunks of the recompute
  __shared__ double S1[KMAX][LDP];  // the partials' sum, then W
  __shared__ double S2[KMAX][LDP];  // C's top rows, then R12; C chunks of the recompute
  __shared__ double s_ss[4][KMAX];  // per-wave sums of squares of the recomputed columns
  __shared__ double s_wk[KMAX], s_wo[KMAX];  // work, workold of the tile's columns
  __shared__ unsigned long long s_rows[4];
  __shared__ int s_rec[KMAX];
  __shared__ int s_any;
  if (dm_stopped(st)) return;
  // the step's end is pending from here on (run by the next k_topk or k_step_end)
  if (blockIdx.x == 0 && threadIdx.x == 0) st->pend = 1;
  const int j = st->j, kk = st->fjb_cmp, c0 = st->c0;
  const int jn = j + kk;
  const int cb = jn + blockIdx.x * KMAX;
  if (cb >= n) return;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, g = lane >> 4, li = lane & 15;
  const double* V = v_cur(bufs, st);
  double* W = w_cur(bufs, st);
  const int ldv = bufs.ldv;
  // every global read of the prologue in flight together: T, the norms, the
  // top rows of C and V, and the slab partials (two slabs per round)
  if (tid < KMAX) {
    const int col = cb + tid;
    s_wk[tid] = col < n ? work[col] : 0.0;
    s_wo[tid] = col < n ? workold[col] : 1.0;
    s_rec[tid] = 0;
  }
  load_mat(S0, T);
  {
    const int c = tid & 63, i0 = tid >> 6, col = cb + c;
    const bool in = col >= c0 && col < n;
    const int cl = col < n ? col : n - 1;
    double sum[16];
#pragma unroll
    for (int q = 0; q < 16; q++) sum[q] = 0.0;
    for (int y = 0; y < ns; y += 2) {
      const int y2 = (y + 1 < ns) ? y + 1 : y;
      double a0[16], a1[16];
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int i = i0 + 4 * q;
        a0[q] = Wp[((size_t)y * KMAX + i) * n + cl];
        a1[q] = Wp[((size_t)y2 * KMAX + i) * n + cl];
      }
#pragma unroll
      for (int q = 0; q < 16; q++) {
        sum[q] += a0[q];
        if (y + 1 < ns) sum[q] += a1[q];
      }
    }
#pragma unroll
    for (int q = 0; q < 16; q++) S1[i0 + 4 * q][c] = in ? sum[q] : 0.0;
  }
  lds_barrier();
  dbl4 acc[4];
  mfma_64(acc, S0, true, S1);  // T^T S1
  // V's top rows and C's top rows (rows j..j+kk) while the product runs
  double vt[KMAX * KMAX / 256], ct[KMAX * KMAX / 256];
#pragma unroll
  for (int q = 0; q < KMAX * KMAX / 256; q++) {
    const int e = tid + 256 * q, r = e & 63, cc = e >> 6;
    vt[q] = (r < kk) ? V[(size_t)cc * ldv + j + r] : 0.0;  // V(j + r, cc)
    ct[q] = (r < kk && cb + cc < n) ? A[(size_t)(cb + cc) * lda + j + r] : 0.0;
  }
  lds_barrier();
  mfma_store(S1, acc);  // S1 = W
#pragma unroll
  for (int q = 0; q < KMAX * KMAX / 256; q++) {
    const int e = tid + 256 * q, r = e & 63, cc = e >> 6;
    S0[r][cc] = vt[q];
    S2[r][cc] = ct[q];
  }
  lds_barrier();
  mfma_64(acc, S0, false, S1);  // V_top W
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++) S2[16 * wv + g + 4 * v][16 * u + li] -= acc[u][v];
  lds_barrier();
  for (int q = 0; q < KMAX * KMAX / 256; q++) {
    const int e = tid + 256 * q, r = e & 63, cc = e >> 6;
    if (r < kk && cb + cc < n) A[(size_t)(cb + cc) * lda + j + r] = S2[r][cc];
  }
  // norm downdate: the sum of squares of column c's kk R12 entries as four
  // partial sums (rows p, p + 4, ...; thread p * 64 + c) added in a fixed order
  {
    const int c = tid & 63, p = tid >> 6;
    double sq = 0.0;
    for (int r = p; r < kk; r += 4) sq = fma(S2[r][c], S2[r][c], sq);
    s_ss[p][c] = sq;
  }
  if (tid == 0) s_any = 0;
  lds_barrier();
  double mx = 0.0;
  if (tid < KMAX) {
    const int c = tid, col = cb + c;
    const double wk = s_wk[c];
    if (col < n && wk != 0.0) {
      const double sq = (s_ss[0][c] + s_ss[1][c]) + (s_ss[2][c] + s_ss[3][c]);
      double temp = sqrt(fabs(sq)) / wk;
      const double d2 = (temp + 1.0) * (1.0 - temp);
      temp = (0.0 >= d2) ? 0.0 : d2;
      const double d1 = wk / s_wo[c];
      const double temp2 = temp * (d1 * d1);
      if (temp2 <= tol3z) {
        s_rec[c] = 1;
        s_any = 1;
      } else {
        const double nw = wk * sqrt(temp);
        s_wk[c] = nw;
        if (nw == nw) mx = nw;
      }
    }
  }
  lds_barrier();
  unsigned long long rows = 0;
  if (s_any) {
    // the recomputed columns take the update on rows [jn, m) now.  Each wave
    // runs its own 16-row blocks (rows jn + 16 wv, then every 64th): D =
    // V_blk W on the matrix cores, V read straight into the A operand, W from
    // LDS, no workgroup barrier inside the loop.  Lane (g, li) keeps the sums
    // of squares of columns 16 u + li over its rows g + 4 v of each block.
    bool fl[4];
#pragma unroll
    for (int u = 0; u < 4; u++) fl[u] = s_rec[16 * u + li] != 0;
    double ps[4] = {0.0, 0.0, 0.0, 0.0};
    for (int rb = jn + 16 * wv; rb < m; rb += 4 * 16) {
      const int ra = rb + li;  // the lane's row of the A operand
      double va[KMAX / 4];
#pragma unroll
      for (int t = 0; t < KMAX / 4; t++)
        va[t] = (ra < m) ? V[(size_t)(4 * t + g) * ldv + ra] : 0.0;  // V(ra, 4t + g)
      double cv[4][4];
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int r = rb + g + 4 * v;
          cv[u][v] = (fl[u] && r < m) ? A[(size_t)(cb + 16 * u + li) * lda + r] : 0.0;
        }
      dbl4 d[4];
#pragma unroll
      for (int u = 0; u < 4; u++) d[u] = (dbl4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < KMAX / 4; t++)
#pragma unroll
        for (int u = 0; u < 4; u++)
          d[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[t], S1[4 * t + g][16 * u + li], d[u], 0,
                                                      0, 0);
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int r = rb + g + 4 * v;
          if (fl[u] && r < m) {
            const double x = cv[u][v] - d[u][v];
            A[(size_t)(cb + 16 * u + li) * lda + r] = x;
            ps[u] = fma(x, x, ps[u]);
          }
        }
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      double x = ps[u];
      x += __shfl_xor(x, 16, 64);
      x += __shfl_xor(x, 32, 64);
      if (g == 0) s_ss[wv][16 * u + li] = x;
    }
    lds_barrier();
    // norms of the recomputed columns (the scaled form where the plain sum
    // of squares leaves the safe range, as dnrm2's scaling does)
    const int mrem = m - jn;
    for (int q = 0; q < 16; q++) {
      const int c = 16 * wv + q, col = cb + c;
      if (!s_rec[c] || col >= n) continue;
      const double sq = (s_ss[0][c] + s_ss[1][c]) + (s_ss[2][c] + s_ss[3][c]);
      double nr;
      if (mrem <= 0) nr = 0.0;
      else if (sq >= 1e-280 && sq <= 1e280) nr = sqrt(sq);
      else nr = wave_nrm2(A + (size_t)col * lda + jn, mrem);
      if (lane == 0) {
        s_wk[c] = nr;
        workold[col] = nr;
      }
      if (nr == nr) mx = fmax(mx, nr);
      rows += (unsigned long long)(mrem > 0 ? mrem : 0);
    }
  }
  lds_barrier();
  // the tile's norms, and its W (zero for the recomputed columns), column-major
  if (tid < KMAX && cb + tid < n) work[cb + tid] = s_wk[tid];
  for (int q = 0; q < KMAX * KMAX / 256; q++) {
    const int e = tid + 256 * q, i = e & 63, c = e >> 6;
    if (cb + c < n) W[(size_t)(cb + c) * KMAX + i] = s_rec[c] ? 0.0 : S1[i][c];
  }
  if (lane == 0) s_rows[wv] = rows;
  // (a lane of wave 0 holds its own column's norm: the wave's maximum goes in)
  block_max_atomic(wave_max(mx), 0, st);
  if (tid == 0) {
    const unsigned long long r = s_rows[0] + s_rows[1] + s_rows[2] + s_rows[3];
    if (r) atomicAdd(&st->recomp_rows, r);
  }
}

void launch_wsum_norm(double* A, int lda, int j0, int m, int n, const double* T,
                      const double* Wp, int ns, const StepBufs& bufs, double* work,
                      double* workold, DevState* st, double tol3z, hipStream_t s) {
  // tiles of [j + kk, n), j + kk >= j0 + 1; one workgroup at least (it marks
  // the step's end pending)
  int tiles = (n - j0 - 1 + KMAX - 1) / KMAX;
  if (tiles < 1) tiles = 1;
  hipLaunchKernelGGL(k_wsum_norm, dim3(tiles), dim3(256), 0, s, A, lda, m, n, T, Wp, ns, bufs,
                     work, workold, st, tol3z);
}

}  // namespace qrdm
