// kernels.hip -- CDNA4 (gfx950) kernels of the DM-QR library.
//
// Wave64 throughout.  FP64 arithmetic; the trailing update runs on the FP64
// matrix cores (v_mfma_f64_16x16x4f64).  Each kernel names the reference code
// whose arithmetic it performs.
#include "qrdm_internal.h"

#include <float.h>
#include <math.h>

namespace qrdm {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// LAPACK dlamch('S') / dlamch('E'): below this |beta| dlarfg rescales.
__device__ constexpr double kSafmin = 2.2250738585072014e-308 / 1.1102230246251565e-16;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Non-negative doubles order like their bit patterns.
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* p, double v) {
  atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

// Two-norm of x[0..n) by one wave; every lane returns it.  Plain sum of
// squares, redone with scaling by max|x| when the sum is out of the range
// where squaring is exact enough (the role of dnrm2's scaling).
__device__ __forceinline__ double wave_nrm2(const double* __restrict__ x, int n) {
  const int lane = local_tid() & 63;
  // four independent loads and partial sums per lane per iteration, so the
  // loads of one column are in flight together
  double s0 = 0.0, s1 = 0.0, s2a = 0.0, s3 = 0.0;
  int i = lane;
  for (; i + 192 < n; i += 256) {
    const double v0 = x[i], v1 = x[i + 64], v2 = x[i + 128], v3 = x[i + 192];
    s0 = fma(v0, v0, s0);
    s1 = fma(v1, v1, s1);
    s2a = fma(v2, v2, s2a);
    s3 = fma(v3, v3, s3);
  }
  for (; i < n; i += 64) {
    const double v = x[i];
    s0 = fma(v, v, s0);
  }
  double s = wave_sum((s0 + s1) + (s2a + s3));
  if (isnan(s)) return s;
  if (s >= 1e-280 && s <= 1e280) return sqrt(s);
  double amax = 0.0;
  for (int i = lane; i < n; i += 64) amax = fmax(amax, fabs(x[i]));
  amax = wave_max(amax);
  if (amax == 0.0 || isinf(amax)) return amax;
  double r = 1.0 / amax, s2 = 0.0;
  for (int i = lane; i < n; i += 64) {
    double v = x[i] * r;
    s2 = fma(v, v, s2);
  }
  return amax * sqrt(wave_sum(s2));
}

// Block-wide sum; red must hold blockDim.x/64 doubles.  All threads return it.
__device__ double block_sum(double v, double* red) {
  const int lane = local_tid() & 63, wave = local_tid() >> 6, nw = blockDim.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < nw; w++) t += red[w];
  return t;
}

__device__ __forceinline__ double block_max(double v, double* red) {
  const int lane = local_tid() & 63, wave = local_tid() >> 6, nw = blockDim.x >> 6;
  v = wave_max(v);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < nw; w++) t = fmax(t, red[w]);
  return t;
}

// Block-wide two-norm of x[0..n) (scaled fallback as in wave_nrm2).
__device__ double block_nrm2(const double* __restrict__ x, int n, double* red) {
  const int nt = blockDim.x;
  double s0 = 0.0, s1 = 0.0;
  int i = local_tid();
  for (; i + nt < n; i += 2 * nt) {
    const double v0 = x[i], v1 = x[i + nt];
    s0 = fma(v0, v0, s0);
    s1 = fma(v1, v1, s1);
  }
  for (; i < n; i += nt) {
    const double v = x[i];
    s0 = fma(v, v, s0);
  }
  double s = block_sum(s0 + s1, red);
  if (isnan(s)) return s;
  if (s >= 1e-280 && s <= 1e280) return sqrt(s);
  double amax = 0.0;
  for (int i = local_tid(); i < n; i += blockDim.x) amax = fmax(amax, fabs(x[i]));
  amax = block_max(amax, red);
  if (amax == 0.0 || isinf(amax)) return amax;
  double r = 1.0 / amax, s2 = 0.0;
  for (int i = local_tid(); i < n; i += blockDim.x) {
    double v = x[i] * r;
    s2 = fma(v, v, s2);
  }
  return amax * sqrt(block_sum(s2, red));
}

// dlapy2 (LAPACK): sqrt(x^2 + y^2) without needless overflow.
__device__ __forceinline__ double dlapy2(double x, double y) {
  double xa = fabs(x), ya = fabs(y);
  double w = fmax(xa, ya), z = fmin(xa, ya);
  if (z == 0.0 || w > DBL_MAX) return w;
  double q = z / w;
  return w * sqrt(1.0 + q * q);
}

// f2c d_sign(a, b) (src/dlarfg.c:21-26): the test is b >= 0, so -0.0 is '+'.
__device__ __forceinline__ double dsign(double a, double b) {
  double x = fabs(a);
  return b >= 0 ? x : -x;
}

// ---------------------------------------------------------------------------
// Initial partial norms: work[c] = workold[c] = ||A(r0:m, c)||, max over all
// columns (src/dgeqrdm_work.c:670-682).  One wave per column.
// One wave per column, NORM_COLS columns per wave; the maximum is reduced in
// the workgroup first so that only one atomic per workgroup reaches the
// single maxnrm word (same-address atomics serialise).
constexpr int NORM_COLS = 4;

__device__ __forceinline__ void block_max_atomic(double v, int bad, DevState* st) {
  __shared__ double wmax[4];
  __shared__ int wbad[4];
  const int lane = local_tid() & 63, wave = local_tid() >> 6;
  if (lane == 0) {
    wmax[wave] = v;
    wbad[wave] = bad;
  }
  __syncthreads();
  if (local_tid() == 0) {
    const double mx = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    const int b = wbad[0] | wbad[1] | wbad[2] | wbad[3];
    if (b) atomicOr(&st->flags, 1);
    if (mx > 0.0) atomic_max_nonneg(&st->maxnrm_bits, mx);
  }
}

__global__ __launch_bounds__(256) void k_init_norms(const double* __restrict__ A, int lda,
                                                     int r0, int m, int n, double* work,
                                                     double* workold, DevState* st) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double mx = 0.0;
  int bad = 0;
  for (int q = 0; q < NORM_COLS; q++) {
    const int col = (blockIdx.x * 4 + wave) * NORM_COLS + q;
    if (col >= n) break;
    const double nr = wave_nrm2(A + (size_t)col * lda + r0, m - r0);
    if (lane == 0) {
      work[col] = nr;
      workold[col] = nr;
    }
    if (!(nr <= DBL_MAX)) bad = 1;
    else mx = fmax(mx, nr);
  }
  block_max_atomic(mx, bad, st);
}

void launch_init_norms(const double* A, int lda, int r0, int m, int n, double* work,
                       double* workold, DevState* st, hipStream_t s) {
  const int waves = (n + NORM_COLS - 1) / NORM_COLS;
  hipLaunchKernelGGL(k_init_norms, dim3((waves + 3) / 4), dim3(256), 0, s, A, lda, r0, m, n,
                     work, workold, st);
}

// ---------------------------------------------------------------------------
// Top-k of the partial norms in the order of the reference's qsort with
// cmpStruct (src/dgeqrdm_work.c:126-142, :345): value descending, and, as
// glibc's merge sort keeps equal keys in place, position ascending on ties.

__device__ __forceinline__ bool key_before(double va, int ia, double vb, int ib) {
  return va > vb || (va == vb && ia < ib);
}

// Value of x in lane (lane ^ JJ) of the wave, without LDS: DPP quad and row
// permutations for JJ < 16, the gfx950 row-swap instructions above.
template <int JJ>
__device__ __forceinline__ int xor_lane(int x) {
  const int lane = local_tid() & 63;
  if constexpr (JJ == 1) {
    return __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
  } else if constexpr (JJ == 2) {
    return __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
  } else if constexpr (JJ == 4) {
    const int a = __builtin_amdgcn_update_dpp(0, x, 0x124, 0xF, 0xF, false);  // row_ror:4
    const int b = __builtin_amdgcn_update_dpp(0, x, 0x12C, 0xF, 0xF, false);  // row_ror:12
    return (lane & 4) ? a : b;
  } else if constexpr (JJ == 8) {
    return __builtin_amdgcn_update_dpp(0, x, 0x128, 0xF, 0xF, false);  // row_ror:8
  } else if constexpr (JJ == 16) {
    // rows of 16: the first result holds rows (0, 0, 2, 2), the second (1, 1, 3, 3)
    const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    return (lane & 16) ? (int)r[0] : (int)r[1];
  } else {
    static_assert(JJ == 32, "xor distances 1..32");
    const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return (lane & 32) ? (int)r[0] : (int)r[1];
  }
}

template <int JJ>
__device__ __forceinline__ double xor_lane(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = xor_lane<JJ>((int)b), hi = xor_lane<JJ>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// One compare-exchange stage across lanes lane and lane ^ JJ: the lane keeps
// the key that comes first (want_first) or second of its pair.  Keys are
// distinct (padding keys carry distinct indices), so one comparison decides.
template <int JJ>
__device__ __forceinline__ void lane_cx(double& v, int& ix, bool want_first) {
  const double pv = xor_lane<JJ>(v);
  const int pi = xor_lane<JJ>(ix);
  if (want_first == key_before(pv, pi, v, ix)) {
    v = pv;
    ix = pi;
  }
}

// Sorts the wave's 64 keys (one per lane) into key_before order: the bitonic
// network, blocks of K keys merged with distances K/2 .. 1.
template <int K, int JJ>
__device__ __forceinline__ void sort_stages(double& v, int& ix) {
  const int lane = local_tid() & 63;
  const bool first = (lane & JJ) == 0;       // lower position of the pair
  const bool desc = (lane & (K & 63)) == 0;  // this K-block sorts descending
  lane_cx<JJ>(v, ix, first == desc);
  if constexpr (JJ > 1) sort_stages<K, JJ / 2>(v, ix);
}

template <int K = 2>
__device__ __forceinline__ void wave_sort64(double& v, int& ix) {
  sort_stages<K, K / 2>(v, ix);
  if constexpr (K < 64) wave_sort64<2 * K>(v, ix);
}

// (v, ix) := the first 64 keys of two sorted 64-key lists (v, ix) and
// (bv, bi), sorted: the lane-wise better of a[lane] and b[63 - lane] is a
// bitonic sequence holding them, which six half-cleaner stages sort.
__device__ __forceinline__ void wave_merge64(double& v, int& ix, double bv, int bi) {
  const int lane = local_tid() & 63;
  // b[63 - lane]: lane ^ 63 through the six distances
  bv = xor_lane<32>(bv);
  bi = xor_lane<32>(bi);
  bv = xor_lane<16>(bv);
  bi = xor_lane<16>(bi);
  bv = xor_lane<8>(bv);
  bi = xor_lane<8>(bi);
  bv = xor_lane<4>(bv);
  bi = xor_lane<4>(bi);
  bv = xor_lane<2>(bv);
  bi = xor_lane<2>(bi);
  bv = xor_lane<1>(bv);
  bi = xor_lane<1>(bi);
  if (key_before(bv, bi, v, ix)) {
    v = bv;
    ix = bi;
  }
  lane_cx<32>(v, ix, (lane & 32) == 0);
  lane_cx<16>(v, ix, (lane & 16) == 0);
  lane_cx<8>(v, ix, (lane & 8) == 0);
  lane_cx<4>(v, ix, (lane & 4) == 0);
  lane_cx<2>(v, ix, (lane & 2) == 0);
  lane_cx<1>(v, ix, (lane & 1) == 0);
}

// Merges the wave lists lv[0..nw) (sorted, 64 keys each) into lv[0] by a
// tree over the waves; every wave of the workgroup calls it.
template <int NW>
__device__ __forceinline__ void block_merge_lists(double (*lv)[KMAX], int (*li)[KMAX], double& v,
                                                  int& ix) {
  const int lane = local_tid() & 63, wave = local_tid() >> 6;
#pragma unroll
  for (int h = NW / 2; h >= 1; h >>= 1) {
    // wave w < h merges lists w and w + h into list w
    if (wave < h) {
      wave_merge64(v, ix, lv[wave + h][lane], li[wave + h][lane]);
      lv[wave][lane] = v;
      li[wave][lane] = ix;
    }
    __syncthreads();
  }
}

// Whether the loop ends after the current step (src/dgeqrdm_work.c:782-785):
// the last column reached, or the stop test met.
__device__ __forceinline__ int step_done(const DevState* st, int m, int n, int enable_stop,
                                         double eta) {
  const int j = st->j, kk = st->fjb_cmp;
  const double maxnrm = __longlong_as_double((long long)st->maxnrm_bits);
  const int minmn = m < n ? m : n;
  int done = (j + kk >= minmn) ? 1 : 0;
  if (enable_stop && maxnrm * sqrt((double)(n - j - kk)) <= eta) done = 1;
  return done;
}

// End of a DM step (run once, by one thread): ncols, the step record and
// work, the stop test (enable_stop, eta absolute), the next j, done.
__device__ void step_end_apply(DevState* st, int* ncols_dev, int m, int n, int enable_stop,
                               double eta) {
  st->pend = 0;
  if (st->done) return;
  const int j = st->j, kk = st->fjb_cmp, fjb = st->fjb;
  double maxnrm;
  {
    const unsigned long long b = st->maxnrm_bits;
    maxnrm = __longlong_as_double((long long)b);
  }
  const int msub = m - j, nsub = n - j;
  StepRec& r = st->last;
  r.j = j;
  r.kmax = st->kmax;
  r.nc = st->nc;
  r.fjb = fjb;
  r.fjb_cmp = kk;
  r.nswap = st->nswap;
  r.panel = (st->fast && !(st->flags & 2)) ? 1 : 0;
  r.maxnrm = maxnrm;
  r.tau = st->tau_step;
  r.delta = st->delta_step;
  const double nct = (double)(n - st->c0);
  StepWork& w = st->work;
  w.select += 8.0 * nsub + 8.0 * (double)msub * st->kmax;
  w.perm += 16.0 * (double)m * st->npair;
  w.panel += 2.0 * msub * (double)fjb * fjb;
  if (st->fused) {
    // the step's fused pass: the previous step's update and this step's
    // V^T C on rows j..m of columns c0..n; W = T^T sum and the R12 rows
    w.gemm1 += 2.0 * msub * nct * (double)(kk + st->kk_prev);
    w.wsum += 4.0 * kk * (double)kk * nct;
  } else {
    w.gemm1 += 2.0 * msub * nct * kk;
    w.wsum += 2.0 * kk * (double)kk * nct;
    w.gemm2 += 16.0 * msub * nct + 8.0 * msub * kk + 8.0 * kk * nct;
    w.norm += 8.0 * (double)(kk + 2) * (nsub - kk);
  }
  st->kk_prev = kk;
  const int done = step_done(st, m, n, enable_stop, eta);  // from j before the advance
  ncols_dev[st->it] = kk;
  st->it += 1;
  st->j = j + kk;
  st->done = done;
  st->maxnrm_bits = 0ull;
  st->flags &= ~2;
}

// The step's candidates: the sorted first kmax keys of work[j..n) (value
// descending, position ascending) and the candidate count (> tau * max,
// src/dgeqrdm_work.c:348-351), after the previous step's end when that is
// pending.  Each workgroup sorts TOPK_BLOCK keys (a 64-key list per wave,
// merged over the waves) into its first 64, written to tmp; the last
// workgroup to finish (atomic ticket) merges the lists and writes the
// results and the pending step end.  Every workgroup reads the state before
// taking its ticket and only the last one writes it.
constexpr int TOPK_WAVES = 8;
constexpr int TOPK_BLOCK = 64 * TOPK_WAVES;

__global__ __launch_bounds__(64 * TOPK_WAVES) void k_topk(const double* __restrict__ work,
                                                          int m_all, int n_all, int nb,
                                                          double tau, double* tmp_val,
                                                          int* tmp_idx, unsigned* ticket,
                                                          DevState* st, int* ncols_dev,
                                                          int enable_stop, double eta) {
  __shared__ double lv[TOPK_WAVES][KMAX];
  __shared__ int li[TOPK_WAVES][KMAX];
  __shared__ int s_j, s_done, s_end, s_last;
  if (threadIdx.x == 0) {
    int j = st->j, done = st->done;
    if (!done && st->pend) {  // the previous step's end, as step_end_apply computes it
      done = step_done(st, m_all, n_all, enable_stop, eta);
      j += st->fjb_cmp;
    }
    s_j = j;
    s_done = dm_stopped(st);  // loop over or halted before this launch: nothing to do
    // a halted step's successors stop before their column moves as well
    if (blockIdx.x == 0 && st->halt == 2) st->halt = 1;
    s_end = done;       // the loop ends with the pending step: no sort
  }
  __syncthreads();
  if (s_done) return;
  const int j = s_j;
  const bool sort = !s_end;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nsub = n_all - j;
  const int nblk = gridDim.x;
  double v = -1.0;
  int ix = 0;
  if (sort) {
    const int t = blockIdx.x * TOPK_BLOCK + 64 * wave + lane;
    // padding keys: -1 (norms are >= 0) with distinct indices past any column
    v = (t < nsub) ? work[j + t] : -1.0;
    ix = (t < nsub) ? t : 0x40000000 + t;
    wave_sort64(v, ix);
    lv[wave][lane] = v;
    li[wave][lane] = ix;
    __syncthreads();
    block_merge_lists<TOPK_WAVES>(lv, li, v, ix);
    if (wave == 0) {
      st_dev(tmp_val + blockIdx.x * KMAX + lane, v);
      st_dev(tmp_idx + blockIdx.x * KMAX + lane, ix);
    }
  }
  handoff_wait();
  __syncthreads();
  if (threadIdx.x == 0) s_last = last_to_arrive(ticket, nblk) ? 1 : 0;
  __syncthreads();
  if (!s_last) return;
  if (sort) {
    // wave w merges lists w, w + TOPK_WAVES, ... into its own
    v = -1.0;  // a sorted list of padding keys
    ix = 0x7fffffc0 + lane;
    // two lists per round, both loads in flight before the merges
    for (int b = wave; b < nblk; b += 2 * TOPK_WAVES) {
      const int b2 = b + TOPK_WAVES;
      const double v1 = ld_dev(tmp_val + b * KMAX + lane);
      const int i1 = ld_dev(tmp_idx + b * KMAX + lane);
      const double v2 = (b2 < nblk) ? ld_dev(tmp_val + b2 * KMAX + lane) : -1.0;
      const int i2 = (b2 < nblk) ? ld_dev(tmp_idx + b2 * KMAX + lane) : 0x7fffff80 + lane;
      wave_merge64(v, ix, v1, i1);
      if (b2 < nblk) wave_merge64(v, ix, v2, i2);
    }
    lv[wave][lane] = v;
    li[wave][lane] = ix;
    __syncthreads();
    block_merge_lists<TOPK_WAVES>(lv, li, v, ix);
  }
  if (threadIdx.x == 0 && st->pend) step_end_apply(st, ncols_dev, m_all, n_all, enable_stop, eta);
  if (sort && wave == 0) {
    // k_max = min(nb, m_sub, n_sub) (src/dgeqrdm_work.c:703), at most KMAX here
    int kmax = nb < KMAX ? nb : KMAX;
    if (m_all - j < kmax) kmax = m_all - j;
    if (nsub < kmax) kmax = nsub;
    const int nk = nsub < kmax ? nsub : kmax;
    st->cand_val[lane] = lane < nk ? v : 0.0;
    st->cand_idx[lane] = lane < nk ? ix : -1;
    // the keys are sorted: those above tau * max form a prefix
    const double lim = tau * __shfl(v, 0, 64);
    const int nc = __popcll(__ballot(lane < nk && v > lim));
    if (lane == 0) {
      st->nc = nc;
      st->kmax = kmax;
    }
  }
}

int topk_blocks(int n) { return (n + TOPK_BLOCK - 1) / TOPK_BLOCK; }

void launch_topk(const double* work, int j0, int m, int n, int nb, double tau, double* tmp_val,
                 int* tmp_idx, unsigned* ticket, DevState* st, int* ncols_dev, int enable_stop,
                 double eta, hipStream_t s) {
  const int nblk = topk_blocks(n - j0);
  hipLaunchKernelGGL(k_topk, dim3(nblk > 0 ? nblk : 1), dim3(64 * TOPK_WAVES), 0, s, work, m, n,
                     nb, tau, tmp_val, tmp_idx, ticket, st, ncols_dev, enable_stop, eta);
}

// ---------------------------------------------------------------------------
// out = X^T X (64 x 64, column-major) for the 64-row slab X held in LDS as
// X[c * (GRAM_ROWS + 1) + r]; 256 threads, wave w computes rows 16w..16w+15
// on the matrix cores.
__device__ __forceinline__ void gram_slab_acc(const double* X, dbl4 (&acc)[4]) {
  const int lane = local_tid() & 63, wv = local_tid() >> 6;
  const int g = lane >> 4, li = lane & 15;
#pragma unroll 4
  for (int k0 = 0; k0 < GRAM_ROWS; k0 += 4) {
    const double a = X[(16 * wv + li) * (GRAM_ROWS + 1) + k0 + g];
#pragma unroll
    for (int u = 0; u < 4; u++)
      acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(
          a, X[(16 * u + li) * (GRAM_ROWS + 1) + k0 + g], acc[u], 0, 0, 0);
  }
}

__device__ __forceinline__ void gram_store(const dbl4 (&acc)[4], double* out) {
  const int lane = local_tid() & 63, wv = local_tid() >> 6;
  const int g = lane >> 4, li = lane & 15;
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++) out[(16 * wv + g + 4 * v) + (16 * u + li) * KMAX] = acc[u][v];
}

// One 16-row block's contribution to a wave's Gram X^T X (64 x 64): xb[u]
// holds X_blk(:, 16u .. 16u + 15) in the matrix cores' output layout (rows
// g + 4v, column 16u + li for lane 16g + li), which is also the operand layout
// of a product over those rows, so the tiles u <= u2 accumulate with no data
// movement.
__device__ __forceinline__ void gram_tiles(const dbl4 (&xb)[4], dbl4 (&gt)[10]) {
  int p = 0;
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int u2 = u; u2 < 4; u2++) {
#pragma unroll
      for (int v = 0; v < 4; v++)
        gt[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(xb[u][v], xb[u2][v], gt[p], 0, 0, 0);
      p++;
    }
}

__device__ __forceinline__ void gram_slab_mfma(const double* X, double* out) {
  dbl4 acc[4];
#pragma unroll
  for (int u = 0; u < 4; u++) acc[u] = (dbl4){0.0, 0.0, 0.0, 0.0};
  gram_slab_acc(X, acc);
  gram_store(acc, out);
}

// ---------------------------------------------------------------------------
// Gram matrix of up to 64 columns over a row range: partial sums per
// GRAM_ROWS-row slab, then a fixed-order reduction.  Used for the candidate
// cosines (cblas_dsyrk, src/dgeqrdm_work.c:376-379) and for V^T V (dlarft).
__global__ __launch_bounds__(256) void k_gram_partial(const double* __restrict__ base, int ld,
                                                      int r0, int r1,
                                                      const int* __restrict__ cols,
                                                      int ncols_max,
                                                      const int* __restrict__ ncols_dev,
                                                      double* partial, const DevState* st,
                                                      int mode) {
  __shared__ double X[KMAX * (GRAM_ROWS + 1)];
  if (!step_runs(st, mode)) return;
  if (mode == RUN_DM) {  // the step's candidates: columns and rows from j
    base += (size_t)st->j * ld;
    r0 = st->j;
  } else if (mode != RUN_ALWAYS) {
    r0 = st->j;
  }
  const int nc = ncols_dev ? *ncols_dev : ncols_max;
  // slabs past r1 (the grid is sized for the lowest j) write zero partials
  const int rs = r0 + blockIdx.x * GRAM_ROWS;
  // 256 threads: 16 elements each, all loads issued before the LDS writes
  double xv[KMAX * GRAM_ROWS / 256];
#pragma unroll
  for (int q = 0; q < KMAX * GRAM_ROWS / 256; q++) {
    const int t = threadIdx.x + 256 * q;
    const int c = t / GRAM_ROWS, r = t % GRAM_ROWS;
    const bool ok = c < nc && rs + r < r1;
    const int cc = ok ? (cols ? cols[c] : c) : 0;
    xv[q] = ok ? base[(size_t)cc * ld + rs + r] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < KMAX * GRAM_ROWS / 256; q++) {
    const int t = threadIdx.x + 256 * q;
    X[(t / GRAM_ROWS) * (GRAM_ROWS + 1) + t % GRAM_ROWS] = xv[q];
  }
  __syncthreads();
  gram_slab_mfma(X, partial + (size_t)blockIdx.x * KMAX * KMAX);
}

__global__ __launch_bounds__(256) void k_gram_reduce(const double* __restrict__ partial,
                                                     int nslab, double* G,
                                                     const DevState* st, int mode) {
  __shared__ double part[4][64];
  if (!step_runs(st, mode)) return;
  const int e = blockIdx.x * 64 + (threadIdx.x & 63);
  // entries below the diagonal take the value of their mirror: G is exactly
  // symmetric, the upper triangle as dsyrk('U') forms it
  const int pr = e & 63, pc = e >> 6;
  const int es = (pr > pc) ? pc + pr * KMAX : e;
  const int q = threadIdx.x >> 6;
  // slabs q, q+4, ... summed eight loads at a time (index clamped, surplus masked)
  double s = 0.0;
  for (int b0 = q; b0 < nslab; b0 += 32) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int b = b0 + 4 * u;
      v[u] = partial[(size_t)(b < nslab ? b : q) * KMAX * KMAX + es];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) s += (b0 + 4 * u < nslab) ? v[u] : 0.0;
  }
  part[q][threadIdx.x & 63] = s;
  __syncthreads();
  if (q == 0) G[e] = (part[0][threadIdx.x] + part[1][threadIdx.x]) +
                     (part[2][threadIdx.x] + part[3][threadIdx.x]);
}

void launch_gram(const double* base, int ld, int r0, int r1, const int* cols, int ncols_max,
                 const int* ncols_dev, double* partial, double* G, hipStream_t s,
                 const DevState* st, int mode) {
  int nslab = (r1 - r0 + GRAM_ROWS - 1) / GRAM_ROWS;
  if (nslab < 1) nslab = 1;
  hipLaunchKernelGGL(k_gram_partial, dim3(nslab), dim3(256), 0, s, base, ld, r0, r1, cols,
                     ncols_max, ncols_dev, partial, st, mode);
  hipLaunchKernelGGL(k_gram_reduce, dim3(KMAX * KMAX / 64), dim3(256), 0, s, partial, nslab, G,
                     st, mode);
}

// ---------------------------------------------------------------------------
// The end of DM_perm (src/dgeqrdm_work.c:354-418), one wave:
// 1. greedy acceptance: the max-norm column first, then, in sorted order,
//    every candidate whose |cos| with all earlier accepts is < delta, while
//    fewer than k_max are accepted.  cos(p,q) = G(p,q) / (norm_p * norm_q)
//    with the downdated norms (the reference scales each candidate by 1/norm
//    before its dsyrk).
// 2. permute_marked's walk (src/dgeqrdm_work.c:149-262, marks 0/1): its
//    placement (the host walk permute_walk in driver.hip) in closed form.
// 3. the plan: (dst, src) pairs of the moved columns into perm[0..4*KMAX),
//    the candidate rank of each panel column into perm[4*KMAX..5*KMAX), and
//    the panel form (Gram-based when fjb >= 2 and m - j >= fast_rows).
// What select_wave reads of the step's state, lane t of the wave: read
// where the caller chooses (the reduce's workgroups read it before their
// ticket, so that it arrives with the partials).
struct SelIn {
  int j, nc, kmax, cidx;
  double delta, cval;
};
__device__ __forceinline__ SelIn select_inputs(const DevState* st, double delta) {
  const int t = local_tid() & 63;
  SelIn in;
  in.j = st->j;
  in.nc = st->nc;
  in.kmax = st->kmax;
  in.delta = st->policy_on ? st->delta_step : delta;
  in.cval = st->cand_val[t];
  in.cidx = st->cand_idx[t];
  return in;
}

// Run by one wave (threadIdx.x < 64); its LDS steps are ordered by
// wave_lds_sync, so the rest of the workgroup may be elsewhere.  Gs: the
// candidate Gram in LDS, column-major with leading dimension KMAX + 1.
__device__ __forceinline__ void select_wave(const double* Gs, const SelIn& in, int m, int n,
                                            int fast_rows, int exact_queued, int* perm,
                                            DevState* st) {
  __shared__ double s_r[KMAX];
  __shared__ int s_sel[KMAX], s_rank[KMAX];
  __shared__ int s_occ[KMAX], s_slot[KMAX];
  const int t = local_tid();
  const int j = in.j, nc = in.nc, kmax = in.kmax;
  const int nsub = n - j;
  const double delta = in.delta;
  // --- 1. greedy.  Lane t builds the mask of earlier candidates s whose
  // |cos(s, t)| >= delta; candidate i is accepted when its mask misses every
  // accepted candidate (max |cos| < delta) and fewer than k_max are accepted.
  unsigned long long acc = 1ull;  // bit i: candidate i accepted
  int nsel = 1;
  if (nc > 1) {
    const double rt = (t < nc) ? 1.0 / in.cval : 0.0;
    s_r[t] = rt;
    wave_lds_sync();
    unsigned long long conf = 0ull;
#pragma unroll 8
    for (int q = 0; q < KMAX; q++) {
      const double c = fabs(Gs[t + q * (KMAX + 1)] * s_r[q] * rt);  // G symmetric: G(q, t)
      conf |= (q < t && q < nc && c >= delta) ? (1ull << q) : 0ull;
    }
    for (int i = 1; i < nc; i++) {
      const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(conf >> 32), i);
      const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)conf, i);
      const unsigned long long ci = ((unsigned long long)hi << 32) | lo;
      if ((ci & acc) == 0ull && nsel < kmax) {
        acc |= 1ull << i;
        nsel++;
      }
    }
  }
  const int fjb = nsel;
  const bool mine = (acc >> t) & 1ull;
  const int spos = __popcll(acc & ((1ull << t) - 1ull));  // acceptance order
  const int cidx = in.cidx;
  if (mine) {
    s_sel[spos] = cidx;
    s_rank[spos] = t;
    st->sel[spos] = cidx;
    st->selpos[spos] = t;
  }
  wave_lds_sync();
  // --- 2. permute_marked's walk, in closed form.  With the accepted
  // positions S (|S| = fjb), jt = n_sub - 1 and the free front slots
  // F = [0, fjb) \ S in increasing order, the walk (host permute_walk)
  // amounts to:
  //  - when jt is in S (and jt > 0), its first pass swaps jt with position 0:
  //    if jt >= fjb the column A at jt goes to 0 and consumes F[0]; the one at
  //    0, if accepted (B), moves on to F[0]; the unmarked column from F[0]
  //    goes to jt.  If jt < fjb every column is accepted and A, B trade places;
  //  - every other accepted column beyond the front (position >= fjb) takes
  //    the next free slot in acceptance order, the slot's unmarked column
  //    going to its position; accepted columns in the front stay.
  // (Checked against the host walk exhaustively for n_sub <= 7 and on random
  // selections up to 300 columns.)  Lane t handles acceptance rank t.
  const unsigned long long below = (1ull << t) - 1ull;
  const int jt = nsub - 1;
  const int ps = (t < fjb) ? s_sel[t] : -1;
  s_occ[t] = 0;
  wave_lds_sync();
  if (t < fjb && ps < fjb) s_occ[ps] = 1;
  wave_lds_sync();
  const unsigned long long freem = __ballot(t < fjb && s_occ[t] == 0);
  if ((freem >> t) & 1ull) s_slot[__popcll(freem & below)] = t;
  const bool zero_sel = s_occ[0] != 0;
  const bool hasA = jt > 0 && __ballot(t < fjb && ps == jt) != 0ull;
  const bool farA = hasA && jt >= fjb;
  const unsigned long long farm = __ballot(t < fjb && ps >= fjb && !(hasA && ps == jt));
  wave_lds_sync();
  int mp = ps, um_dst = 0, um_src = 0;
  bool um = false;
  if (t < fjb) {
    if (hasA && ps == jt) {  // A
      mp = 0;
      if (farA) {
        um = true;
        um_dst = jt;
        um_src = s_slot[0];
      }
    } else if (ps >= fjb) {
      const int k = __popcll(farm & below) + (farA ? 1 : 0);
      mp = s_slot[k];
      um = true;
      um_dst = ps;
      um_src = mp;
    } else if (hasA && ps == 0) {  // B
      mp = farA ? s_slot[0] : jt;
    }
  }
  const int nswap = __popcll(farm) + (farA ? (zero_sel ? 2 : 1) : (hasA ? 1 : 0));
  // --- 3. plan: a (dst, src) pair for every accepted column and every
  // unmarked column that moved; the candidate rank of each panel column; the
  // panel form.
  const bool mvm = t < fjb && mp != ps;
  const unsigned long long bm = __ballot(mvm), bu = __ballot(um);
  if (mvm) {
    const int q = __popcll(bm & below);
    perm[2 * q] = j + mp;
    perm[2 * q + 1] = j + ps;
  }
  if (um) {
    const int q = __popcll(bm) + __popcll(bu & below);
    perm[2 * q] = j + um_dst;
    perm[2 * q + 1] = j + um_src;
  }
  // every accepted column now sits in [0, fjb)
  if (t < fjb) perm[4 * KMAX + mp] = s_rank[t];
  if (t == 0) {
    st->fjb = fjb;
    st->npair = __popcll(bm) + __popcll(bu);
    st->nswap = nswap;
    const int fast = (fjb >= 2 && m - j >= fast_rows) ? 1 : 0;
    st->fast = fast;
    if (!fast && !exact_queued) st->halt = 2;
  }
}


// The DM step's candidate Gram over rows j..m (cblas_dsyrk,
// src/dgeqrdm_work.c:376-379) as per-slab partials, one 64-row slab per
// workgroup and a 16-row block per wave, read straight into the matrix cores'
// layout (the Gram's tiles accumulate from it, gram_tiles; only the upper
// triangle of a partial is written, all the reduce reads).  With a pending
// update (the previous step's, deferred: st->fused)
// the candidate columns first receive it on the slab, X -= V_prev(slab rows)
// W_prev(:, candidates) on the matrix cores, and the updated rows are written
// back: the cosines, the panel and the later steps read updated columns.
// Slabs past m (the grid is sized for the lowest j) write zero partials.
__global__ __launch_bounds__(256) void k_gram_cand(double* A, int lda, int m, double* partial,
                                                   const DevState* st, StepBufs bufs) {
  __shared__ double Ws[KMAX][KMAX + 1];       // W_prev(i, candidate c) at Ws[i][c]
  __shared__ double Gp[4][10][4][64];         // each wave's tiles, in its lanes' layout
  if (dm_stopped(st)) return;
  const int j = st->j;
  // the greedy takes candidate 0 even when nc is 0 (all norms zero), so it
  // receives the update too; the cosines read the Gram only when nc > 1
  const int nc = max(st->nc, 1);
  const int tid = local_tid(), lane = tid & 63, wv = tid >> 6, g = lane >> 4, li = lane & 15;
  // wave wv takes rows rb .. rb + 15 of the workgroup's 64-row slab, read
  // straight into the matrix cores' layout: X(rb + g + 4v, candidate 16u + li)
  const int rb = j + blockIdx.x * GRAM_ROWS + 16 * wv;
  const bool pend = update_pending(st);
  double* cc[4];
  bool cok[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int c = 16 * u + li;
    cok[u] = c < nc;
    cc[u] = A + (size_t)(j + (cok[u] ? st->cand_idx[c] : 0)) * lda;
  }
  dbl4 xb[4];
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = rb + g + 4 * v;
      xb[u][v] = (cok[u] && r < m) ? cc[u][r] : 0.0;
    }
  if (pend) {
    // X -= V_prev(rows, :) W_prev(:, candidates) on the matrix cores: V rows
    // straight into the A operand, W through LDS; the product comes out in
    // X's layout.  The updated rows are written back.
    const double* Vp = v_prev(bufs, st);
    const double* Wq = w_prev(bufs, st);
    {
      // wave wv gathers candidates wv, wv + 4, ... (a candidate's 64 entries
      // are contiguous): the indices from one load per lane, every W load
      // in flight before the LDS stores
      const int cid = st->cand_idx[lane];
      double wc[KMAX / 4];
#pragma unroll
      for (int q = 0; q < KMAX / 4; q++) {
        const int c = wv + 4 * q;
        const int ci = __builtin_amdgcn_readlane(cid, c);
        wc[q] = (c < nc) ? Wq[(size_t)(j + ci) * KMAX + lane] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < KMAX / 4; q++) Ws[lane][wv + 4 * q] = wc[q];
    }
    double va[KMAX / 4];
#pragma unroll
    for (int t = 0; t < KMAX / 4; t++)
      va[t] = (rb + li < m) ? Vp[(size_t)(4 * t + g) * bufs.ldv + rb + li] : 0.0;
    __syncthreads();
    dbl4 d[4];
#pragma unroll
    for (int u = 0; u < 4; u++) d[u] = (dbl4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < KMAX / 4; t++)
#pragma unroll
      for (int u = 0; u < 4; u++)
        d[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[t], Ws[4 * t + g][16 * u + li], d[u], 0, 0,
                                                    0);
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = rb + g + 4 * v;
        xb[u][v] -= d[u][v];
        if (cok[u] && r < m) cc[u][r] = xb[u][v];
      }
  }
  // the slab's Gram: each wave's upper tiles from X's layout, then their sum
  // over the waves in a fixed order.  The reduce reads the upper triangle
  // only, so the tiles below the diagonal blocks are not written.
  dbl4 gt[10];
#pragma unroll
  for (int p = 0; p < 10; p++) gt[p] = (dbl4){0.0, 0.0, 0.0, 0.0};
  gram_tiles(xb, gt);
#pragma unroll
  for (int p = 0; p < 10; p++)
#pragma unroll
    for (int v = 0; v < 4; v++) Gp[wv][p][v][lane] = gt[p][v];
  __syncthreads();
  double* out = partial + (size_t)blockIdx.x * KMAX * KMAX;
  for (int idx = tid; idx < 10 * 4 * 64; idx += 256) {
    const int p = idx >> 8, v = (idx >> 6) & 3, ln = idx & 63;
    const int u = p < 4 ? 0 : (p < 7 ? 1 : (p < 9 ? 2 : 3));
    const int u2 = u + p - (u == 0 ? 0 : (u == 1 ? 4 : (u == 2 ? 7 : 9)));
    const double s = ((Gp[0][p][v][ln] + Gp[1][p][v][ln]) + Gp[2][p][v][ln]) + Gp[3][p][v][ln];
    out[(16 * u + (ln >> 4) + 4 * v) + (size_t)(16 * u2 + (ln & 15)) * KMAX] = s;
  }
}

// The DM step's candidate Gram reduce (as k_gram_reduce, RUN_DM), then the
// last workgroup to finish (atomic ticket, reset for the next launch) runs
// the greedy acceptance and the step's plan on it (select_wave).
__global__ __launch_bounds__(256) void k_reduce_select(const double* __restrict__ partial,
                                                       int nslab, double* G, double delta,
                                                       int m, int n, int fast_rows,
                                                       int exact_queued, int* perm,
                                                       unsigned* ticket, DevState* st,
                                                       StepBufs bufs) {
  __shared__ double part[4][64];
  __shared__ double Gs[KMAX * (KMAX + 1)];
  __shared__ int s_last;
  if (dm_stopped(st)) return;
  // the step's state for the selection and the W columns to clear, in
  // flight with the partials
  SelIn sin;
  if (threadIdx.x < 64) sin = select_inputs(st, delta);
  const bool pend = update_pending(st);
  double* Wq = w_prev(bufs, st);
  const int jw = st->j, ncw = max(st->nc, 1);
  const int cid_lane = st->cand_idx[threadIdx.x & 63];
  const int e = blockIdx.x * 64 + (threadIdx.x & 63);
  const int pr = e & 63, pc = e >> 6;
  const int es = (pr > pc) ? pc + pr * KMAX : e;  // symmetric: the upper triangle's value
  const int q = threadIdx.x >> 6;
  double sacc = 0.0;
  for (int b0 = q; b0 < nslab; b0 += 32) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int b = b0 + 4 * u;
      v[u] = partial[(size_t)(b < nslab ? b : q) * KMAX * KMAX + es];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) sacc += (b0 + 4 * u < nslab) ? v[u] : 0.0;
  }
  part[q][threadIdx.x & 63] = sacc;
  __syncthreads();
  if (q == 0)
    st_dev(G + e, (part[0][threadIdx.x] + part[1][threadIdx.x]) +
                      (part[2][threadIdx.x] + part[3][threadIdx.x]));
  handoff_wait();
  __syncthreads();
  if (threadIdx.x == 0) s_last = last_to_arrive(ticket, gridDim.x) ? 1 : 0;
  __syncthreads();
  if (!s_last) return;
  {
    // G from the launch's workgroups into LDS, every load in flight at once
    double gv[KMAX * KMAX / 256];
#pragma unroll
    for (int q = 0; q < KMAX * KMAX / 256; q++) gv[q] = ld_dev(G + threadIdx.x + 256 * q);
#pragma unroll
    for (int q = 0; q < KMAX * KMAX / 256; q++) {
      const int e2 = threadIdx.x + 256 * q;
      Gs[(e2 & 63) + (e2 >> 6) * (KMAX + 1)] = gv[q];
    }
  }
  __syncthreads();
  if (threadIdx.x < 64) select_wave(Gs, sin, m, n, fast_rows, exact_queued, perm, st);
  if (pend) {
    // the candidates received the deferred update in k_gram_cand: wave w
    // clears the W columns of candidates w, w + 4, ... (64 entries each)
    const int lane = threadIdx.x & 63;
    for (int cq = threadIdx.x >> 6; cq < ncw; cq += 4) {
      const int ci = __builtin_amdgcn_readlane(cid_lane, cq);
      Wq[(size_t)(jw + ci) * KMAX + lane] = 0.0;
    }
  }
}

void launch_gram_select(double* A, int lda, int j0, int m, int n, double* partial,
                        double* G, double delta, int fast_rows, int exact_queued, int* perm,
                        unsigned* ticket, DevState* st, const StepBufs& bufs, hipStream_t s) {
  int nslab = (m - j0 + GRAM_ROWS - 1) / GRAM_ROWS;
  if (nslab < 1) nslab = 1;
  hipLaunchKernelGGL(k_gram_cand, dim3(nslab), dim3(256), 0, s, A, lda, m, partial, st, bufs);
  hipLaunchKernelGGL(k_reduce_select, dim3(KMAX * KMAX / 64), dim3(256), 0, s, partial, nslab, G,
                     delta, m, n, fast_rows, exact_queued, perm, ticket, st, bufs);
}

// ---------------------------------------------------------------------------
// Column permutation (the data movement of permute_marked,
// src/dgeqrdm_work.c:149-262): perm holds st->npair (dst, src) pairs, a
// permutation of the touched columns; column dst receives the old column src
// (all m rows), work[dst] the old work[src] and jp[dst] the old jp[src].
// Each workgroup owns a row slab, reads every source first and then writes,
// so no slab races another.
constexpr int PERM_ROWS = 16;  // rows per workgroup: m / 16 workgroups move the columns

// Row slab blk of the step's column moves (256 threads); slab 0 also moves
// the work and jpvt entries.
__device__ void apply_perm_rows(double* A, int lda, int m, double* work, int* jp,
                                const int* __restrict__ perm, const DevState* st, int blk,
                                int nblk, const StepBufs& bufs) {
  __shared__ double buf[2 * KMAX * PERM_ROWS];
  if (st->done || st->halt == 1) return;  // halt 2: this step's moves still run
  const int npair = st->npair;
  if (update_pending(st)) {
    // the deferred update's W columns move with their columns: block blk
    // takes entries blk, blk + nblk, ... of every moved W column (all reads of
    // an entry before its writes; npair <= 2 KMAX <= the block's threads)
    double* Wq = w_prev(bufs, st);
    for (int i = blk; i < KMAX; i += nblk) {
      const int p = local_tid();
      const double v = (p < npair) ? Wq[(size_t)perm[2 * p + 1] * KMAX + i] : 0.0;
      __syncthreads();
      if (p < npair) Wq[(size_t)perm[2 * p] * KMAX + i] = v;
    }
  }
  const int rs = blk * PERM_ROWS;
  // npair <= 2 * KMAX: at most 8 elements per thread (256 threads), all
  // loaded before they are stored
  const int tot = npair * PERM_ROWS;
  for (int q0 = 0; q0 < 2 * KMAX * PERM_ROWS / 256; q0 += 8) {
    if (q0 * 256 >= tot) break;
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int t = local_tid() + 256 * (q0 + q);
      const int p = t / PERM_ROWS, r = rs + t % PERM_ROWS;
      v[q] = (t < tot && r < m) ? A[(size_t)perm[2 * p + 1] * lda + r] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int t = local_tid() + 256 * (q0 + q);
      if (t < tot) buf[t] = v[q];
    }
  }
  double wv = 0.0;
  int jv = 0;
  const bool side = blk == 0 && local_tid() < npair;
  if (side) {
    wv = work[perm[2 * local_tid() + 1]];
    if (jp) jv = jp[perm[2 * local_tid() + 1]];
  }
  __syncthreads();
  for (int t = local_tid(); t < tot; t += 256) {
    const int p = t / PERM_ROWS, r = rs + t % PERM_ROWS;
    if (r < m) A[(size_t)perm[2 * p] * lda + r] = buf[t];
  }
  if (side) {
    work[perm[2 * local_tid()]] = wv;
    if (jp) jp[perm[2 * local_tid()]] = jv;
  }
}

__global__ __launch_bounds__(256) void k_apply_perm(double* A, int lda, int m, double* work,
                                                    int* jp, const int* __restrict__ perm,
                                                    const DevState* st, StepBufs bufs) {
  apply_perm_rows(A, lda, m, work, jp, perm, st, blockIdx.x, gridDim.x, bufs);
}

int apply_perm_blocks(int m) { return (m + PERM_ROWS - 1) / PERM_ROWS; }

void launch_apply_perm(double* A, int lda, int m, double* work, int* jp, const int* perm,
                       DevState* st, const StepBufs& bufs, hipStream_t s) {
  hipLaunchKernelGGL(k_apply_perm, dim3((m + PERM_ROWS - 1) / PERM_ROWS), dim3(256), 0, s, A,
                     lda, m, work, jp, perm, st, bufs);
}

// ---------------------------------------------------------------------------
// Exact Householder panel: dgeqr2_mia (src/dgeqr2.c:144-191) with dlarfg_mia
// (src/dlarfg.c:116-185) and dlarf_ (src/dlarf.c:161-185), column by column
// in one workgroup.  Truncation: for column i > 0 stop when ||x|| < thres,
// thres = 5e-14, or tau_dm*|R(0,0)| once column 0 is done (tau_dm > 0).
// truncate = 0 disables the test (unpivoted QR).  RUN_EXACT: j and fjb from
// st, tau indexed from column 0 (the step writes tau[j..]).
// The panel's work for the whole workgroup (any size up to 1024 threads;
// red holds blockDim.x / 64 doubles): returns fjb_cmp, the columns reduced.
// GROUP: block columns a wave updates at a time (1 for many-wave workgroups,
// 4 for few waves).
template <int GROUP>
__device__ __forceinline__ int panel_exact_body(double* A, int lda, int j, int m, int fjb, double tau_dm,
                                int truncate, double* tau, double* red) {
  const int msub = m - j;
  const int tid = local_tid(), lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  double* P = A + j + (size_t)j * lda;
  double thres = 5e-14;
  int i;
  for (i = 0; i < fjb; i++) {
    const int len = msub - i;
    double* col = P + i + (size_t)i * lda;
    double taui = 0.0;
    if (len > 1) {
      double xnorm = block_nrm2(col + 1, len - 1, red);
      if (truncate && i > 0 && xnorm < thres) break;
      if (xnorm != 0.0) {
        double alpha = col[0];
        double beta = -dsign(dlapy2(alpha, xnorm), alpha);
        int knt = 0;
        if (fabs(beta) < kSafmin) {
          const double rsafmn = 1.0 / kSafmin;
          do {
            ++knt;
            __syncthreads();
            for (int r = 1 + tid; r < len; r += blockDim.x) col[r] *= rsafmn;
            beta *= rsafmn;
            alpha *= rsafmn;
          } while (fabs(beta) < kSafmin);
          __syncthreads();
          xnorm = block_nrm2(col + 1, len - 1, red);
          beta = -dsign(dlapy2(alpha, xnorm), alpha);
        }
        taui = (beta - alpha) / beta;
        const double sc = 1.0 / (alpha - beta);
        __syncthreads();
        for (int r = 1 + tid; r < len; r += blockDim.x) col[r] *= sc;
        for (int k = 0; k < knt; k++) beta *= kSafmin;
        if (tid == 0) col[0] = beta;
      }
    }
    if (tid == 0) tau[i] = taui;
    __syncthreads();
    if (i + 1 < fjb) {
      const double aii = col[0];
      if (i == 0 && tau_dm > 0.0) thres = tau_dm * fabs(aii);
      if (GROUP == 1 && taui != 0.0) {
        // dlarf_ on the later block columns: one column per wave at a time
        for (int c = i + 1 + wave; c < fjb; c += nw) {
          double* cc = P + i + (size_t)c * lda;
          double s = 0.0;
          for (int r = lane; r < len; r += 64) s = fma(r == 0 ? 1.0 : col[r], cc[r], s);
          const double f = -taui * wave_sum(s);
          for (int r = lane; r < len; r += 64) cc[r] = fma(r == 0 ? 1.0 : col[r], f, cc[r]);
        }
      } else if (GROUP == 4 && taui != 0.0) {
        // wave w takes columns i+1+w, i+1+w+nw, ..., four at a time (one
        // read of the reflector for four columns); each column's sums keep
        // their order, as in the one-column form
        for (int cq = i + 1 + wave; cq < fjb; cq += 4 * nw) {
          double* cc[4];
          bool on[4];
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const int c = cq + q * nw;
            on[q] = c < fjb;
            cc[q] = P + i + (size_t)(on[q] ? c : cq) * lda;
          }
          double s[4] = {0.0, 0.0, 0.0, 0.0};
          for (int r = lane; r < len; r += 64) {
            const double v = r == 0 ? 1.0 : col[r];
#pragma unroll
            for (int q = 0; q < 4; q++)
              if (on[q]) s[q] = fma(v, cc[q][r], s[q]);
          }
          double f[4];
#pragma unroll
          for (int q = 0; q < 4; q++) f[q] = -taui * wave_sum(s[q]);
          for (int r = lane; r < len; r += 64) {
            const double v = r == 0 ? 1.0 : col[r];
#pragma unroll
            for (int q = 0; q < 4; q++)
              if (on[q]) cc[q][r] = fma(v, f[q], cc[q][r]);
          }
        }
      }
      __syncthreads();
    }
  }
  return i;
}

__global__ __launch_bounds__(1024) void k_panel_exact(double* A, int lda, int j, int m,
                                                      int fjb, double tau_dm, int truncate,
                                                      double* tau, DevState* st, int mode) {
  __shared__ double red[16];
  if (!step_runs(st, mode)) return;
  if (mode != RUN_ALWAYS) {
    j = st->j;
    fjb = st->fjb;
    tau += j;
  }
  const int i = panel_exact_body<1>(A, lda, j, m, fjb, tau_dm, truncate, tau, red);
  if (threadIdx.x == 0) {
    st->fjb_cmp = i;
    // dlarf_ inside the panel already updated the rejected columns
    // j+fjb_cmp..j+fjb-1; the trailing update starts at j+fjb (:764-766)
    st->c0 = j + fjb;
  }
}

void launch_panel_exact(double* A, int lda, int j, int m, int fjb, double tau_dm,
                        int truncate, double* tau, DevState* st, hipStream_t s, int mode) {
  hipLaunchKernelGGL(k_panel_exact, dim3(1), dim3(1024), 0, s, A, lda, j, m, fjb, tau_dm,
                     truncate, tau, st, mode);
}

// ---------------------------------------------------------------------------
// Dense copy of the block reflector V (unit lower trapezoidal, zero padded to
// KMAX columns and to mpad rows, zero above row j) into V[r + c*ldv].
__global__ __launch_bounds__(256) void k_build_v(const double* __restrict__ A, int lda, int j,
                                                 int m, int mpad, const DevState* st,
                                                 double* V, int ldv, int mode) {
  if (!step_runs(st, mode)) return;
  if (mode != RUN_ALWAYS) j = st->j;
  const int kk = st->fjb_cmp;
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= mpad) return;
  for (int c = 0; c < KMAX; c++) {
    double v = 0.0;
    int rr = r - j;  // row within the panel
    if (c < kk && r < m && rr >= c) v = (rr == c) ? 1.0 : A[(size_t)(j + c) * lda + r];
    V[(size_t)c * ldv + r] = v;
  }
}

void launch_build_v(const double* A, int lda, int j, int m, int mpad, const DevState* st,
                    double* V, int ldv, hipStream_t s, int mode) {
  hipLaunchKernelGGL(k_build_v, dim3((mpad + 255) / 256), dim3(256), 0, s, A, lda, j, m, mpad,
                     st, V, ldv, mode);
}

// ---------------------------------------------------------------------------
// The DM loop's exact-panel V (RUN_EXACT): each workgroup builds the V rows
// [rs, rs + EX_ROWS) from the panel (unit lower trapezoidal, fjb_cmp
// columns, zero padded) in 64-row slabs, writes them to V and their Gram
// V^T V to partial[blockIdx.x] (zero past m).  Workgroup 0 also zeroes the V
// rows [j - 64, j) above the panel, which the trailing update relies on.
constexpr int EX_ROWS = 4 * GRAM_ROWS;

__global__ __launch_bounds__(256) void k_vgram_exact(const double* __restrict__ A, int lda,
                                                     int m, const DevState* st, StepBufs bufs,
                                                     double* partial) {
  __shared__ double X[KMAX * (GRAM_ROWS + 1)];
  if (!step_runs(st, RUN_EXACT)) return;
  double* V = v_cur(bufs, st);
  const int ldv = bufs.ldv;
  const int j = st->j, kk = st->fjb_cmp;
  const int tid = threadIdx.x;
  dbl4 acc[4];
#pragma unroll
  for (int u = 0; u < 4; u++) acc[u] = (dbl4){0.0, 0.0, 0.0, 0.0};
  for (int sub = 0; sub < EX_ROWS / GRAM_ROWS; sub++) {
    const int rs = j + blockIdx.x * EX_ROWS + sub * GRAM_ROWS;
    if (rs >= m) break;
    double xv[KMAX * GRAM_ROWS / 256];
#pragma unroll
    for (int q = 0; q < KMAX * GRAM_ROWS / 256; q++) {
      const int e = tid + 256 * q, c = e / GRAM_ROWS, r = rs + e % GRAM_ROWS;
      const int rr = r - j;
      xv[q] = (c < kk && r < m && rr >= c) ? (rr == c ? 1.0 : A[(size_t)(j + c) * lda + r]) : 0.0;
    }
    __syncthreads();  // the previous slab's products are done with X
#pragma unroll
    for (int q = 0; q < KMAX * GRAM_ROWS / 256; q++) {
      const int e = tid + 256 * q, c = e / GRAM_ROWS, r = e % GRAM_ROWS;
      X[c * (GRAM_ROWS + 1) + r] = xv[q];
      if (rs + r < m) V[(size_t)c * ldv + rs + r] = xv[q];
    }
    __syncthreads();
    gram_slab_acc(X, acc);
  }
  if (blockIdx.x == 0) {
    const int z0 = j >= KMAX ? j - KMAX : 0;
    for (int e = tid; e < KMAX * KMAX; e += 256) {
      const int r = z0 + (e & 63), c = e >> 6;
      if (r < j) V[(size_t)c * ldv + r] = 0.0;
    }
  }
  gram_store(acc, partial + (size_t)blockIdx.x * KMAX * KMAX);
}

int vgram_exact_slabs(int j0, int m) {
  const int ns = (m - j0 + EX_ROWS - 1) / EX_ROWS;
  return ns > 0 ? ns : 1;
}

void launch_vgram_exact(const double* A, int lda, int j0, int m, const DevState* st,
                        const StepBufs& bufs, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_vgram_exact, dim3(vgram_exact_slabs(j0, m)), dim3(256), 0, s, A, lda, m,
                     st, bufs, partial);
}

// ---------------------------------------------------------------------------
// dlarft, DIRECT='F', STOREV='C' (call site src/dgeqrdm_work.c:751-754):
// T(0:i,i) = -tau_i T(0:i,0:i) (V^T v_i)(0:i), T(i,i) = tau_i, from S = V^T V.
// With nslab > 0, S is first summed from nslab partial Grams (the exact
// panel's V slabs); 256 threads, the recurrence on the first 64.
// The recurrence for the whole workgroup (at least 64 threads): Ss = V^T V
// and Ts (zero on entry) column-major with leading dimension KMAX, ts the
// kk scalars tau, tmp KMAX doubles; all in LDS.
__device__ __forceinline__ void larft_recurrence(const double* Ss, double* Ts, const double* ts, double* tmp,
                                 int kk) {
  const int p = local_tid();
  for (int i = 0; i < kk; i++) {
    const double ti = ts[i];
    if (ti != 0.0) {
      if (p < KMAX) tmp[p] = (p < i) ? -ti * Ss[p + i * KMAX] : 0.0;
      __syncthreads();
      double sacc = 0.0;
      if (p < i)
        for (int q = p; q < i; q++) sacc += Ts[p + q * KMAX] * tmp[q];
      __syncthreads();
      if (p < i) Ts[p + i * KMAX] = sacc;
      if (p == i) Ts[i + i * KMAX] = ti;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_larft(const double* __restrict__ S,
                                               const double* __restrict__ tau,
                                               const DevState* st, double* T, int mode,
                                               int nslab) {
  __shared__ double Ss[KMAX * KMAX];
  __shared__ double Ts[KMAX * KMAX];
  __shared__ double tmp[KMAX], ts[KMAX];
  if (!step_runs(st, mode)) return;
  if (mode != RUN_ALWAYS) tau += st->j;
  const int tid = threadIdx.x;
  const int kk = st->fjb_cmp;
  for (int e = tid; e < KMAX * KMAX; e += 256) {
    double v = 0.0;
    if (nslab > 0) {
      for (int b0 = 0; b0 < nslab; b0 += 8) {  // eight loads in flight
        double pv[8];
#pragma unroll
        for (int u = 0; u < 8; u++)
          pv[u] = S[(size_t)(b0 + u < nslab ? b0 + u : 0) * KMAX * KMAX + e];
#pragma unroll
        for (int u = 0; u < 8; u++) v += (b0 + u < nslab) ? pv[u] : 0.0;
      }
    } else {
      v = S[e];
    }
    Ss[e] = v;
    Ts[e] = 0.0;
  }
  if (tid < KMAX) ts[tid] = tid < kk ? tau[tid] : 0.0;
  __syncthreads();
  larft_recurrence(Ss, Ts, ts, tmp, kk);
  for (int e = tid; e < KMAX * KMAX; e += 256) T[e] = Ts[e];
}

void launch_larft(const double* S, const double* tau, const DevState* st, double* T,
                  hipStream_t s, int mode, int nslab) {
  hipLaunchKernelGGL(k_larft, dim3(1), dim3(256), 0, s, S, tau, st, T, mode, nslab);
}

// ---------------------------------------------------------------------------
// Trailing update (LAPACKE_dlarfb_mia, src/dlarfb.c:142 -> dlarfb 'L','T','F',
// 'C'): C := C - V (C^T V T)^T on rows [j, m), columns [c0, n); with
// trans = 0 the block reflector itself, C := C - V T V^T C.
//   k_gemm1: W = op(T) V^T C, the rows split in slabs; each workgroup writes
//            its slab's partial V^T C, the last one of a column tile (atomic
//            ticket) sums the partials in slab order and applies op(T);
//   k_gemm2: C -= V W.
// Matrix cores: v_mfma_f64_16x16x4f64, D(16x16) += A(16x4) B(4x16); lane l
// holds A[l%16][l/16], B[l/16][l%16] and D[l/16 + 4*v][l%16], v = 0..3 (the
// output rows of one lane are interleaved by 4, as measured on gfx950).
constexpr int G1_CHUNK = 32;
constexpr int G1_MAX_SPLIT = 16;  // row slabs of the split reduction (Wp holds this many)

// k_gemm1: the partial products Wp[y] = V^T C over row slab y.  A workgroup
// covers 64 * G1_NWG columns of C: wave w owns the 64-column tile
// G1_NWG*blockIdx.x + w and the waves walk the same 16-row chunks of the slab,
// so the V chunk (16 rows x 64 columns) is loaded once per workgroup into LDS
// (double buffered, one barrier per chunk) while each wave streams its own C
// tile through registers one chunk ahead of its MFMAs.  Two-wave workgroups,
// about two per CU, measured faster than four-wave ones, one per CU, at the
// same number of row slabs (k_gemm1 0.44 -> 0.49 of FP64 MFMA peak at 8192^2).  Lane group g = lane/16
// holds rows 4g..4g+3 of a chunk (MFMA k-step q is row 4g + q).  Only the
// slab's last chunk can run past m; its rows >= re are masked.  Each wave
// writes its 64 x 64 partial tile straight from its accumulators.
// With st (the DM loop) j, c0 come from st and ncol = ldw - c0; the grid is
// sized for the lowest j, so whole workgroups past ncol return and row slabs
// past m write zero partials.
constexpr int G1_VPAD = 4;  // LDS row padding of the V chunk (keeps 32-byte alignment)
constexpr int G1_NWG = 2;   // waves (64-column tiles) per workgroup

template <bool VEC>
__global__ __launch_bounds__(64 * G1_NWG) void k_gemm1(const double* __restrict__ A, int lda, int j,
                                               int m, int c0, int ncol,
                                               const double* __restrict__ V, int ldv,
                                               int slab, double* Wp, int ldw,
                                               const DevState* st) {
  __shared__ __align__(32) double Vs[2][KMAX][16 + G1_VPAD];
  if (st) {
    if (dm_stopped(st)) return;
    j = st->j;
    c0 = st->c0;
    ncol = ldw - c0;
  }
  if ((int)blockIdx.x * 64 * G1_NWG >= ncol) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, li = lane & 15;
  const int base = j & ~(G1_CHUNK - 1);
  const int rs = base + blockIdx.y * slab;
  const int re = max(min(rs + slab, m), rs);  // a slab past m is empty
  const int cbw = blockIdx.x * 64 * G1_NWG + wave * 64;  // this wave's tile (relative to c0)
  // columns past ncol read a valid column instead; they are never stored
  const double* Cc[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int c = cbw + 16 * u + li;
    Cc[u] = A + (size_t)(c0 + (c < ncol ? c : ncol - 1)) * lda;
  }
  dbl4 acc[4][4];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int u = 0; u < 4; u++) acc[t][u] = (dbl4){0.0, 0.0, 0.0, 0.0};
  constexpr int CH = 16;
  const int nch = (re - rs + CH - 1) / CH;  // the last chunk may end past re
  // last 4-row group that starts inside the column: with lda % 4 == 0 (the
  // vector path's requirement) rows up to round_up(m, 4) - 1 exist in memory
  const int rclamp = ((m + 3) & ~3) - 4;
  // V chunk loader: thread tid brings rows 4*(tid&3).. of the V columns
  // (tid>>2) + 16 * G1_NWG * h, h < VPT
  constexpr int VPT = 4 / G1_NWG;
  const int vq = 4 * (tid & 3);
  auto load_v = [&](int c, int h) -> dbl4 {
    const double* Vcol = V + (size_t)((tid >> 2) + 16 * G1_NWG * h) * ldv;
    const int r = rs + CH * c + vq;
    dbl4 v;
    if constexpr (VEC) {
      v = *reinterpret_cast<const dbl4*>(Vcol + min(r, rclamp));
    } else {
#pragma unroll
      for (int q = 0; q < 4; q++) v[q] = Vcol[min(r + q, m - 1)];
    }
    return v;
  };
  // rows >= re (past m, where V may hold anything) are stored as zero
  auto store_v = [&](int buf, int c, int h, dbl4 v) {
    const int r = rs + CH * c + vq;
#pragma unroll
    for (int q = 0; q < 4; q++) v[q] = (r + q < re) ? v[q] : 0.0;
    *reinterpret_cast<dbl4*>(&Vs[buf][(tid >> 2) + 16 * G1_NWG * h][vq]) = v;
  };
  auto load_c = [&](int c, dbl4 (&cc)[4]) {
    const int r = rs + CH * c + 4 * g;
    if constexpr (VEC) {
      const int rr = min(r, rclamp);
#pragma unroll
      for (int u = 0; u < 4; u++) cc[u] = *reinterpret_cast<const dbl4*>(Cc[u] + rr);
    } else {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int rq = min(r + q, m - 1);
#pragma unroll
        for (int u = 0; u < 4; u++) cc[u][q] = Cc[u][rq];
      }
    }
  };
  dbl4 ca[4];
  if (nch > 0) {
#pragma unroll
    for (int h = 0; h < VPT; h++) store_v(0, 0, h, load_v(0, h));
    load_c(0, ca);
  }
  __syncthreads();
#pragma clang loop unroll(disable)
  for (int c = 0; c < nch; c++) {
    const int buf = c & 1;
    // chunk c + 1 (clamped to the last chunk, then unused) in flight
    const int cn_idx = min(c + 1, nch - 1);
    dbl4 vnext[VPT];
#pragma unroll
    for (int h = 0; h < VPT; h++) vnext[h] = load_v(cn_idx, h);
    dbl4 cn[4];
    load_c(cn_idx, cn);
    dbl4 vc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) vc[t] = *reinterpret_cast<const dbl4*>(&Vs[buf][16 * t + li][4 * g]);
    if (rs + CH * c + CH > re) {
      // the chunk runs past re: its C rows there may be anything (padding)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const bool ok = rs + CH * c + 4 * g + q < re;
#pragma unroll
        for (int u = 0; u < 4; u++) ca[u][q] = ok ? ca[u][q] : 0.0;
      }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int u = 0; u < 4; u++)
          acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(vc[t][q], ca[u][q], acc[t][u], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int h = 0; h < VPT; h++) store_v(buf ^ 1, cn_idx, h, vnext[h]);
#pragma unroll
    for (int u = 0; u < 4; u++) ca[u] = cn[u];
    __syncthreads();
  }
  // partial tile: W row (V column) 16t + g + 4v, column cbw + 16u + li
  double* out = Wp + (size_t)blockIdx.y * KMAX * ldw;
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int cc = cbw + 16 * u + li;
    if (cc < ncol) {
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) out[(size_t)(16 * t + g + 4 * v) * ldw + cc] = acc[t][u][v];
    }
  }
}

// k_wsum: W = op(T) sum_y Wp[y] for one 64-column tile per workgroup; the
// partials are summed in slab order (deterministic), op(T) applied on the
// matrix cores (T upper triangular, zero below).  trans = 1: T^T.  NSMAX:
// 8 or 16, the smaller one not below the slab count.  (Up to 32 slabs made
// k_gemm1 faster on narrow steps, 0.49 -> 0.51 of peak, but the longer sums
// cost more than that.)
template <int NSMAX>
__global__ __launch_bounds__(256) void k_wsum(const double* __restrict__ Wp, int ns, int ldw,
                                              int c0, int ncol, const double* __restrict__ T,
                                              int trans, double* W, const DevState* st) {
  __shared__ double Ts[KMAX][KMAX + 1];
  __shared__ double Ws[KMAX][KMAX + 1];
  if (st) {
    if (dm_stopped(st)) return;
    c0 = st->c0;
    ncol = ldw - c0;
  }
  const int cb = blockIdx.x * 64;
  if (cb >= ncol) return;
  const int tid = threadIdx.x;
  {
    double tv[KMAX * KMAX / 256];
#pragma unroll
    for (int q = 0; q < KMAX * KMAX / 256; q++) tv[q] = T[tid + 256 * q];
#pragma unroll
    for (int q = 0; q < KMAX * KMAX / 256; q++) {
      const int e = tid + 256 * q;
      Ts[e & 63][e >> 6] = tv[q];
    }
  }
#pragma unroll 4
  for (int q = 0; q < KMAX * KMAX / 256; q++) {
    const int e = tid + 256 * q;
    const int i = e >> 6, cc = cb + (e & 63);
    const int ccl = cc < ncol ? cc : ncol - 1;
    // all NSMAX loads in flight (slabs past ns re-read the last one and
    // are not added)
    double pv[NSMAX];
#pragma unroll
    for (int y = 0; y < NSMAX; y++)
      pv[y] = Wp[((size_t)(y < ns ? y : ns - 1) * KMAX + i) * ldw + ccl];
    double sum = 0.0;
#pragma unroll
    for (int y = 0; y < NSMAX; y++) sum += (y < ns) ? pv[y] : 0.0;
    Ws[i][e & 63] = sum;
  }
  __syncthreads();
  // out(r, c) = sum_i op(T)(r, i) Ws(i, c); wave w rows 16w..16w+15
  const int lane = tid & 63, wv = tid >> 6, li = lane & 15, kq = lane >> 4;
  dbl4 acc[4];
#pragma unroll
  for (int u = 0; u < 4; u++) acc[u] = (dbl4){0.0, 0.0, 0.0, 0.0};
  const int ri = 16 * wv + li;
#pragma unroll 4
  for (int k0 = 0; k0 < KMAX; k0 += 4) {
    const double av = trans ? Ts[k0 + kq][ri] : Ts[ri][k0 + kq];
#pragma unroll
    for (int u = 0; u < 4; u++)
      acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, Ws[k0 + kq][16 * u + li], acc[u], 0, 0, 0);
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int cc = cb + 16 * u + li;
    if (cc < ncol) {
#pragma unroll
      for (int v = 0; v < 4; v++) W[(size_t)(16 * wv + kq + 4 * v) * ldw + cc] = acc[u][v];
    }
  }
}

// k_gemm2: C -= V W computed as C^T -= W^T V^T on the matrix cores, so that
// the accumulator's 16 lanes of one register hold 16 consecutive rows of a
// column of C (coalesced loads and stores).  Workgroup tile 64 rows x 64
// columns; wave w owns columns 16w..16w+15; the V tile (64 rows x 64) is
// shared through LDS.
__global__ __launch_bounds__(256) void k_gemm2(double* A, int lda, int j, int m, int c0,
                                               int ncol, const double* __restrict__ V, int ldv,
                                               const double* __restrict__ W, int ldw,
                                               const DevState* st) {
  __shared__ double Vs[64][KMAX + 1];
  if (st) {  // the DM loop: geometry of the current step, grid sized for the lowest j
    if (dm_stopped(st)) return;
    j = st->j;
    c0 = st->c0;
    ncol = ldw - c0;
  }
  if (j + (int)blockIdx.x * 64 >= m || (int)blockIdx.y * 64 >= ncol) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, li = lane & 15;
  const int rbase = j + blockIdx.x * 64;
  const int cbase = blockIdx.y * 64 + wave * 16;
  // C values this lane updates: rows rbase + 16t + li, columns cbase + g + 4v
  double cval[4][4];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int cc = cbase + g + 4 * v;
    const double* col = A + (size_t)(c0 + (cc < ncol ? cc : 0)) * lda;
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int r = rbase + 16 * t + li;
      cval[t][v] = (cc < ncol && r < m) ? col[r] : 0.0;
    }
  }
  {
    double vv[64 * KMAX / 256];
#pragma unroll
    for (int q = 0; q < 64 * KMAX / 256; q++) {
      const int e = tid + 256 * q;
      const int k = e >> 6, r = e & 63;
      vv[q] = (rbase + r < m) ? V[(size_t)k * ldv + rbase + r] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 64 * KMAX / 256; q++) {
      const int e = tid + 256 * q;
      Vs[e & 63][e >> 6] = vv[q];
    }
  }
  __syncthreads();
  dbl4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; t++) acc[t] = (dbl4){0.0, 0.0, 0.0, 0.0};
  const int wc = cbase + li;
  const bool wok = wc < ncol;
#pragma unroll 4
  for (int k0 = 0; k0 < KMAX; k0 += 4) {
    const double a = wok ? W[(size_t)(k0 + g) * ldw + wc] : 0.0;
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const double b = Vs[16 * t + li][k0 + g];
      acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
    }
  }
  // acc[t][v]: row (of C^T) = column cbase + g + 4v, column = row rbase + 16t + li
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int cc = cbase + g + 4 * v;
    if (cc >= ncol) continue;
    double* col = A + (size_t)(c0 + cc) * lda;
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int r = rbase + 16 * t + li;
      if (r < m) col[r] = cval[t][v] - acc[t][v];
    }
  }
}

// k_gemm2 with 32-byte row vectors (lda % 4 == 0, 32-byte aligned A): the
// tile starts at the 4-row boundary at or below j, and lane li of a lane
// group holds rows 4li..4li+3 of its columns (MFMA tile t = row 4li + t), so
// each column of C moves as one dbl4 per lane (16 lanes: 512 contiguous
// bytes).  Rows above j (V is zero there, so they would be rewritten with
// their own value) and rows past m are not stored.  (Sweeping 2, 4 or 8
// column tiles per workgroup with the V tile kept in LDS was measured
// slower: fewer workgroups, one wave per SIMD.)
__global__ __launch_bounds__(256) void k_gemm2_vec(double* A, int lda, int j, int m, int c0,
                                                   int ncol, const double* __restrict__ V,
                                                   int ldv, const double* __restrict__ W,
                                                   int ldw, const DevState* st) {
  __shared__ double Vs[64][KMAX + 1];
  if (st) {
    if (dm_stopped(st)) return;
    j = st->j;
    c0 = st->c0;
    ncol = ldw - c0;
  }
  const int r0 = j & ~3;
  if (r0 + (int)blockIdx.x * 64 >= m || (int)blockIdx.y * 64 >= ncol) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, li = lane & 15;
  const int rbase = r0 + blockIdx.x * 64;
  const int cbase = blockIdx.y * 64 + wave * 16;
  // the lane's rows rbase + 4li .. +3 (a row group starting past m reads the
  // last group inside the column, never stored)
  const int rclamp = ((m + 3) & ~3) - 4;
  const int rl = rbase + 4 * li;
  const int rld = rl < rclamp ? rl : rclamp;
  dbl4 cval[4];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int cc = cbase + g + 4 * v;
    const double* col = A + (size_t)(c0 + (cc < ncol ? cc : 0)) * lda;
    cval[v] = *reinterpret_cast<const dbl4*>(col + rld);
  }
  {
    double vv[64 * KMAX / 256];
#pragma unroll
    for (int q = 0; q < 64 * KMAX / 256; q++) {
      const int e = tid + 256 * q;
      const int k = e >> 6, r = e & 63;
      vv[q] = (rbase + r < m) ? V[(size_t)k * ldv + rbase + r] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 64 * KMAX / 256; q++) {
      const int e = tid + 256 * q;
      Vs[e & 63][e >> 6] = vv[q];
    }
  }
  __syncthreads();
  dbl4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; t++) acc[t] = (dbl4){0.0, 0.0, 0.0, 0.0};
  const int wc = cbase + li;
  const bool wok = wc < ncol;
#pragma unroll 4
  for (int k0 = 0; k0 < KMAX; k0 += 4) {
    const double a = wok ? W[(size_t)(k0 + g) * ldw + wc] : 0.0;
#pragma unroll
    for (int t = 0; t < 4; t++)
      acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Vs[4 * li + t][k0 + g], acc[t], 0, 0, 0);
  }
  // acc[t][v]: column cbase + g + 4v, row rbase + 4li + t
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int cc = cbase + g + 4 * v;
    if (cc >= ncol) continue;
    double* col = A + (size_t)(c0 + cc) * lda;
    dbl4 o;
#pragma unroll
    for (int t = 0; t < 4; t++) o[t] = cval[v][t] - acc[t][v];
    if (rl >= j && rl + 3 < m) {
      *reinterpret_cast<dbl4*>(col + rl) = o;
    } else {
#pragma unroll
      for (int t = 0; t < 4; t++)
        if (rl + t >= j && rl + t < m) col[rl + t] = o[t];
    }
  }
}

// Row slabs of k_gemm1's split reduction for columns [c0, n) and rows
// [j & ~(G1_CHUNK - 1), m): about 256 * 4 / G1_NWG workgroups (two per CU),
// so as many slabs as one 256-column workgroup per CU would have.
struct G1Geom {
  int cgroups, ns, slab;
};

static G1Geom gemm1_geom(int j, int m, int c0, int n, int nsplit) {
  const int ncol = n - c0;
  G1Geom g;
  g.cgroups = (ncol + 64 * G1_NWG - 1) / (64 * G1_NWG);
  int want = (1024 / G1_NWG + g.cgroups - 1) / g.cgroups;
  if (want > nsplit) want = nsplit;
  if (want > G1_MAX_SPLIT) want = G1_MAX_SPLIT;
  const int base = j & ~(G1_CHUNK - 1);
  const int span = m - base;
  int slab = (span + want - 1) / want;
  slab = ((slab + G1_CHUNK - 1) / G1_CHUNK) * G1_CHUNK;
  if (slab < G1_CHUNK) slab = G1_CHUNK;
  g.slab = slab;
  g.ns = (span + slab - 1) / slab;
  return g;
}

void launch_gemm1(double* A, int lda, int j, int m, int c0, int n, const double* V, int ldv,
                  double* Wp, int nsplit, hipStream_t s, const DevState* st) {
  // with st: j, c0 are lower bounds (c0 >= j + 1 then) of the step's values
  const int ncol = n - c0, mrow = m - j;
  if (ncol <= 0 || mrow <= 0) return;
  const G1Geom g = gemm1_geom(j, m, c0, n, nsplit);
  // 32-byte vector loads need 32-byte aligned columns of at least 4 rows
  const int aligned = (lda % 4 == 0) && (ldv % 4 == 0) && m >= 4 &&
                      ((reinterpret_cast<uintptr_t>(A) & 31) == 0) &&
                      ((reinterpret_cast<uintptr_t>(V) & 31) == 0);
  if (aligned)
    hipLaunchKernelGGL(k_gemm1<true>, dim3(g.cgroups, g.ns), dim3(64 * G1_NWG), 0, s, A, lda, j, m, c0,
                       ncol, V, ldv, g.slab, Wp, n, st);
  else
    hipLaunchKernelGGL(k_gemm1<false>, dim3(g.cgroups, g.ns), dim3(64 * G1_NWG), 0, s, A, lda, j, m, c0,
                       ncol, V, ldv, g.slab, Wp, n, st);
}

void launch_wsum(int j, int m, int c0, int n, const double* T, int trans, const double* Wp,
                 double* W, int nsplit, hipStream_t s, const DevState* st) {
  const int ncol = n - c0, mrow = m - j;
  if (ncol <= 0 || mrow <= 0) return;
  const G1Geom g = gemm1_geom(j, m, c0, n, nsplit);
  const dim3 grid((ncol + 63) / 64);
  static_assert(G1_MAX_SPLIT == 16, "k_wsum instances cover 16 slabs");
  if (g.ns <= 8)
    hipLaunchKernelGGL(k_wsum<8>, grid, dim3(256), 0, s, Wp, g.ns, n, c0, ncol, T, trans, W, st);
  else
    hipLaunchKernelGGL(k_wsum<16>, grid, dim3(256), 0, s, Wp, g.ns, n, c0, ncol, T, trans, W, st);
}

void launch_larfb_w(double* A, int lda, int j, int m, int c0, int n, const double* V, int ldv,
                    const double* T, int trans, double* Wp, double* W, int nsplit, hipStream_t s,
                    const DevState* st) {
  launch_gemm1(A, lda, j, m, c0, n, V, ldv, Wp, nsplit, s, st);
  launch_wsum(j, m, c0, n, T, trans, Wp, W, nsplit, s, st);
}

void launch_larfb_apply(double* A, int lda, int j, int m, int c0, int n, const double* V, int ldv,
                        const double* W, hipStream_t s, const DevState* st) {
  const int ncol = n - c0, mrow = m - j;
  if (ncol <= 0 || mrow <= 0) return;
  const int ctiles = (ncol + 63) / 64;
  const bool vec = (lda % 4 == 0) && m >= 4 && ((reinterpret_cast<uintptr_t>(A) & 31) == 0);
  if (vec)  // the tile starts at j & ~3: one more row tile at most
    hipLaunchKernelGGL(k_gemm2_vec, dim3((m - (j & ~3) + 63) / 64, ctiles), dim3(256), 0, s, A,
                       lda, j, m, c0, ncol, V, ldv, W, n, st);
  else
    hipLaunchKernelGGL(k_gemm2, dim3((mrow + 63) / 64, ctiles), dim3(256), 0, s, A, lda, j, m,
                       c0, ncol, V, ldv, W, n, st);
}

void launch_larfb(double* A, int lda, int j, int m, int c0, int n, const double* V, int ldv,
                  const double* T, int trans, double* Wp, double* W, int nsplit, hipStream_t s) {
  launch_larfb_w(A, lda, j, m, c0, n, V, ldv, T, trans, Wp, W, nsplit, s, nullptr);
  launch_larfb_apply(A, lda, j, m, c0, n, V, ldv, W, s, nullptr);
}

// ---------------------------------------------------------------------------
// norm_update (src/dgeqrdm_work.c:80-109): Drmac-Bujanovic downdate of the
// partial norms of columns [j+kk, n) after kk rows were triangularised; when
// temp2 <= tol3z the norm is recomputed from the rows below and workold
// reset.  workold is indexed by position and is not permuted, as in the
// reference.  One wave per column.
__global__ __launch_bounds__(256) void k_norm_update(const double* __restrict__ A, int lda,
                                                     int m, int n, double* work,
                                                     double* workold, DevState* st,
                                                     double tol3z) {
  __shared__ unsigned long long s_rows[4];
  if (dm_stopped(st)) return;
  // the step's end is pending from here on (run by the next k_topk or k_step_end)
  if (blockIdx.x == 0 && threadIdx.x == 0) st->pend = 1;
  const int j = st->j;
  const int kk = st->fjb_cmp;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int mrem = m - j - kk;
  double mx = 0.0;
  unsigned long long rows = 0;
  for (int q = 0; q < NORM_COLS; q++) {
    const int col = j + kk + (blockIdx.x * 4 + wave) * NORM_COLS + q;
    if (col >= n) break;
    double wk = work[col];
    if (wk == 0.0) continue;
    const double* a = A + (size_t)col * lda + j;
    // sum of squares of the kk new R rows (src/dgeqrdm_work.c:84-86): per
    // lane rows lane, lane + 64, ... (one term for kk <= 64), then the wave
    double x2 = 0.0;
    for (int r = lane; r < kk; r += 64) x2 = fma(a[r], a[r], x2);
    const double s = wave_sum(x2);
    double temp = sqrt(fabs(s)) / wk;
    const double d2 = (temp + 1.0) * (1.0 - temp);
    temp = (0.0 >= d2) ? 0.0 : d2;
    const double d1 = wk / workold[col];
    const double temp2 = temp * (d1 * d1);
    if (temp2 <= tol3z) {
      wk = (mrem > 0) ? wave_nrm2(a + kk, mrem) : 0.0;
      if (lane == 0) workold[col] = wk;
      rows += (unsigned long long)(mrem > 0 ? mrem : 0);
    } else {
      wk *= sqrt(temp);
    }
    if (lane == 0) work[col] = wk;
    if (wk == wk) mx = fmax(mx, wk);
  }
  if (lane == 0) s_rows[wave] = rows;
  block_max_atomic(mx, 0, st);
  if (threadIdx.x == 0) {
    const unsigned long long r = s_rows[0] + s_rows[1] + s_rows[2] + s_rows[3];
    if (r) atomicAdd(&st->recomp_rows, r);
  }
}

void launch_norm_update(const double* A, int lda, int j0, int m, int n, double* work,
                        double* workold, DevState* st, double tol3z, hipStream_t s) {
  const int ncols = n - j0;  // upper bound; the kernel starts at j + fjb_cmp
  if (ncols <= 0) return;
  const int waves = (ncols + NORM_COLS - 1) / NORM_COLS;
  hipLaunchKernelGGL(k_norm_update, dim3((waves + 3) / 4), dim3(256), 0, s, A, lda, m, n, work,
                     workold, st, tol3z);
}

// ---------------------------------------------------------------------------
// End of a DM step (src/dgeqrdm_work.c:740, :774-785): ncols[it] = fjb_cmp,
// the step's record and algorithmic work, j += fjb_cmp, the stop test
// maxnrm * sqrt(n_sub - fjb_cmp) <= eta, and the loop's end at j >= min(m,n).
// The pending end of the last queued step (the next step's k_topk runs it
// otherwise).
__global__ void k_step_end(DevState* st, int* ncols_dev, int m, int n, int enable_stop,
                           double eta) {
  if (st->pend) step_end_apply(st, ncols_dev, m, n, enable_stop, eta);
}

void launch_step_end(DevState* st, int* ncols_dev, int m, int n, int enable_stop, double eta,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1), 0, s, st, ncols_dev, m, n, enable_stop, eta);
}

__global__ void k_dm_begin(DevState* st, int j, int minmn, double tau, double delta, int fused) {
  st->fused = fused;
  st->kk_prev = 0;
  st->policy_on = 0;
  st->tau_step = tau;
  st->delta_step = delta;
  st->pend = 0;
  st->halt = 0;
  st->j = j;
  st->it = 0;
  st->done = (j >= minmn) ? 1 : 0;
  st->maxnrm_bits = 0ull;
  st->flags &= ~2;
  st->work = StepWork{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
}

void launch_dm_begin(DevState* st, int j, int minmn, double tau, double delta, int fused,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_dm_begin, dim3(1), dim3(1), 0, s, st, j, minmn, tau, delta, fused);
}

// ---------------------------------------------------------------------------
// Small helpers.
__global__ void k_set_kk(DevState* st, int kk) { st->fjb_cmp = kk; }

void launch_set_kk(DevState* st, int kk, hipStream_t s) {
  hipLaunchKernelGGL(k_set_kk, dim3(1), dim3(1), 0, s, st, kk);
}

__global__ void k_resume(DevState* st) { st->halt = 0; }

void launch_resume(DevState* st, hipStream_t s) {
  hipLaunchKernelGGL(k_resume, dim3(1), dim3(1), 0, s, st);
}

// The trailing update's form from the next step on: 0 = applied within its
// step (the three-kernel form), after the pending merged update was applied.
__global__ void k_set_fused(DevState* st, int fused) { st->fused = fused; }

void launch_set_fused(DevState* st, int fused, hipStream_t s) {
  hipLaunchKernelGGL(k_set_fused, dim3(1), dim3(1), 0, s, st, fused);
}



// Swap two whole columns (cblas_dswap of the fixed-column pre-pass,
// src/dgeqrdm_work.c:599).
__global__ __launch_bounds__(256) void k_swap2(double* A, int lda, int m, int p, int q) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= m) return;
  double* cp = A + (size_t)p * lda;
  double* cq = A + (size_t)q * lda;
  const double t = cp[r];
  cp[r] = cq[r];
  cq[r] = t;
}

void launch_swap2(double* A, int lda, int m, int p, int q, hipStream_t s) {
  hipLaunchKernelGGL(k_swap2, dim3((m + 255) / 256), dim3(256), 0, s, A, lda, m, p, q);
}

// Zero fill of a workspace region (4-byte words): the clearing that a memset
// node would do, as an ordinary kernel.
__global__ __launch_bounds__(256) void k_zero_words(unsigned* p, size_t nw) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < nw; i += (size_t)gridDim.x * 256)
    p[i] = 0u;
}

void launch_zero(void* p, size_t bytes, hipStream_t s) {
  const size_t nw = bytes / 4;
  if (nw == 0) return;
  size_t blocks = (nw + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_zero_words, dim3((unsigned)blocks), dim3(256), 0, s,
                     reinterpret_cast<unsigned*>(p), nw);
}

}  // namespace qrdm
