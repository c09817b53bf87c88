// wide.hip -- DM steps wider than 64 columns (nb > 64), driven from the host.
//
// The reference accepts any nb > 0 (src/dgeqrdm_work.c:572-574) and takes up
// to k_max = min(nb, m - j, n - j) candidates per step (:703).  The device
// loop of kernels.hip, panel.hip and trailing.hip is built around one
// 64-column tile per step; a call with 64 < nb <= KW runs each step with the
// kernels below and the host walking the step (one synchronisation per
// step):
//   k_topk_wide    the top k_max of the partial norms (value descending,
//                  position ascending, the order of the reference's qsort):
//                  passes of 1024-key bitonic sorts, each keeping its first
//                  k_max keys, until one workgroup holds them all;
//   k_gram_wide    the candidates' Gram over rows j..m on the matrix cores,
//                  one 64 x 64 tile pair per workgroup (rows in a fixed order);
//   k_select_wide  the candidate count (> tau * max) and the greedy cosine
//                  acceptance (:354-403), with the cosine of the 64-wide
//                  path: |G(q, t) / norm_q / norm_t| >= delta marks a conflict;
//   k_swap_pairs   the column moves of permute_marked's walk (computed on the
//                  host, src/dgeqrdm_work.c:149-262), with work and jpvt;
// then the column-by-column panel on all fjb columns (k_panel_exact, any
// width), the block reflector applied in 64-column chunks of the kept
// reflectors (the same product H_0 .. H_{k-1} as one dlarfb of width k), and
// the norm downdate over the step's fjb_cmp new rows.
#include "qrdm_internal.h"

#include <climits>

namespace qrdm {

constexpr int WT_KEYS = 1024;  // keys per workgroup of a sort pass

// One pass: workgroup b sorts keys [1024 b, 1024 (b + 1)) of the L input
// keys (key_before's order; index idx_in[g], or g itself when idx_in is
// null; missing keys sort last) and writes its first K to out[b K ..].
__global__ __launch_bounds__(256) void k_topk_wide(const double* __restrict__ val_in,
                                                   const int* __restrict__ idx_in, int L,
                                                   double* val_out, int* idx_out, int K) {
  __shared__ double sv[WT_KEYS];
  __shared__ int si[WT_KEYS];
  const int tid = threadIdx.x;
  const int base = blockIdx.x * WT_KEYS;
  for (int e = tid; e < WT_KEYS; e += 256) {
    const int g = base + e;
    const bool in = g < L;
    sv[e] = in ? val_in[g] : -1.0;  // norms are >= 0
    si[e] = in ? (idx_in ? idx_in[g] : g) : INT_MAX;
  }
  __syncthreads();
  for (int size = 2; size <= WT_KEYS; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int e = tid; e < WT_KEYS / 2; e += 256) {
        const int lo = 2 * e - (e & (stride - 1)), hi = lo + stride;
        const bool up = (lo & size) == 0;  // this run in key order
        const double va = sv[lo], vb = sv[hi];
        const int ia = si[lo], ib = si[hi];
        if (up ? key_before(vb, ib, va, ia) : key_before(va, ia, vb, ib)) {
          sv[lo] = vb;
          sv[hi] = va;
          si[lo] = ib;
          si[hi] = ia;
        }
      }
      __syncthreads();
    }
  }
  for (int e = tid; e < K; e += 256) {
    val_out[(size_t)blockIdx.x * K + e] = sv[e];
    idx_out[(size_t)blockIdx.x * K + e] = si[e];
  }
}

// G(p, q) = sum_{r = j..m-1} A(r, j + idx[p]) A(r, j + idx[q]) for the 64 x 64
// tile pair (ti <= tj) of this workgroup; G column-major with leading
// dimension ldg.  Rows go through LDS in 64-row chunks; wave w accumulates
// tile rows 16w..16w+15 on the matrix cores.
__global__ __launch_bounds__(256) void k_gram_wide(const double* __restrict__ A, int lda, int j,
                                                   int m, const int* __restrict__ idx, int nc,
                                                   double* G, int ldg) {
  __shared__ double X[64][65];
  __shared__ double Y[64][65];
  const int nt = (nc + 63) / 64;
  int t = blockIdx.x, ti = 0;
  while (t >= nt - ti) {
    t -= nt - ti;
    ti++;
  }
  const int tj = ti + t;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, g = lane >> 4, li = lane & 15;
  dbl4 acc[4];
#pragma unroll
  for (int cb = 0; cb < 4; cb++) acc[cb] = (dbl4){0.0, 0.0, 0.0, 0.0};
  for (int r0 = j; r0 < m; r0 += 64) {
    for (int e = tid; e < 64 * 64; e += 256) {
      const int rr = e & 63, c = e >> 6, r = r0 + rr;  // consecutive threads: consecutive rows
      const int pc = 64 * ti + c, qc = 64 * tj + c;
      X[rr][c] = (r < m && pc < nc) ? A[(size_t)(j + idx[pc]) * lda + r] : 0.0;
      Y[rr][c] = (r < m && qc < nc) ? A[(size_t)(j + idx[qc]) * lda + r] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int k0 = 0; k0 < 64; k0 += 4) {
      const double a = X[k0 + g][16 * wave + li];
#pragma unroll
      for (int cb = 0; cb < 4; cb++)
        acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Y[k0 + g][16 * cb + li], acc[cb], 0, 0,
                                                       0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int cb = 0; cb < 4; cb++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int p = 64 * ti + 16 * wave + g + 4 * v, q = 64 * tj + 16 * cb + li;
      if (p < nc && q < nc) G[p + (size_t)q * ldg] = acc[cb][v];
    }
}

// The step's selection from the sorted candidates (val: their norms, the
// first kmax): out[0] = nc (> tau * val[0], at most kmax), out[1] = fjb,
// out[2 + s] = the candidate rank accepted s-th.  Candidate 0 is always
// accepted; candidate i after it when |cos| < delta against every accepted
// one and fewer than kmax are accepted.
__global__ __launch_bounds__(256) void k_select_wide(const double* __restrict__ G, int ldg,
                                                     const double* __restrict__ val, int kmax,
                                                     double tau, double delta, int* out) {
  __shared__ unsigned long long conf[KW][KW / 64];
  __shared__ double r[KW];
  __shared__ int s_nc;
  const int t = threadIdx.x;
  if (t == 0) {
    const double lim = tau * val[0];
    int nc = 0;
    while (nc < kmax && val[nc] > lim) nc++;
    s_nc = nc;
  }
  __syncthreads();
  const int nc = s_nc;
  for (int i = t; i < KW; i += blockDim.x) r[i] = (i < nc) ? 1.0 / val[i] : 0.0;
  __syncthreads();
  for (int i = t; i < KW; i += blockDim.x) {
    for (int w = 0; w < KW / 64; w++) {
      unsigned long long bits = 0ull;
      for (int b = 0; b < 64; b++) {
        const int q = 64 * w + b;
        if (q < i && i < nc) {
          const double c = fabs(G[q + (size_t)i * ldg] * r[q] * r[i]);  // G(q, i), q < i
          if (c >= delta) bits |= 1ull << b;
        }
      }
      conf[i][w] = bits;
    }
  }
  __syncthreads();
  if (t == 0) {
    unsigned long long acc[KW / 64];
    for (int w = 0; w < KW / 64; w++) acc[w] = 0ull;
    acc[0] = 1ull;
    int nsel = 1;
    out[2] = 0;
    for (int i = 1; i < nc; i++) {
      unsigned long long hit = 0ull;
      for (int w = 0; w < KW / 64; w++) hit |= conf[i][w] & acc[w];
      if (hit == 0ull && nsel < kmax) {
        acc[i >> 6] |= 1ull << (i & 63);
        out[2 + nsel] = i;
        nsel++;
      }
    }
    out[0] = nc;
    out[1] = nsel;
  }
}

// Position swaps (relative to column j), in order: whole columns of A, work
// and jpvt.
__global__ __launch_bounds__(256) void k_swap_pairs(double* A, int lda, int m, int j,
                                                    double* work, int* jp,
                                                    const int* __restrict__ pairs, int np) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r < m) {
    for (int p = 0; p < np; p++) {
      double* ca = A + (size_t)(j + pairs[2 * p]) * lda;
      double* cb = A + (size_t)(j + pairs[2 * p + 1]) * lda;
      const double x = ca[r];
      ca[r] = cb[r];
      cb[r] = x;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int p = 0; p < np; p++) {
      const int a = j + pairs[2 * p], b = j + pairs[2 * p + 1];
      const double w = work[a];
      work[a] = work[b];
      work[b] = w;
      const int q = jp[a];
      jp[a] = jp[b];
      jp[b] = q;
    }
  }
}

// The state the host-driven kernels read: the step's column and width, a
// fresh maximum norm.
__global__ void k_wide_state(DevState* st, int j, int kk) {
  st->j = j;
  st->fjb_cmp = kk;
  st->maxnrm_bits = 0ull;
  st->done = 0;
  st->halt = 0;
}

int wide_sort_keys(int n) { return ((n + WT_KEYS - 1) / WT_KEYS) * KW; }

void launch_topk_wide(const double* work, int nsub, int kmax, double* val[2], int* idx[2],
                      const double** out_val, hipStream_t s) {
  const double* vin = work;
  const int* iin = nullptr;
  int L = nsub, b = 0, grid;
  do {
    grid = (L + WT_KEYS - 1) / WT_KEYS;
    hipLaunchKernelGGL(k_topk_wide, dim3(grid), dim3(256), 0, s, vin, iin, L, val[b], idx[b],
                       kmax);
    vin = val[b];
    iin = idx[b];
    L = grid * kmax;
    b ^= 1;
  } while (grid > 1);
  *out_val = vin;
}

const int* wide_topk_idx(double* val[2], int* idx[2], const double* out_val) {
  return out_val == val[0] ? idx[0] : idx[1];
}

void launch_gram_wide(const double* A, int lda, int j, int m, const int* idx, int nc, double* G,
                      int ldg, hipStream_t s) {
  const int nt = (nc + 63) / 64;
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_gram_wide, dim3(nt * (nt + 1) / 2), dim3(256), 0, s, A, lda, j, m, idx,
                     nc, G, ldg);
}

void launch_select_wide(const double* G, int ldg, const double* val, int kmax, double tau,
                        double delta, int* out, hipStream_t s) {
  hipLaunchKernelGGL(k_select_wide, dim3(1), dim3(256), 0, s, G, ldg, val, kmax, tau, delta,
                     out);
}

void launch_swap_pairs(double* A, int lda, int m, int j, double* work, int* jp, const int* pairs,
                       int np, hipStream_t s) {
  if (np <= 0) return;
  hipLaunchKernelGGL(k_swap_pairs, dim3((m + 255) / 256), dim3(256), 0, s, A, lda, m, j, work, jp,
                     pairs, np);
}

void launch_wide_state(DevState* st, int j, int kk, hipStream_t s) {
  hipLaunchKernelGGL(k_wide_state, dim3(1), dim3(1), 0, s, st, j, kk);
}

}  // namespace qrdm
