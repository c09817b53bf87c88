// batched.hip -- many independent DM-QR factorisations in one persistent
// launch: each workgroup factorises whole matrices, one after another,
// taking the next one (in decreasing flop count, ordered by the host) from a
// device counter, with the DM state of the current matrix in LDS.
//
// The arithmetic of a step is the plain path's (src/dgeqrdm_work.c:694-787):
// the same device routines for the norms (wave_nrm2), the top-k order
// (wave_sort64 / wave_merge64), the greedy acceptance and permute_marked
// plan (select_wave), and the panel in either of the plain path's forms --
// on panels of at least gram_rows rows the Gram-based one (chol1_core,
// recon_core: CholeskyQR2 with Householder reconstruction, T from V^T V),
// otherwise dgeqr2_mia column by column (panel_exact_body) with dlarft
// (larft_recurrence).  The candidate Gram, the panel's products and Grams
// and the trailing update run on the FP64 matrix cores inside the
// workgroup.  Limits: n <= BT_NMAX (the per-column state lives in LDS); m
// is free.

namespace qrdm {

constexpr int BT_NMAX = BATCH_NMAX;
constexpr double kEpsD = 1.1102230246251565e-16;  // LAPACK dlamch('E') = 2^-53

// One matrix of a batch; layout of qrdm_batch_item (include/qrdm_amd.h).
struct BatchItem {
  int m, n, lda, pad;
  long long a_off, tau_off, col_off;
};

struct BatchParams {
  double delta, tau_dm, eta3, tol3z;
  int nb, stop_mode;
  const PolicyNet* policy;  // learned (tau, delta) per step, or nullptr
  float* trace;             // per-step training records (POL_REC floats), or nullptr
  unsigned* trace_count;    // records written (may exceed trace_cap)
  int trace_cap;
  float tau_lo, tau_hi, delta_lo, delta_hi;  // label normalisation
  int gram_rows;  // panels of at least this many rows (and 2+ columns) take the Gram-based form
};

template <int NW>
struct BatchLds {
  double work[BT_NMAX], wold[BT_NMAX];
  int jp[BT_NMAX];
  union {
    struct {
      double lv[NW][KMAX];
      int li[NW][KMAX];
    } tk;                       // the top-k's per-wave lists
    double E[2][KMAX][LDP];     // the Gram-based panel's matrices
  } u;
  double R1[KMAX * LDP];  // Gram slab / G / V^T V / trailing-update W
  double Ts[KMAX * LDP];  // T (column-major, leading dimension KMAX)
  double ts[KMAX], tmp[KMAX];
  double rinv[KMAX], dsg[KMAX], tl[KMAX], scv[KMAX];
  double red[16];
  int perm[4 * KMAX + KMAX];
  int s_done, s_why;
  DevState st;
  float feat[POL_IN];
  float hbuf[POL_H];
  int hist[POL_BINS];
};

// Unit lower trapezoidal V of the panel at column j: V(row, c) for c < kk.
__device__ __forceinline__ double bt_v(const double* a, int lda, int j, int kk, int row, int c,
                                       int m) {
  const int rr = row - j;
  if (c >= kk || row >= m || rr < c) return 0.0;
  return rr == c ? 1.0 : a[(size_t)(j + c) * lda + row];
}

// C := C - V T^T V^T C on rows [j, m), columns [c0, n), in 64-column tiles.
// P = V^T C: wave w takes the 16-row chunks w % 4, w % 4 + 4, ... and the
// 16-column groups u of its share (all four with 4 waves; two, by w / 4, with
// 8), accumulating its partial on the matrix cores (lane group g holds rows
// 4g..4g+3 of a chunk, as k_gemm1); the partials are summed in wave order in
// LDS.  W = T^T P (waves 0..3, rows 16w..), then C -= V W as C^T -= W^T V^T,
// wave w taking the 16-row blocks w, w + NW, ...
template <int NW>
__device__ __forceinline__ void bt_trailing(double* a, int lda, int j, int m, int kk, int c0, int n,
                            BatchLds<NW>& L) {
  constexpr int UPW = 16 / NW;  // 16-column groups per wave in P
  const int tid = local_tid(), lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, li = lane & 15;
  const int wr = wave & 3, u0 = UPW * (wave >> 2);
  double* Wl = L.R1;  // [k][c], leading dimension KMAX + 1
  // V columns of this lane (16t + li), clamped to a valid column
  const double* vcol[4];
#pragma unroll
  for (int t = 0; t < 4; t++) vcol[t] = a + (size_t)(j + min(16 * t + li, kk - 1)) * lda;
  for (int cb = c0; cb < n; cb += 64) {
    const int ncl = min(64, n - cb);
    const double* ccol[UPW];
#pragma unroll
    for (int u = 0; u < UPW; u++) ccol[u] = a + (size_t)(cb + min(16 * (u0 + u) + li, ncl - 1)) * lda;
    dbl4 acc[4][UPW];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
      for (int u = 0; u < UPW; u++) acc[t][u] = (dbl4){0.0, 0.0, 0.0, 0.0};
    for (int rs = j + 16 * wr; rs < m; rs += 64) {
      double vr[4][4], cr[UPW][4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int row = rs + 4 * g + q, rr = row - j;
        const bool rok = row < m;
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const int c = 16 * t + li;
          const double x = (rok && c < kk && rr > c) ? vcol[t][row] : 0.0;
          vr[t][q] = (rr == c && c < kk) ? 1.0 : x;
        }
#pragma unroll
        for (int u = 0; u < UPW; u++)
          cr[u][q] = (rok && 16 * (u0 + u) + li < ncl) ? ccol[u][row] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
          for (int u = 0; u < UPW; u++)
            acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(vr[t][q], cr[u][q], acc[t][u], 0, 0,
                                                             0);
    }
    // sum of the wave partials, in wave order: P(16t + g + 4v, 16u + li)
    __syncthreads();  // the previous tile is done with Wl
    for (int w = 0; w < NW; w++) {
      if (wave == w) {
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
          for (int u = 0; u < UPW; u++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
              double* p = &Wl[(16 * t + g + 4 * v) * (KMAX + 1) + 16 * (u0 + u) + li];
              *p = (wr == 0) ? acc[t][u][v] : *p + acc[t][u][v];
            }
      }
      __syncthreads();
    }
    // W = T^T P: T(i, r) = Ts[i + r * KMAX]
    dbl4 wacc[4];
    if (wave < 4) {
#pragma unroll
      for (int u = 0; u < 4; u++) wacc[u] = (dbl4){0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
      for (int k0 = 0; k0 < KMAX; k0 += 4) {
        const double av = L.Ts[(k0 + g) + (16 * wave + li) * KMAX];
#pragma unroll
        for (int u = 0; u < 4; u++)
          wacc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(
              av, Wl[(k0 + g) * (KMAX + 1) + 16 * u + li], wacc[u], 0, 0, 0);
      }
    }
    __syncthreads();
    if (wave < 4) {
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) Wl[(16 * wave + g + 4 * v) * (KMAX + 1) + 16 * u + li] = wacc[u][v];
    }
    __syncthreads();
    // C -= V W on 16-row blocks: D_u(c, r) = sum_k W(k, 16u + c) V(r, k)
    const double* cbase = a + (size_t)cb * lda;
    for (int rb = j + 16 * wave; rb < m; rb += 16 * NW) {
      // W is read from LDS in every row block: kept out of the loop (64
      // doubles per lane), it pushed the kernel past 512 VGPRs into scratch
      __asm__ volatile("" ::: "memory");
      const int row = rb + li;
      double vk[KMAX / 4];
#pragma unroll
      for (int s = 0; s < KMAX / 4; s++) vk[s] = bt_v(a, lda, j, kk, row, 4 * s + g, m);
      double cval[4][4];
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int c = 16 * u + g + 4 * v;
          cval[u][v] = (c < ncl && row < m) ? cbase[(size_t)c * lda + row] : 0.0;
        }
      dbl4 d[4];
#pragma unroll
      for (int u = 0; u < 4; u++) d[u] = (dbl4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < KMAX / 4; s++)
#pragma unroll
        for (int u = 0; u < 4; u++)
          d[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(Wl[(4 * s + g) * (KMAX + 1) + 16 * u + li],
                                                     vk[s], d[u], 0, 0, 0);
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int c = 16 * u + g + 4 * v;
          if (c < ncl && row < m) a[(size_t)(cb + c) * lda + row] = cval[u][v] - d[u][v];
        }
    }
  }
}

// S = the sum of the four waves' tile Grams gt (in wave order), both
// triangles; four waves.  Ends with a barrier.
__device__ __forceinline__ void bt_gram_store(double (*S)[LDP], const dbl4p (&gt)[10]) {
  const int t = local_tid(), lane = t & 63, wv = t >> 6, g = lane >> 4, li = lane & 15;
  for (int e = t; e < KMAX * KMAX; e += 256) S[e >> 6][e & 63] = 0.0;
  lds_barrier();
  for (int w = 0; w < 4; w++) {
    if (wv == w) {
      int p = 0;
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int u2 = u; u2 < 4; u2++) {
#pragma unroll
          for (int v = 0; v < 4; v++) S[16 * u + g + 4 * v][16 * u2 + li] += gt[p][v];
          p++;
        }
    }
    lds_barrier();
  }
  // the diagonal tiles are whole; mirror the tiles below them
  for (int e = t; e < KMAX * KMAX; e += 256) {
    const int r = e >> 6, c = e & 63;
    if ((r >> 4) > (c >> 4)) S[r][c] = S[c][r];
  }
  lds_barrier();
}

// Gram X^T X of up to 64 columns over rows [r0, m), X(r, c) = colval(c, r),
// into R1 (both triangles, leading dimension KMAX + 1); four waves.  Each
// wave takes its own 16-row blocks (rows r0 + 16 wv, then every 64th), reads
// them straight into the matrix cores' operand layout and accumulates the
// upper tiles in registers (bt_gram_tiles); no barrier until the waves' sums
// go into R1 in wave order.
template <int NW, typename ColFn>
__device__ __forceinline__ void bt_gram(int r0, int m, int ncols, ColFn colval, BatchLds<NW>& L) {
  static_assert(NW == 4, "the Gram's tiles are spread over four waves");
  const int tid = local_tid(), lane = tid & 63, wv = tid >> 6, g = lane >> 4, li = lane & 15;
  dbl4p gt[10];
#pragma unroll
  for (int p = 0; p < 10; p++) gt[p] = (dbl4p){0.0, 0.0, 0.0, 0.0};
  for (int rb = r0 + 16 * wv; rb < m; rb += NW * 16) {
    dbl4p xb[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int c = 16 * u + li, r = rb + g + 4 * v;
        xb[u][v] = (c < ncols && r < m) ? colval(c, r) : 0.0;
      }
    gram_tiles(xb, gt);
  }
  bt_gram_store(reinterpret_cast<double (*)[LDP]>(L.R1), gt);
}

// The Gram-based panel (panel.hip's tall-panel form) by the workgroup: R1 =
// chol of the panel's Gram, gathered from the candidate Gram in L.R1 through
// the ranks select_wave left in perm[4 KMAX ..] (chol1_core); Q = P R1^-1
// and G2 = Q^T Q over 64-row slabs; the reconstruction (recon_core: R, V1,
// tau, M); V2 = P_bottom M written into the panel, with V^T V; then tau and
// T from V^T V (the compact WY identity, as k_reduce_t), T into L.Ts.
// Returns the columns kept, or -1 with A and tau unchanged when the form
// falls back (the caller then runs the column-by-column panel).
template <int NW>
__device__ __forceinline__ int bt_panel_gram(double* a, int lda, int j, int m, int fjb, double tau_dm,
                             double* tau, BatchLds<NW>& L) {
  static_assert(NW == 4, "the one-workgroup factorisations run on four waves");
  const int tid = local_tid(), lane = tid & 63, wv = tid >> 6;
  double (*E0)[LDP] = L.u.E[0];
  double (*E1)[LDP] = L.u.E[1];
  double (*RT)[LDP] = reinterpret_cast<double (*)[LDP]>(L.R1);
  double (*TS)[LDP] = reinterpret_cast<double (*)[LDP]>(L.Ts);
  auto Yof = [](double* p) { return reinterpret_cast<double (*)[BLK][BLK + 1]>(p); };
  // 1. R1 and R1^-1 (E0, E1; zero outside k1 x k1)
  const int* rank = L.perm + 4 * KMAX;
  if (tid < KMAX) {
    const int rk = tid < fjb ? rank[tid] : 0;
    L.tmp[tid] = tid < fjb ? L.R1[rk * (KMAX + 2)] : 0.0;  // G(rk, rk)
  }
  for (int e = tid; e < KMAX * KMAX; e += NW * 64) {
    const int r = e >> 6, c = e & 63;
    E0[r][c] = (r < fjb && c < fjb) ? L.R1[rank[r] + rank[c] * (KMAX + 1)] : (r == c ? 1.0 : 0.0);
  }
  lds_barrier();
  const int k1 = chol1_core(E0, E1, Yof(L.Ts), L.rinv, L.tmp, fjb, tau_dm, &L.s_done, &L.s_why);
  if (k1 < 0) return -1;
  for (int e = tid; e < KMAX * KMAX; e += NW * 64) {
    const int r = e >> 6, c = e & 63;
    if (!(r < k1 && c < k1)) {
      E0[r][c] = 0.0;
      E1[r][c] = 0.0;
    }
  }
  lds_barrier();
  // 2. G2 = (P R1^-1)^T (P R1^-1): each wave takes its own 16-row blocks (rows
  // j + 16 wv, then every 64th) with no barrier: Q_blk = P_blk R1^-1 on the
  // matrix cores, P read straight into the A operand; Q_blk^T Q_blk from the
  // product's own layout (bt_gram_tiles).  The waves' sums go into RT in wave
  // order; identity beyond k1.
  const int g = lane >> 4, li = lane & 15;
  dbl4p gt[10];
#pragma unroll
  for (int p = 0; p < 10; p++) gt[p] = (dbl4p){0.0, 0.0, 0.0, 0.0};
  for (int rb = j + 16 * wv; rb < m; rb += NW * 16) {
    const int ra = rb + li;
    double pa[KMAX / 4];
#pragma unroll
    for (int t = 0; t < KMAX / 4; t++)
      pa[t] = (ra < m && 4 * t + g < k1) ? a[(size_t)(j + 4 * t + g) * lda + ra] : 0.0;
    dbl4p qb[4];
#pragma unroll
    for (int u = 0; u < 4; u++) qb[u] = (dbl4p){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < KMAX / 4; t++)
#pragma unroll
      for (int u = 0; u < 4; u++)
        qb[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[t], E1[4 * t + g][16 * u + li], qb[u], 0, 0,
                                                     0);
    gram_tiles(qb, gt);
  }
  bt_gram_store(RT, gt);
  if (tid >= k1 && tid < KMAX) RT[tid][tid] = 1.0;
  lds_barrier();
  // 3. the reconstruction: SR = E0, SA = E1 (R1^-1), SW = TS (also its
  // scratch Y, and M out), SX = RT
  double* Mf = &TS[0][0];  // M(q, c) at Mf[q + c * KMAX]
  const int k = recon_core(
      E0, E1, TS, RT, Yof(L.Ts), L.rinv, L.dsg, L.tl, L.scv, &L.s_done, &L.s_why,
      [&](double(*S)[LDP]) {  // E1 holds R1^-1: S takes it before E1 is overwritten
        for (int e = tid; e < KMAX * KMAX; e += NW * 64) S[e & 63][e >> 6] = E1[e & 63][e >> 6];
        lds_barrier();
      },
      [&](double(*S)[LDP]) { load_slab(S, a, lda, j, j, j + k1, k1); }, a, lda, j, fjb, k1, tau_dm,
      tau + j, Mf);
  if (k < 0) return -1;
  __syncthreads();  // the top block the reconstruction wrote to A is read below
  for (int e = tid; e < KMAX * KMAX; e += NW * 64) E0[e & 63][e >> 6] = Mf[e];  // M, padded
  lds_barrier();
  // 4. V2 = P_bottom M: each wave its own 16-row blocks as in 2, P straight
  // into the A operand, M (E0) the B operand; the top block's rows take V1
  // (unit lower, written into A by the reconstruction), V2's rows are
  // written into the panel, and V^T V accumulates from the product's layout
  // (summed into E1 in wave order).
#pragma unroll
  for (int p = 0; p < 10; p++) gt[p] = (dbl4p){0.0, 0.0, 0.0, 0.0};
  for (int rb = j + 16 * wv; rb < m; rb += NW * 16) {
    const int ra = rb + li;
    double pa[KMAX / 4];
#pragma unroll
    for (int t = 0; t < KMAX / 4; t++)
      pa[t] = (ra < m && 4 * t + g < k) ? a[(size_t)(j + 4 * t + g) * lda + ra] : 0.0;
    dbl4p vb[4];
#pragma unroll
    for (int u = 0; u < 4; u++) vb[u] = (dbl4p){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < KMAX / 4; t++)
#pragma unroll
      for (int u = 0; u < 4; u++)
        vb[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[t], E0[4 * t + g][16 * u + li], vb[u], 0, 0,
                                                     0);
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = rb + g + 4 * v, c = 16 * u + li, pr = r - j;
        double x = vb[u][v];
        if (c >= k || r >= m) x = 0.0;
        else if (pr < k) x = (pr < c) ? 0.0 : (pr == c ? 1.0 : a[(size_t)(j + c) * lda + r]);
        else a[(size_t)(j + c) * lda + r] = x;
        vb[u][v] = x;
      }
    gram_tiles(vb, gt);
  }
  bt_gram_store(E1, gt);
  // 5. tau_i = 2 / S_ii and T = (striu(S) + diag(S) / 2)^-1 from S = V^T V
  for (int e = tid; e < KMAX * KMAX; e += NW * 64) {
    const int r = e & 63, c = e >> 6;
    const double sv = E1[r][c];
    E0[r][c] = (r < k && c < k) ? ((r < c) ? sv : (r == c ? 0.5 * sv : 0.0)) : (r == c ? 1.0 : 0.0);
  }
  lds_barrier();
  if (tid < k) tau[j + tid] = 1.0 / E0[tid][tid];
  lu_inverse<false>(E0, E1, L.rinv, Yof(L.R1));  // E1 = T (upper)
  for (int e = tid; e < KMAX * KMAX; e += NW * 64) {
    const int r = e & 63, c = e >> 6;
    L.Ts[r + c * KMAX] = (r < k && c < k && r <= c) ? E1[r][c] : 0.0;
  }
  __syncthreads();  // V2 in A and T in LDS, for the trailing update
  return k;
}

// One DM-QR (dgeqrdm semantics, free columns) of the matrix it by the whole
// workgroup.  Returns info (0, or -13 for a non-finite entry, A untouched).
template <int NW>
__device__ __forceinline__ int bt_factor(const BatchItem& it, double* A, double* tau_all, int* jp_all,
                         int* nc_all, const BatchParams& P, BatchLds<NW>& L) {
  constexpr int NT = 64 * NW;
  const int tid = local_tid(), lane = tid & 63, wave = tid >> 6;
  const int m = it.m, n = it.n, lda = it.lda;
  double* a = A + it.a_off;
  double* tau = tau_all + it.tau_off;
  const int minmn = m < n ? m : n;
  // initial norms (src/dgeqrdm_work.c:670-682), as k_init_norms
  double mx = 0.0, bad = 0.0;
  for (int c = wave; c < n; c += NW) {
    const double nr = wave_nrm2(a + (size_t)c * lda, m);
    if (lane == 0) {
      L.work[c] = nr;
      L.wold[c] = nr;
    }
    if (!(nr <= DBL_MAX)) bad = 1.0;
    else mx = fmax(mx, nr);
  }
  for (int c = tid; c < n; c += NT) L.jp[c] = c + 1;
  const double maxnrm0 = block_max(mx, L.red);
  if (block_max(bad, L.red) != 0.0) return -13;
  double eta = 0.0;
  const int enable_stop = P.stop_mode;
  if (P.stop_mode == 1) eta = kEpsD * n;
  else if (P.stop_mode == 2) eta = kEpsD * sqrt((double)n);
  else if (P.stop_mode == 3) eta = P.eta3;
  eta *= maxnrm0;
  int j = 0, itn = 0;
  while (j < minmn) {
    // the workitem id afresh for each step: what the step derives from it is
    // not held in registers across the steps
    const int tid = local_tid(), lane = tid & 63, wave = tid >> 6;
    const int nsub = n - j;
    int kmax = P.nb < KMAX ? P.nb : KMAX;
    if (m - j < kmax) kmax = m - j;
    if (nsub < kmax) kmax = nsub;
    // --- candidates: the first kmax keys of work[j..n) in qsort/cmpStruct
    // order (value descending, position ascending), count > tau * max
    {
      double v = -1.0;
      int ix = 0x7fffffc0 + lane;  // a sorted list of padding keys
      for (int b = wave; 64 * b < nsub; b += NW) {
        const int t = 64 * b + lane;
        double bv = (t < nsub) ? L.work[j + t] : -1.0;
        int bi = (t < nsub) ? t : 0x40000000 + t;
        wave_sort64(bv, bi);
        wave_merge64(v, ix, bv, bi);
      }
      L.u.tk.lv[wave][lane] = v;
      L.u.tk.li[wave][lane] = ix;
      __syncthreads();
      block_merge_lists<NW>(L.u.tk.lv, L.u.tk.li, v, ix);
      if (wave == 0) {
        const double cv = L.u.tk.lv[0][lane];
        const int ci = L.u.tk.li[0][lane];
        L.st.cand_val[lane] = lane < kmax ? cv : 0.0;
        L.st.cand_idx[lane] = lane < kmax ? ci : -1;
        const double lim = P.tau_dm * L.u.tk.lv[0][0];
        const int nc = __popcll(__ballot(lane < kmax && cv > lim));
        if (lane == 0) {
          L.st.nc = nc;
          L.st.kmax = kmax;
          L.st.j = j;
          L.st.halt = 0;
          L.st.policy_on = 0;
          L.st.tau_step = P.tau_dm;
          L.st.delta_step = P.delta;
        }
      }
      __syncthreads();
      // the learned proposal (k_policy's arithmetic), and the inputs of a
      // training record
      if (P.policy || P.trace)
        policy_features(L.work + j, nsub, L.st.cand_val[0], j, m, n, L.feat, L.hist);
      if (P.policy && wave == 0) {
        const double cv = L.u.tk.lv[0][lane];
        const float2 y = policy_forward(P.policy, L.feat, L.hbuf);
        double tp = P.policy->tau_lo + (P.policy->tau_hi - P.policy->tau_lo) * (double)y.x;
        double dp = P.policy->delta_lo + (P.policy->delta_hi - P.policy->delta_lo) * (double)y.y;
        if (!(tp >= 0.0 && tp <= 1.0 && dp >= 0.0 && dp <= 1.0)) {
          tp = P.tau_dm;
          dp = P.delta;
        }
        const int nc2 = __popcll(__ballot(lane < kmax && cv > tp * L.u.tk.lv[0][0]));
        if (lane == 0) {
          L.st.nc = nc2;
          L.st.policy_on = 1;
          L.st.tau_step = tp;
          L.st.delta_step = dp;
        }
      }
      __syncthreads();
    }
    const int nc = L.st.nc;
    // --- cosines (src/dgeqrdm_work.c:365-379): Gram of the candidate
    // columns over rows j..m, scaled by the reciprocal norms in select_wave
    // (over all kmax candidates when a training record scores other tau)
    const int ng = P.trace ? kmax : nc;
    if (ng > 1) {
      const DevState& st = L.st;
      bt_gram<NW>(j, m, ng,
                  [&](int c, int r) { return a[(size_t)(j + st.cand_idx[c]) * lda + r]; }, L);
    }
    if (P.trace && wave == 0) {
      unsigned slot = 0;
      if (lane == 0) slot = atomicAdd(P.trace_count, 1u);
      slot = (unsigned)__builtin_amdgcn_readfirstlane((int)slot);
      if (slot < (unsigned)P.trace_cap)
        label_record(L.R1, L.st.cand_val, kmax, P.tau_dm, P.delta, L.Ts, L.feat, j, nsub,
                     P.tau_lo, P.tau_hi, P.delta_lo, P.delta_hi,
                     P.trace + (size_t)slot * POL_REC);
    }
    // --- greedy acceptance and the permute_marked plan
    if (wave == 0) select_wave(L.R1, select_inputs(&L.st, P.delta), m, n, INT_MAX, 1, L.perm, &L.st);
    __syncthreads();
    const int fjb = L.st.fjb, npair = L.st.npair;
    // --- column moves (whole columns; work and jpvt with them, not workold)
    if (npair > 0) {
      const int rb_rows = max(1, min(m, (8 * NT) / npair));
      for (int r0 = 0; r0 < m; r0 += rb_rows) {
        const int rows = min(rb_rows, m - r0), tot = npair * rows;
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
          const int e = tid + NT * q;
          v[q] = (e < tot) ? a[(size_t)L.perm[2 * (e / rows) + 1] * lda + r0 + e % rows] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; q++) {
          const int e = tid + NT * q;
          if (e < tot) a[(size_t)L.perm[2 * (e / rows)] * lda + r0 + e % rows] = v[q];
        }
      }
      double wv = 0.0;
      int jv = 0;
      if (tid < npair) {
        wv = L.work[L.perm[2 * tid + 1]];
        jv = L.jp[L.perm[2 * tid + 1]];
      }
      __syncthreads();
      if (tid < npair) {
        L.work[L.perm[2 * tid]] = wv;
        L.jp[L.perm[2 * tid]] = jv;
      }
      __syncthreads();
    }
    // --- dgeqr2_mia with truncation (src/dgeqrdm_work.c:735-738): the
    // Gram-based form on tall panels (T from its V^T V; the trailing update
    // also reaches the rejected in-block columns, as dlarf_ did inside the
    // reference's panel), else, or when it falls back, column by column
    int kk = -1, c0 = 0;
    if (fjb >= 2 && m - j >= P.gram_rows) {
      kk = bt_panel_gram<NW>(a, lda, j, m, fjb, P.tau_dm, tau, L);
      c0 = j + kk;
    }
    if (kk < 0) {
      kk = panel_exact_body<4>(a, lda, j, m, fjb, P.tau_dm, 1, tau + j, L.red);
      __syncthreads();
      c0 = j + fjb;
      // --- dlarft from S = V^T V (src/dgeqrdm_work.c:751-754)
      if (c0 < n) {
        bt_gram<NW>(j, m, kk, [&](int c, int r) { return bt_v(a, lda, j, kk, r, c, m); }, L);
        // R1 holds S with leading dimension KMAX + 1: repack to KMAX for the recurrence
        double sv[KMAX * KMAX / NT];
#pragma unroll
        for (int q = 0; q < KMAX * KMAX / NT; q++) {
          const int e = tid + NT * q;
          sv[q] = L.R1[(e & 63) + (e >> 6) * (KMAX + 1)];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < KMAX * KMAX / NT; q++) {
          const int e = tid + NT * q;
          L.R1[e] = sv[q];
          L.Ts[e] = 0.0;
        }
        if (tid < KMAX) L.ts[tid] = tid < kk ? tau[j + tid] : 0.0;
        __syncthreads();
        larft_recurrence(L.R1, L.Ts, L.ts, L.tmp, kk);
        __syncthreads();
      }
    }
    // --- dlarfb on columns c0..n (src/dgeqrdm_work.c:762-767)
    if (c0 < n) {
      bt_trailing<NW>(a, lda, j, m, kk, c0, n, L);
      __syncthreads();
    }
    // --- norm_update (src/dgeqrdm_work.c:80-109), as k_norm_update
    double nmx = 0.0;
    const int mrem = m - j - kk;
    for (int col = j + kk + wave; col < n; col += NW) {
      double wk = L.work[col];
      if (wk == 0.0) continue;
      const double* ac = a + (size_t)col * lda + j;
      const double x = (lane < kk) ? ac[lane] : 0.0;
      const double s = wave_sum(x * x);
      double temp = sqrt(fabs(s)) / wk;
      const double d2 = (temp + 1.0) * (1.0 - temp);
      temp = (0.0 >= d2) ? 0.0 : d2;
      const double d1 = wk / L.wold[col];
      const double temp2 = temp * (d1 * d1);
      if (temp2 <= P.tol3z) {
        wk = (mrem > 0) ? wave_nrm2(ac + kk, mrem) : 0.0;
        if (lane == 0) L.wold[col] = wk;
      } else {
        wk *= sqrt(temp);
      }
      if (lane == 0) L.work[col] = wk;
      if (wk == wk) nmx = fmax(nmx, wk);
    }
    const double maxnrm = block_max(nmx, L.red);
    // --- step end (src/dgeqrdm_work.c:740, :780-785)
    if (tid == 0) nc_all[it.col_off + itn] = kk;
    itn++;
    int done = (j + kk >= minmn) ? 1 : 0;
    if (enable_stop && maxnrm * sqrt((double)(n - j - kk)) <= eta) done = 1;
    j += kk;
    __syncthreads();
    if (done) break;
  }
  for (int c = tid; c < n; c += NT) jp_all[it.col_off + c] = L.jp[c];
  return 0;
}

// Persistent batch kernel: the workgroups factorise the matrices
// order[0 .. count), each taking the next index from queue[0].
template <int NW>
__global__ __launch_bounds__(64 * NW) void k_dm_batch(const BatchItem* __restrict__ items,
                                                      const int* __restrict__ order, int count,
                                                      int* queue, double* A, double* tau,
                                                      int* jpvt, int* ncols, int* info,
                                                      BatchParams P) {
  __shared__ BatchLds<NW> L;
  __shared__ int s_next;
  // Each workgroup takes the next matrix of order (decreasing flop count)
  // until none is left: list scheduling on the actual durations.  A
  // matrix's arithmetic does not depend on the workgroup that runs it.
  for (;;) {
    if (threadIdx.x == 0) s_next = atomicAdd(&queue[0], 1);
    __syncthreads();
    const int b = s_next;
    if (b >= count) break;  // every thread read the same s_next
    const int idx = order[b];
    const BatchItem it = items[idx];
    const int r = bt_factor<NW>(it, A, tau, jpvt, ncols, P, L);
    if (threadIdx.x == 0) info[idx] = r;
    __syncthreads();  // s_next is read by all before thread 0 takes the next
  }
  // the last workgroup out zeroes the queue for the next launch (a graph
  // replay included): every other workgroup has taken its last index
  if (threadIdx.x == 0 && atomicAdd(&queue[1], 1) == (int)gridDim.x - 1) {
    atomicExch(&queue[0], 0);
    atomicExch(&queue[1], 0);
  }
}

int launch_dm_batch(int waves, int grid, const void* items, const int* order, int count,
                     int* queue, double* A, double* tau, int* jpvt, int* ncols,
                     int* info, int nb, double delta, double tau_dm, int stop_mode, double eta3,
                     double tol3z, const void* policy, float* trace, unsigned* trace_count,
                     int trace_cap, const float* ranges, int gram_rows, hipStream_t s) {
  BatchParams P{delta, tau_dm, eta3, tol3z, nb, stop_mode,
                static_cast<const PolicyNet*>(policy), trace, trace_count, trace_cap,
                ranges ? ranges[0] : 0.f, ranges ? ranges[1] : 1.f,
                ranges ? ranges[2] : 0.f, ranges ? ranges[3] : 1.f,
                gram_rows > 0 ? gram_rows : INT_MAX};
  const BatchItem* it = static_cast<const BatchItem*>(items);
  // Only the 4-wave shape is instantiated: runs of k_dm_batch<8> on a
  // 4096-matrix batch ended in host-heap aborts of unidentified cause.
  if (waves != 4) return -1;
  hipLaunchKernelGGL(k_dm_batch<4>, dim3(grid), dim3(256), 0, s, it, order, count, queue, A,
                     tau, jpvt, ncols, info, P);
  return 0;
}

}  // namespace qrdm
