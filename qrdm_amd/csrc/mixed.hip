// mixed.hip -- the trailing update C := C - V T^T V^T C with its two large
// products on the FP32 or split-BF16 matrix cores (FP32 accumulation).
//
// Only the products change precision: C stays FP64 in HBM (each value is
// rounded to FP32 as it is loaded and the FP32 product is subtracted from the
// FP64 value), T^T (sum of the partials) stays FP64 (k_wsum), and the panel,
// the norms and the cosine test are untouched.  Two forms:
//   TRAIL_FP32   v_mfma_f32_16x16x4f32 on FP32-rounded operands;
//   TRAIL_BF16X3 each FP32 operand x split as x = hi + lo + O(2^-17 |x|)
//                with hi, lo bfloat16 (round to nearest), and the product
//                taken as hi*hi + hi*lo + lo*hi (lo*lo, below 2^-16 of the
//                product, is dropped) on v_mfma_f32_16x16x32_bf16.
// The kernels have the geometry of k_gemm1 / k_gemm2 (kernels.hip); the row
// chunk is 32 rows (the bf16 MFMA's k), so k_gemm1's slabs (multiples of 32
// rows) hold whole chunks.
// Output layout of the FP32-accumulating 16x16 forms: lane l holds
// D[4*(l/16) + v][l%16] in register v (consecutive rows per lane group).

namespace qrdm {

typedef float flt4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ constexpr int mx_row(int g, int v) { return 4 * g + v; }

// x = hi + lo + O(2^-17 |x|), hi and lo bfloat16 rounded to nearest (the
// difference x - hi is exact in FP32).
__device__ __forceinline__ void split_bf16(const float (&x)[8], bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const __bf16 h = (__bf16)x[i];
    hi[i] = h;
    lo[i] = (__bf16)(x[i] - (float)h);
  }
}

constexpr int MX_CH = 32;  // rows per chunk

// Partial products Wp[y] = V^T C over row slab y (k_gemm1's split), FP32
// accumulation.  Workgroup: 64 * G1_NWG columns (as k_gemm1), wave w the
// 64-column tile G1_NWG*blockIdx.x + w; lane group g = lane/16 holds rows
// 8g..8g+7 of a 32-row chunk.  The V chunk is converted once per workgroup
// into LDS (double buffered); each wave streams its C tile through registers
// one chunk ahead.
template <int PREC, bool VEC>
__global__ __launch_bounds__(64 * G1_NWG) void k_gemm1_mx(const double* __restrict__ A, int lda, int j,
                                                  int m, int c0, int ncol,
                                                  const double* __restrict__ V, int ldv,
                                                  int slab, double* Wp, int ldw,
                                                  const DevState* st) {
  constexpr bool F32 = PREC == TRAIL_FP32;
  __shared__ __align__(16) float Vf[F32 ? 2 : 1][KMAX][MX_CH + 4];
  __shared__ __align__(16) __bf16 Vh[F32 ? 1 : 2][KMAX][MX_CH + 8];
  __shared__ __align__(16) __bf16 Vl[F32 ? 1 : 2][KMAX][MX_CH + 8];
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
  const int re = max(min(rs + slab, m), rs);
  const int cbw = blockIdx.x * 64 * G1_NWG + wave * 64;
  // columns past ncol read a valid column instead; they are never stored
  const double* Cc[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int c = cbw + 16 * u + li;
    Cc[u] = A + (size_t)(c0 + (c < ncol ? c : ncol - 1)) * lda;
  }
  flt4 acc[4][4];
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int u = 0; u < 4; u++) acc[t][u] = (flt4){0.f, 0.f, 0.f, 0.f};
  const int nch = (re - rs + MX_CH - 1) / MX_CH;
  // last 4-row group that starts inside a column (lda % 4 == 0 on the vector path)
  const int rclamp = ((m + 3) & ~3) - 4;
  auto load8 = [&](const double* col, int r, double (&x)[8]) {
    if constexpr (VEC) {
      const dbl4 a = *reinterpret_cast<const dbl4*>(col + min(r, rclamp));
      const dbl4 b = *reinterpret_cast<const dbl4*>(col + min(r + 4, rclamp));
#pragma unroll
      for (int q = 0; q < 4; q++) {
        x[q] = a[q];
        x[4 + q] = b[q];
      }
    } else {
#pragma unroll
      for (int q = 0; q < 8; q++) x[q] = col[min(r + q, m - 1)];
    }
  };
  // V chunk loader: thread tid brings rows 8*(tid&3).. of the V columns
  // (tid>>2) + 16 * G1_NWG * h, h < VPT; rows >= re (past m, where V may hold
  // anything) are stored as zero
  constexpr int VPT = 4 / G1_NWG;
  const int vq = 8 * (tid & 3);
  auto vcolp = [&](int h) { return V + (size_t)((tid >> 2) + 16 * G1_NWG * h) * ldv; };
  auto store_v = [&](int buf, int c, int h, const double (&x)[8]) {
    const int vcol = (tid >> 2) + 16 * G1_NWG * h;
    const int r = rs + MX_CH * c + vq;
    float f[8];
#pragma unroll
    for (int q = 0; q < 8; q++) f[q] = (r + q < re) ? (float)x[q] : 0.f;
    if constexpr (F32) {
      *reinterpret_cast<flt4*>(&Vf[buf][vcol][vq]) = (flt4){f[0], f[1], f[2], f[3]};
      *reinterpret_cast<flt4*>(&Vf[buf][vcol][vq + 4]) = (flt4){f[4], f[5], f[6], f[7]};
    } else {
      bf16x8 h, l;
      split_bf16(f, h, l);
      *reinterpret_cast<bf16x8*>(&Vh[buf][vcol][vq]) = h;
      *reinterpret_cast<bf16x8*>(&Vl[buf][vcol][vq]) = l;
    }
  };
  double ca[4][8];
  if (nch > 0) {
#pragma unroll
    for (int h = 0; h < VPT; h++) {
      double x[8];
      load8(vcolp(h), rs + vq, x);
      store_v(0, 0, h, x);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) load8(Cc[u], rs + 8 * g, ca[u]);
  }
  __syncthreads();
#pragma clang loop unroll(disable)
  for (int c = 0; c < nch; c++) {
    const int buf = c & 1;
    const int cn_idx = min(c + 1, nch - 1);  // chunk c + 1 (clamped, then unused) in flight
    double vnext[VPT][8];
#pragma unroll
    for (int h = 0; h < VPT; h++) load8(vcolp(h), rs + MX_CH * cn_idx + vq, vnext[h]);
    double cn[4][8];
#pragma unroll
    for (int u = 0; u < 4; u++) load8(Cc[u], rs + MX_CH * cn_idx + 8 * g, cn[u]);
    if (rs + MX_CH * c + MX_CH > re) {
      // the chunk runs past re: its C rows there may be anything (padding)
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const bool ok = rs + MX_CH * c + 8 * g + q < re;
#pragma unroll
        for (int u = 0; u < 4; u++) ca[u][q] = ok ? ca[u][q] : 0.0;
      }
    }
    if constexpr (F32) {
      // k-step q: lane group g supplies row 8g + q of V and of C
      float vf[4][8], cf[4][8];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const flt4 a = *reinterpret_cast<const flt4*>(&Vf[buf][16 * t + li][8 * g]);
        const flt4 b = *reinterpret_cast<const flt4*>(&Vf[buf][16 * t + li][8 * g + 4]);
#pragma unroll
        for (int q = 0; q < 4; q++) {
          vf[t][q] = a[q];
          vf[t][4 + q] = b[q];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int q = 0; q < 8; q++) cf[u][q] = (float)ca[u][q];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int q = 0; q < 8; q++)
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
          for (int u = 0; u < 4; u++)
            acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x4f32(vf[t][q], cf[u][q], acc[t][u], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
    } else {
      // one k-step of 32: lane group g supplies rows 8g..8g+7
      bf16x8 vh[4], vl[4], ch[4], cl[4];
#pragma unroll
      for (int t = 0; t < 4; t++) {
        vh[t] = *reinterpret_cast<const bf16x8*>(&Vh[buf][16 * t + li][8 * g]);
        vl[t] = *reinterpret_cast<const bf16x8*>(&Vl[buf][16 * t + li][8 * g]);
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        float f[8];
#pragma unroll
        for (int q = 0; q < 8; q++) f[q] = (float)ca[u][q];
        split_bf16(f, ch[u], cl[u]);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int u = 0; u < 4; u++) {
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vl[t], ch[u], acc[t][u], 0, 0, 0);
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vh[t], cl[u], acc[t][u], 0, 0, 0);
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vh[t], ch[u], acc[t][u], 0, 0, 0);
        }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int h = 0; h < VPT; h++) store_v(buf ^ 1, cn_idx, h, vnext[h]);
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int q = 0; q < 8; q++) ca[u][q] = cn[u][q];
    __syncthreads();
  }
  // partial tile: W row (V column) 16t + mx_row(g, v), column cbw + 16u + li
  double* out = Wp + (size_t)blockIdx.y * KMAX * ldw;
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int cc = cbw + 16 * u + li;
    if (cc < ncol) {
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int v = 0; v < 4; v++)
          out[(size_t)(16 * t + mx_row(g, v)) * ldw + cc] = (double)acc[t][u][v];
    }
  }
}

// C -= V W with the product in FP32 (C^T -= W^T V^T, as k_gemm2): workgroup
// tile 64 rows x 64 columns, wave w columns 16w..16w+15; the V tile is
// converted into LDS once per workgroup (row-major: [row][k]).
template <int PREC>
__global__ __launch_bounds__(256) void k_gemm2_mx(double* A, int lda, int j, int m, int c0,
                                                  int ncol, const double* __restrict__ V, int ldv,
                                                  const double* __restrict__ W, int ldw,
                                                  const DevState* st) {
  constexpr bool F32 = PREC == TRAIL_FP32;
  __shared__ __align__(16) float Vf[F32 ? 64 : 1][KMAX + 4];
  __shared__ __align__(16) __bf16 Vh[F32 ? 1 : 64][KMAX + 8];
  __shared__ __align__(16) __bf16 Vl[F32 ? 1 : 64][KMAX + 8];
  if (st) {
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
  // C values this lane updates: rows rbase + 16t + li, columns cbase + mx_row(g, v)
  double cval[4][4];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int cc = cbase + mx_row(g, v);
    const double* col = A + (size_t)(c0 + (cc < ncol ? cc : 0)) * lda;
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int r = rbase + 16 * t + li;
      cval[t][v] = (cc < ncol && r < m) ? col[r] : 0.0;
    }
  }
  {
    // thread: row r = tid & 63 of the tile, k = 4*kg + 16p + q (p, q < 4) for
    // kg = tid >> 6; rows past m are zero
    const int r = tid & 63, kg = tid >> 6;
    const bool rok = rbase + r < m;
    float f[16];
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int k = 16 * p + 4 * kg + q;
        f[4 * p + q] = rok ? (float)V[(size_t)k * ldv + rbase + r] : 0.f;
      }
    if constexpr (F32) {
#pragma unroll
      for (int p = 0; p < 4; p++)
        *reinterpret_cast<flt4*>(&Vf[r][16 * p + 4 * kg]) =
            (flt4){f[4 * p], f[4 * p + 1], f[4 * p + 2], f[4 * p + 3]};
    } else {
#pragma unroll
      for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int k = 16 * p + 4 * kg + q;
          const __bf16 h = (__bf16)f[4 * p + q];
          Vh[r][k] = h;
          Vl[r][k] = (__bf16)(f[4 * p + q] - (float)h);
        }
    }
  }
  __syncthreads();
  flt4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; t++) acc[t] = (flt4){0.f, 0.f, 0.f, 0.f};
  const int wc = cbase + li;
  const bool wok = wc < ncol;
  if constexpr (F32) {
#pragma unroll 4
    for (int k0 = 0; k0 < KMAX; k0 += 4) {
      const float a = wok ? (float)W[(size_t)(k0 + g) * ldw + wc] : 0.f;
#pragma unroll
      for (int t = 0; t < 4; t++)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, Vf[16 * t + li][k0 + g], acc[t], 0, 0, 0);
    }
  } else {
#pragma unroll
    for (int ks = 0; ks < KMAX / 32; ks++) {
      const int kb = 32 * ks + 8 * g;
      float wf[8];
#pragma unroll
      for (int q = 0; q < 8; q++) wf[q] = wok ? (float)W[(size_t)(kb + q) * ldw + wc] : 0.f;
      bf16x8 wh, wl;
      split_bf16(wf, wh, wl);
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const bf16x8 vh = *reinterpret_cast<const bf16x8*>(&Vh[16 * t + li][kb]);
        const bf16x8 vl = *reinterpret_cast<const bf16x8*>(&Vl[16 * t + li][kb]);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wl, vh, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh, vl, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wh, vh, acc[t], 0, 0, 0);
      }
    }
  }
  // acc[t][v]: column cbase + mx_row(g, v), row rbase + 16t + li
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int cc = cbase + mx_row(g, v);
    if (cc >= ncol) continue;
    double* col = A + (size_t)(c0 + cc) * lda;
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int r = rbase + 16 * t + li;
      if (r < m) col[r] = cval[t][v] - (double)acc[t][v];
    }
  }
}

void launch_gemm1_prec(int prec, double* A, int lda, int j, int m, int c0, int n,
                       const double* V, int ldv, double* Wp, int nsplit, hipStream_t s,
                       const DevState* st) {
  if (prec == TRAIL_FP64) {
    launch_gemm1(A, lda, j, m, c0, n, V, ldv, Wp, nsplit, s, st);
    return;
  }
  const int ncol = n - c0, mrow = m - j;
  if (ncol <= 0 || mrow <= 0) return;
  const G1Geom g = gemm1_geom(j, m, c0, n, nsplit);  // k_gemm1's geometry
  const bool aligned = (lda % 4 == 0) && (ldv % 4 == 0) && m >= 4 &&
                       ((reinterpret_cast<uintptr_t>(A) & 31) == 0) &&
                       ((reinterpret_cast<uintptr_t>(V) & 31) == 0);
  const dim3 grid(g.cgroups, g.ns);
#define QRDM_G1MX(P, VEC)                                                                     \
  hipLaunchKernelGGL((k_gemm1_mx<P, VEC>), grid, dim3(64 * G1_NWG), 0, s, A, lda, j, m, c0, ncol, V, \
                     ldv, g.slab, Wp, n, st)
  if (prec == TRAIL_FP32) {
    if (aligned) QRDM_G1MX(TRAIL_FP32, true);
    else QRDM_G1MX(TRAIL_FP32, false);
  } else {
    if (aligned) QRDM_G1MX(TRAIL_BF16X3, true);
    else QRDM_G1MX(TRAIL_BF16X3, false);
  }
#undef QRDM_G1MX
}

void launch_apply_prec(int prec, double* A, int lda, int j, int m, int c0, int n,
                       const double* V, int ldv, const double* W, hipStream_t s,
                       const DevState* st) {
  if (prec == TRAIL_FP64) {
    launch_larfb_apply(A, lda, j, m, c0, n, V, ldv, W, s, st);
    return;
  }
  const int ncol = n - c0, mrow = m - j;
  if (ncol <= 0 || mrow <= 0) return;
  const dim3 grid((mrow + 63) / 64, (ncol + 63) / 64);
  if (prec == TRAIL_FP32)
    hipLaunchKernelGGL(k_gemm2_mx<TRAIL_FP32>, grid, dim3(256), 0, s, A, lda, j, m, c0, ncol, V,
                       ldv, W, n, st);
  else
    hipLaunchKernelGGL(k_gemm2_mx<TRAIL_BF16X3>, grid, dim3(256), 0, s, A, lda, j, m, c0, ncol,
                       V, ldv, W, n, st);
}

}  // namespace qrdm
