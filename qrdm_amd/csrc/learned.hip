// learned.hip -- the learned (tau, delta) proposal: a 34-64-64-2 perceptron
// with bfloat16 weights and FP32 accumulation, run once per DM step, and the
// kernels that train it (forward + backward + Adam).
//
// Inputs (34): a 32-bin histogram of the partial norms of the remaining
// columns divided by their maximum (bin floor(32 r), clamped to 31; counts
// divided by n - j), the step fraction j / min(m, n), and the remaining
// aspect ratio min((m - j) / (n - j), 8) / 8.  Outputs: two sigmoids, mapped
// to tau in [tau_lo, tau_hi] and delta in [delta_lo, delta_hi].  The panel's
// truncation test keeps the default tau (src/dgeqr2.c:176-177): a column the
// proposal admits but that is nearly dependent on the accepted ones is cut
// from the block there, so the default acceptance rules judge the proposal.
// A non-finite proposal falls back to the default parameters.

namespace qrdm {

constexpr int POL_IN = 34, POL_H = 64, POL_OUT = 2, POL_BINS = 32;
// parameter vector of training (FP32 master weights), in this order
constexpr int POL_W1 = 0, POL_B1 = POL_W1 + POL_H * POL_IN, POL_W2 = POL_B1 + POL_H,
              POL_B2 = POL_W2 + POL_H * POL_H, POL_W3 = POL_B2 + POL_H,
              POL_B3 = POL_W3 + POL_OUT * POL_H, POL_NPARAM = POL_B3 + POL_OUT;

// Layout of qrdm_policy (include/qrdm_amd.h).
struct PolicyNet {
  unsigned short w1[POL_H * POL_IN];  // [out][in], bfloat16 bits
  unsigned short w2[POL_H * POL_H];
  unsigned short w3[POL_OUT * POL_H];
  float b1[POL_H], b2[POL_H], b3[POL_OUT];
  float tau_lo, tau_hi, delta_lo, delta_hi;
};

__device__ __forceinline__ float bf16_bits_to_float(unsigned short h) {
  return __uint_as_float((unsigned)h << 16);
}

// bfloat16 rounding (nearest, ties to even) of a finite float, as bits.
__device__ __forceinline__ unsigned short float_to_bf16_bits(float x) {
  const unsigned u = __float_as_uint(x);
  return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float bf16_round(float x) {
  return bf16_bits_to_float(float_to_bf16_bits(x));
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// The 34 inputs of step j into feat (LDS), by the whole workgroup (every
// thread calls it; it ends with a barrier); work holds the n - j partial
// norms of the remaining columns, maxw their maximum; hist is LDS scratch of
// POL_BINS ints.
__device__ __forceinline__ void policy_features(const double* work, int nsub, double maxw, int j, int m, int n,
                                float* feat, int* hist) {
  const int tid = local_tid(), nt = blockDim.x;
  if (tid < POL_BINS) hist[tid] = 0;
  __syncthreads();
  for (int t = tid; t < nsub; t += nt) {
    const double r = maxw > 0.0 ? work[t] / maxw : 0.0;
    int b = (int)(r * (double)POL_BINS);
    b = b < 0 ? 0 : (b >= POL_BINS ? POL_BINS - 1 : b);
    atomicAdd(&hist[b], 1);
  }
  __syncthreads();
  if (tid < POL_BINS) feat[tid] = (float)hist[tid] / (float)nsub;
  if (tid == POL_BINS) feat[tid] = (float)j / (float)(m < n ? m : n);
  if (tid == POL_BINS + 1) {
    const double ar = (double)(m - j) / (double)(n - j);
    feat[tid] = (float)((ar < 8.0 ? ar : 8.0) / 8.0);
  }
  __syncthreads();
}

// Forward pass by one wave (lane = hidden unit); feat and hbuf (64 floats)
// in LDS.  Returns the two sigmoid outputs in every lane.
__device__ __forceinline__ float2 policy_forward(const PolicyNet* __restrict__ P, const float* feat,
                                 float* hbuf) {
  const int l = local_tid() & 63;
  float z1 = P->b1[l];
#pragma unroll 2
  for (int k = 0; k < POL_IN; k++) z1 = fmaf(bf16_bits_to_float(P->w1[l * POL_IN + k]), feat[k], z1);
  hbuf[l] = fmaxf(z1, 0.f);
  wave_lds_sync();
  float z2 = P->b2[l];
#pragma unroll 4
  for (int k = 0; k < POL_H; k++) z2 = fmaf(bf16_bits_to_float(P->w2[l * POL_H + k]), hbuf[k], z2);
  const float h2 = fmaxf(z2, 0.f);
  const float o0 = wave_sum_f(bf16_bits_to_float(P->w3[l]) * h2) + P->b3[0];
  const float o1 = wave_sum_f(bf16_bits_to_float(P->w3[POL_H + l]) * h2) + P->b3[1];
  wave_lds_sync();  // hbuf may be reused after return
  return make_float2(1.f / (1.f + expf(-o0)), 1.f / (1.f + expf(-o1)));
}

// The step's proposal: features of work[j..n) (the top-k stage has left the
// sorted candidates and kmax in st), the forward pass, and the candidate
// count for the proposed tau; the greedy acceptance reads delta_step.
constexpr int POL_THREADS = 1024;

__global__ __launch_bounds__(POL_THREADS) void k_policy(const double* __restrict__ work, int m,
                                                        int n, const PolicyNet* __restrict__ P,
                                                        DevState* st) {
  __shared__ float feat[POL_IN];
  __shared__ float hbuf[POL_H];
  __shared__ int hist[POL_BINS];
  if (dm_stopped(st)) return;
  const int lane = threadIdx.x;
  const int j = st->j, kmax = st->kmax;
  const double maxw = st->cand_val[0];
  policy_features(work + j, n - j, maxw, j, m, n, feat, hist);
  if (threadIdx.x >= 64) return;
  const float2 y = policy_forward(P, feat, hbuf);
  double tau = P->tau_lo + (P->tau_hi - P->tau_lo) * (double)y.x;
  double delta = P->delta_lo + (P->delta_hi - P->delta_lo) * (double)y.y;
  if (!(tau >= 0.0 && tau <= 1.0 && delta >= 0.0 && delta <= 1.0)) {
    tau = st->tau_step;  // fallback: the defaults (set at the loop's start)
    delta = st->delta_step;
  }
  const double v = st->cand_val[lane];
  const int nc = __popcll(__ballot(lane < kmax && v > tau * maxw));
  if (lane == 0) {
    st->nc = nc;
    st->policy_on = 1;
    st->tau_step = tau;
    st->delta_step = delta;
  }
}

void launch_policy(const double* work, int m, int n, const void* net, DevState* st,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_policy, dim3(1), dim3(POL_THREADS), 0, s, work, m, n,
                     static_cast<const PolicyNet*>(net), st);
}

// Batch inference: out[2 s + o] for samples x[34 s ..], one wave per sample.
__global__ __launch_bounds__(256) void k_policy_predict(const PolicyNet* __restrict__ P,
                                                        const float* __restrict__ x, int ns,
                                                        float* out) {
  __shared__ float feat[4][POL_IN];
  __shared__ float hbuf[4][POL_H];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int s = blockIdx.x * 4 + wave; s < ns; s += gridDim.x * 4) {
    if (lane < POL_IN) feat[wave][lane] = x[(size_t)s * POL_IN + lane];
    wave_lds_sync();
    const float2 y = policy_forward(P, feat[wave], hbuf[wave]);
    if (lane == 0) {
      out[2 * (size_t)s] = y.x;
      out[2 * (size_t)s + 1] = y.y;
    }
  }
}

// ---------------------------------------------------------------------------
// Training.  Loss = (1/N) sum_s sum_o (sigmoid(z3_o) - y_o)^2 over a
// minibatch of N samples, with the forward pass on bfloat16-rounded copies
// of the FP32 master weights (the gradient is applied to the masters).
// k_policy_grad: wave w takes samples w, w + W, ... (W waves in all); lane l
// is hidden unit l of both layers and accumulates the gradients of its rows
// in registers; each wave writes its partial gradient (POL_NPARAM floats) and
// partial loss.
constexpr int PG_WAVES = 4;

__global__ __launch_bounds__(64 * PG_WAVES) void k_policy_grad(
    const float* __restrict__ theta, const float* __restrict__ x, const float* __restrict__ y,
    int ns, float* partial, float* loss_partial) {
  __shared__ float W2s[POL_H][POL_H + 1];  // bf16-rounded W2 [out][in]
  __shared__ float hs[PG_WAVES][POL_H];
  __shared__ float ds[PG_WAVES][POL_H];
  __shared__ float xs[PG_WAVES][POL_IN];
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  for (int e = threadIdx.x; e < POL_H * POL_H; e += 64 * PG_WAVES)
    W2s[e / POL_H][e % POL_H] = bf16_round(theta[POL_W2 + e]);
  float w1r[POL_IN];
#pragma unroll
  for (int k = 0; k < POL_IN; k++) w1r[k] = bf16_round(theta[POL_W1 + l * POL_IN + k]);
  const float b1 = theta[POL_B1 + l], b2 = theta[POL_B2 + l];
  const float w30 = bf16_round(theta[POL_W3 + l]), w31 = bf16_round(theta[POL_W3 + POL_H + l]);
  const float b30 = theta[POL_B3], b31 = theta[POL_B3 + 1];
  __syncthreads();
  float g1[POL_IN], g2[POL_H];
#pragma unroll
  for (int k = 0; k < POL_IN; k++) g1[k] = 0.f;
#pragma unroll
  for (int k = 0; k < POL_H; k++) g2[k] = 0.f;
  float gb1 = 0.f, gb2 = 0.f, g30 = 0.f, g31 = 0.f, gb30 = 0.f, gb31 = 0.f, loss = 0.f;
  const float inv_n = 1.f / (float)ns;
  const int nw = gridDim.x * PG_WAVES;
  for (int s = blockIdx.x * PG_WAVES + wave; s < ns; s += nw) {
    if (l < POL_IN) xs[wave][l] = x[(size_t)s * POL_IN + l];
    wave_lds_sync();
    float z1 = b1;
#pragma unroll
    for (int k = 0; k < POL_IN; k++) z1 = fmaf(w1r[k], xs[wave][k], z1);
    const float h1 = fmaxf(z1, 0.f);
    hs[wave][l] = h1;
    wave_lds_sync();
    float z2 = b2;
#pragma unroll 8
    for (int k = 0; k < POL_H; k++) z2 = fmaf(W2s[l][k], hs[wave][k], z2);
    const float h2 = fmaxf(z2, 0.f);
    const float o0 = wave_sum_f(w30 * h2) + b30;
    const float o1 = wave_sum_f(w31 * h2) + b31;
    const float p0 = 1.f / (1.f + expf(-o0)), p1 = 1.f / (1.f + expf(-o1));
    const float e0 = p0 - y[2 * (size_t)s], e1 = p1 - y[2 * (size_t)s + 1];
    loss += (e0 * e0 + e1 * e1) * inv_n;
    const float d0 = 2.f * e0 * inv_n * p0 * (1.f - p0);
    const float d1 = 2.f * e1 * inv_n * p1 * (1.f - p1);
    g30 = fmaf(d0, h2, g30);
    g31 = fmaf(d1, h2, g31);
    gb30 += d0;
    gb31 += d1;
    const float dz2 = (z2 > 0.f) ? (w30 * d0 + w31 * d1) : 0.f;
    gb2 += dz2;
#pragma unroll
    for (int k = 0; k < POL_H; k++) g2[k] = fmaf(dz2, hs[wave][k], g2[k]);
    ds[wave][l] = dz2;
    wave_lds_sync();
    float dh1 = 0.f;
#pragma unroll 8
    for (int o = 0; o < POL_H; o++) dh1 = fmaf(W2s[o][l], ds[wave][o], dh1);
    const float dz1 = (z1 > 0.f) ? dh1 : 0.f;
    gb1 += dz1;
#pragma unroll
    for (int k = 0; k < POL_IN; k++) g1[k] = fmaf(dz1, xs[wave][k], g1[k]);
    wave_lds_sync();  // xs, hs, ds are rewritten by the next sample
  }
  float* out = partial + (size_t)(blockIdx.x * PG_WAVES + wave) * POL_NPARAM;
#pragma unroll
  for (int k = 0; k < POL_IN; k++) out[POL_W1 + l * POL_IN + k] = g1[k];
  out[POL_B1 + l] = gb1;
#pragma unroll
  for (int k = 0; k < POL_H; k++) out[POL_W2 + l * POL_H + k] = g2[k];
  out[POL_B2 + l] = gb2;
  out[POL_W3 + l] = g30;
  out[POL_W3 + POL_H + l] = g31;
  if (l == 0) {
    out[POL_B3] = gb30;  // gb30, gb31 are the same in every lane
    out[POL_B3 + 1] = gb31;
    loss_partial[blockIdx.x * PG_WAVES + wave] = loss;
  }
}

// grad = sum of the wave partials in wave order; with adam != 0, one Adam
// step (bias-corrected, step number t >= 1) on theta and its moments m1, m2.
// loss_out[0] = sum of the partial losses.
__global__ __launch_bounds__(256) void k_policy_update(float* theta, float* m1, float* m2,
                                                       const float* __restrict__ partial,
                                                       const float* __restrict__ loss_partial,
                                                       int nwaves, float* grad, float* loss_out,
                                                       int adam, float lr, float beta1,
                                                       float beta2, float eps, int t) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p == 0) {
    float ls = 0.f;
    for (int w = 0; w < nwaves; w++) ls += loss_partial[w];
    loss_out[0] = ls;
  }
  if (p >= POL_NPARAM) return;
  float g = 0.f;
  for (int w = 0; w < nwaves; w++) g += partial[(size_t)w * POL_NPARAM + p];
  if (grad) grad[p] = g;
  if (!adam) return;
  const float a = beta1 * m1[p] + (1.f - beta1) * g;
  const float b = beta2 * m2[p] + (1.f - beta2) * g * g;
  m1[p] = a;
  m2[p] = b;
  const float ah = a / (1.f - powf(beta1, (float)t));
  const float bh = b / (1.f - powf(beta2, (float)t));
  theta[p] -= lr * ah / (sqrtf(bh) + eps);
}

// The inference layout (bfloat16 weights) of theta.
__global__ __launch_bounds__(256) void k_policy_pack(const float* __restrict__ theta,
                                                     PolicyNet* net, float tau_lo, float tau_hi,
                                                     float delta_lo, float delta_hi) {
  for (int e = blockIdx.x * 256 + threadIdx.x; e < POL_NPARAM; e += gridDim.x * 256) {
    const float v = theta[e];
    if (e < POL_B1) net->w1[e - POL_W1] = float_to_bf16_bits(v);
    else if (e < POL_W2) net->b1[e - POL_B1] = v;
    else if (e < POL_B2) net->w2[e - POL_W2] = float_to_bf16_bits(v);
    else if (e < POL_W3) net->b2[e - POL_B2] = v;
    else if (e < POL_B3) net->w3[e - POL_W3] = float_to_bf16_bits(v);
    else net->b3[e - POL_B3] = v;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    net->tau_lo = tau_lo;
    net->tau_hi = tau_hi;
    net->delta_lo = delta_lo;
    net->delta_hi = delta_hi;
  }
}

// ---------------------------------------------------------------------------
// Training labels from the method's own steps.  For a step whose candidate
// Gram over the first kmax keys is in G (leading dimension KMAX + 1, columns
// unscaled) and whose sorted candidate norms are cand[0..kmax), each (tau,
// delta) of a grid is scored by the block it would reduce: the candidates
// above tau * max, the greedy |cos| < delta acceptance, then the accepted
// columns' Cholesky factor in acceptance order, cut at the first diagonal
// below tau_def * r_00 (the panel's truncation test with the default tau).
// The label is the grid point with the largest such block (ties: the
// defaults first, then larger tau, then smaller delta).  One wave; L is
// KMAX * (KMAX + 1) doubles of LDS scratch.
constexpr int LBL_NT = 5, LBL_ND = 6;
__device__ constexpr double kLblTau[LBL_NT] = {0.25, 0.15, 0.10, 0.05, 0.01};
__device__ constexpr double kLblDelta[LBL_ND] = {0.80, 0.90, 0.95, 0.98, 0.99, 0.999};

__device__ int label_block(const double* G, const double* cand, int kmax, double tau,
                           double delta, double tau_def, double* L) {
  const int t = local_tid() & 63;
  const double maxw = cand[0];
  const double vt = cand[t];
  const int nc = __popcll(__ballot(t < kmax && vt > tau * maxw));
  unsigned long long acc = 1ull;
  if (nc > 1) {
    const double rt = (t < nc) ? 1.0 / vt : 0.0;
    unsigned long long conf = 0ull;
    for (int q = 0; q < nc && q < t; q++) {
      const double c = fabs(G[t + q * (KMAX + 1)] * (1.0 / cand[q]) * rt);
      conf |= (c >= delta) ? (1ull << q) : 0ull;
    }
    for (int i = 1; i < nc; i++) {
      const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(conf >> 32), i);
      const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)conf, i);
      const unsigned long long ci = ((unsigned long long)hi << 32) | lo;
      if ((ci & acc) == 0ull) acc |= 1ull << i;
    }
  }
  const int f = __popcll(acc);
  // lane i < f: the i-th accepted candidate (acceptance order = index order)
  int ai = 0;
  {
    unsigned long long r = acc;
    for (int q = 0; q < t && r; q++) r &= r - 1;
    ai = r ? __ffsll((long long)r) - 1 : 0;
  }
  // row-per-lane Cholesky of G(a, a), cut by the truncation test
  double r0 = 0.0;
  for (int k = 0; k < f; k++) {
    const int ak = __builtin_amdgcn_readlane(ai, k);
    double d = 0.0;
    if (t == k) {
      d = G[ak + ak * (KMAX + 1)];
      for (int p = 0; p < k; p++) d -= L[k * (KMAX + 1) + p] * L[k * (KMAX + 1) + p];
      d = sqrt(d > 0.0 ? d : 0.0);
      L[k * (KMAX + 1) + k] = d;
    }
    wave_lds_sync();
    const double rk = rdlane(d, k);
    if (k == 0) r0 = rk;
    else if (!(rk >= tau_def * r0)) return k;
    if (t > k && t < f) {
      double x = G[ai + ak * (KMAX + 1)];
      for (int p = 0; p < k; p++) x -= L[t * (KMAX + 1) + p] * L[k * (KMAX + 1) + p];
      L[t * (KMAX + 1) + k] = rk > 0.0 ? x / rk : 0.0;
    }
    wave_lds_sync();
  }
  return f;
}

// Writes the step's record (POL_REC floats: the 34 inputs, the label
// normalised to [0, 1] over the output ranges, the default's block, the
// label's block, j, n - j) to rec.  feat must hold the inputs.
constexpr int POL_REC = 40;
__device__ __forceinline__ void label_record(const double* G, const double* cand, int kmax, double tau_def,
                             double delta_def, double* L, const float* feat, int j, int nsub,
                             float tau_lo, float tau_hi, float delta_lo, float delta_hi,
                             float* rec) {
  const int t = local_tid() & 63;
  const int eff_def = label_block(G, cand, kmax, tau_def, delta_def, tau_def, L);
  int best = eff_def;
  double bt = tau_def, bd = delta_def;
  for (int a = 0; a < LBL_NT; a++)
    for (int b = 0; b < LBL_ND; b++) {
      const int e = label_block(G, cand, kmax, kLblTau[a], kLblDelta[b], tau_def, L);
      if (e > best) {
        best = e;
        bt = kLblTau[a];
        bd = kLblDelta[b];
      }
    }
  if (t < POL_IN) rec[t] = feat[t];
  if (t == 0) {
    rec[POL_IN] = (float)((bt - tau_lo) / (tau_hi - tau_lo));
    rec[POL_IN + 1] = (float)((bd - delta_lo) / (delta_hi - delta_lo));
    rec[POL_IN + 2] = (float)eff_def;
    rec[POL_IN + 3] = (float)best;
    rec[POL_IN + 4] = (float)j;
    rec[POL_IN + 5] = (float)nsub;
  }
}

constexpr int PG_GRID = 64;  // workgroups of k_policy_grad

int policy_nparam() { return POL_NPARAM; }
size_t policy_net_bytes() { return sizeof(PolicyNet); }
int policy_partials() { return PG_GRID * PG_WAVES; }

void launch_policy_grad(const float* theta, const float* x, const float* y, int ns,
                        float* partial, float* loss_partial, hipStream_t s) {
  hipLaunchKernelGGL(k_policy_grad, dim3(PG_GRID), dim3(64 * PG_WAVES), 0, s, theta, x, y, ns,
                     partial, loss_partial);
}

void launch_policy_update(float* theta, float* m1, float* m2, const float* partial,
                          const float* loss_partial, float* grad, float* loss_out, int adam,
                          float lr, float beta1, float beta2, float eps, int t, hipStream_t s) {
  hipLaunchKernelGGL(k_policy_update, dim3((POL_NPARAM + 255) / 256), dim3(256), 0, s, theta, m1,
                     m2, partial, loss_partial, PG_GRID * PG_WAVES, grad, loss_out, adam, lr,
                     beta1, beta2, eps, t);
}

void launch_policy_pack(const float* theta, void* net, float tau_lo, float tau_hi,
                        float delta_lo, float delta_hi, hipStream_t s) {
  hipLaunchKernelGGL(k_policy_pack, dim3((POL_NPARAM + 255) / 256), dim3(256), 0, s, theta,
                     static_cast<PolicyNet*>(net), tau_lo, tau_hi, delta_lo, delta_hi);
}

void launch_policy_predict(const void* net, const float* x, int ns, float* out, hipStream_t s) {
  const int grid = (ns + 3) / 4 < 1024 ? (ns + 3) / 4 : 1024;
  hipLaunchKernelGGL(k_policy_predict, dim3(grid > 0 ? grid : 1), dim3(256), 0, s,
                     static_cast<const PolicyNet*>(net), x, ns, out);
}

}  // namespace qrdm
