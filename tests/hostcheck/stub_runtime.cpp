// Host-only stand-in for the HIP runtime and for the kernel launchers of
// libqrdm_amd.so, so that the library's host code (driver.hip, compiled for
// the host with AddressSanitizer) runs on a CPU with no GPU.
//
// "Device" memory is host heap memory (malloc, checked by ASan), streams run
// every call at once and in order, events and graphs are records.  The
// launchers do no arithmetic: they model only what the host reads back of
// the device state (DevState), i.e. the step protocol of the DM loop --
//   k_dm_begin        j, it, done;
//   the panels        fjb_cmp (at least one column per step), c0; a tall
//                     step's Gram-based panel falls back now and then (halt,
//                     resumed by the host with the exact panel);
//   the update        pend (the step's end is due);
//   k_step_end/k_topk the pending end: ncols_dev[it] = kk, it += 1, j += kk,
//                     done = j >= min(m, n);
// and they write the device arrays the real kernels write, over the extents
// those write (ncols, jpvt, tau, info of the batch), so that an extent the
// host sized wrongly is an ASan error here.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

#include "../../include/qrdm_amd.h"
#include "../../qrdm_amd/csrc/qrdm_internal.h"

struct ihipStream_t {
  bool capturing = false;
  std::vector<std::function<void()>>* rec = nullptr;
};
struct ihipEvent_t {
  int unused = 0;
};
struct ihipGraph {
  std::vector<std::function<void()>> nodes;
};
struct hipGraphExec {
  std::vector<std::function<void()>> nodes;
};

namespace {
int g_launches = 0;
hipError_t g_last = hipSuccess;
// a launch on a capturing stream is recorded in its graph, else run now
void submit(hipStream_t s, std::function<void()> f) {
  g_launches++;
  if (s && s->capturing) s->rec->push_back(std::move(f));
  else f();
}
}  // namespace

extern "C" {
// ---- registration of the library's device variable (its fixed workspace)
void** __hipRegisterFatBinary(const void*) {
  static void* h = nullptr;
  return &h;
}
void __hipUnregisterFatBinary(void**) {}
void __hipRegisterVar(void**, void*, char*, char*, int, size_t, int, int) {}
void __hipRegisterFunction(void**, const void*, char*, const char*, unsigned int, void*, void*,
                           void*, void*, int*) {}
int stub_launch_count() { return g_launches; }
}

// ---- runtime API (the declarations of hip_runtime_api.h)
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
  *v = (a == hipDeviceAttributeMultiprocessorCount) ? 256 : 0;
  return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetLastError() {
  hipError_t e = g_last;
  g_last = hipSuccess;
  return e;
}
const char* hipGetErrorString(hipError_t) { return "stub error"; }
hipError_t hipMalloc(void** p, size_t n) {
  *p = malloc(n ? n : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
  free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned int) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  if (n) memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
  return hipMemcpy(d, s, n, k);
}
hipError_t hipMemcpy2D(void* d, size_t dp, const void* s, size_t sp, size_t w, size_t h,
                       hipMemcpyKind) {
  for (size_t r = 0; r < h; r++) memmove((char*)d + r * dp, (const char*)s + r * sp, w);
  return hipSuccess;
}
hipError_t hipMemset(void* p, int v, size_t n) {
  memset(p, v, n);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void* p, int v, size_t n, hipStream_t) { return hipMemset(p, v, n); }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
  *s = new ihipStream_t;
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = new ihipEvent_t;
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { return e ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipEventSynchronize(hipEvent_t e) { return e ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipEventElapsedTime(float* t, hipEvent_t, hipEvent_t) {
  *t = 1.0f;
  return hipSuccess;
}
hipError_t hipGetSymbolAddress(void** p, const void* sym) {
  *p = const_cast<void*>(sym);  // the host copy of the array stands for the device one
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  s->capturing = true;
  s->rec = new std::vector<std::function<void()>>;
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
  *g = new ihipGraph;
  (*g)->nodes = std::move(*s->rec);
  delete s->rec;
  s->rec = nullptr;
  s->capturing = false;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  *x = new hipGraphExec;
  (*x)->nodes = g->nodes;
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t) {
  for (auto& f : x->nodes) f();
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t x) {
  delete x;
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}

// ---- the launchers' model of the DM step protocol
namespace qrdm {
namespace {
int minmn(int m, int n) { return m < n ? m : n; }
void put_maxnrm(DevState* st, double v) { memcpy(&st->maxnrm_bits, &v, sizeof v); }
// the panel keeps all fjb columns but now and then one fewer (truncation)
int kept(int j, int fjb) { return (j % 3 == 0 && fjb > 1) ? fjb - 1 : fjb; }
void end_step(DevState* st, int* ncols_dev, int m, int n) {
  if (!st->pend) return;
  st->pend = 0;
  const int kk = st->fjb_cmp, j = st->j;
  st->last.j = j;
  st->last.fjb = st->fjb;
  st->last.fjb_cmp = kk;
  st->last.kmax = st->kmax;
  ncols_dev[st->it] = kk;
  st->it += 1;
  st->j = j + kk;
  st->done = (j + kk >= minmn(m, n)) ? 1 : 0;
  st->kk_prev = kk;
}
// the kernels' step_runs: stopped, or the other panel form
bool runs(const DevState* st, int mode) {
  if (st->done || st->halt) return false;
  if (mode == RUN_EXACT) return !st->fast || (st->flags & 2);
  if (mode == RUN_FAST) return st->fast && !(st->flags & 2);
  return true;
}
}  // namespace

int topk_blocks(int n) { return (n + 255) / 256; }
int vgram_exact_slabs(int j0, int m) {
  const int ns = (m - j0 + 255) / 256;
  return ns > 0 ? ns : 1;
}
int fused_slabs(int, int, int, int) { return 4; }
int fused_max_slabs() { return 16; }
int apply_perm_blocks(int m) { return (m + 255) / 256; }
int wide_sort_keys(int n) { return ((n + 1023) / 1024) * KW; }
int policy_nparam() { return 64 * 34 + 64 + 64 * 64 + 64 + 2 * 64 + 2; }
size_t policy_net_bytes() { return sizeof(qrdm_policy); }
int policy_partials() { return 64; }

void launch_init_norms(const double*, int, int r0, int m, int n, double* work, double* workold,
                       DevState* st, hipStream_t s) {
  submit(s, [=] {
    for (int c = 0; c < n; c++) work[c] = workold[c] = 1.0;
    st->flags = 0;
    put_maxnrm(st, 1.0);
    (void)r0;
    (void)m;
  });
}
void launch_dm_begin(DevState* st, int j, int mn, double tau, double delta, int fused,
                     hipStream_t s) {
  submit(s, [=] {
    st->fused = fused;
    st->kk_prev = 0;
    st->tau_step = tau;
    st->delta_step = delta;
    st->pend = 0;
    st->halt = 0;
    st->j = j;
    st->it = 0;
    st->done = (j >= mn) ? 1 : 0;
    st->flags &= ~2;
  });
}
void launch_topk(const double* work, int, int m, int n, int nb, double, double*, int*, unsigned*,
                 DevState* st, int* ncols_dev, int, double, hipStream_t s) {
  submit(s, [=] {
    end_step(st, ncols_dev, m, n);
    if (!runs(st, RUN_ALWAYS)) return;
    int k = nb < KMAX ? nb : KMAX;
    k = std::min(k, std::min(m - st->j, n - st->j));
    volatile double x = work[n - 1];  // the kernel reads work[j..n)
    (void)x;
    st->kmax = k;
    st->nc = k;
    st->flags &= ~2;
  });
}
void launch_gram_select(double*, int, int, int m, int, double*, double*, double, int fast_rows,
                        int exact_queued, int*, unsigned*, DevState* st, const StepBufs&,
                        hipStream_t s) {
  submit(s, [=] {
    if (!runs(st, RUN_ALWAYS)) return;
    const int j = st->j;
    st->fjb = (j % 5 == 1 && st->kmax > 1) ? st->kmax - 1 : st->kmax;
    st->npair = 0;
    st->nswap = 0;
    st->fast = (m - j >= fast_rows && st->fjb >= 2) ? 1 : 0;
    st->c0 = j + st->fjb;
    if (!exact_queued && !st->fast) st->halt = 1;
  });
}
void launch_apply_perm(double*, int, int, double*, int*, const int*, DevState*, const StepBufs&,
                       hipStream_t) {}
void launch_perm_chol1(double*, int, int, double*, int*, const int*, const double*, double,
                       double*, double*, DevState*, const StepBufs&, hipStream_t) {}
void launch_panel_fast(double*, int, int, int, double, double* tau, double*, double*, double*,
                       double*, double*, double*, const StepBufs&, unsigned*, DevState* st,
                       hipStream_t s) {
  submit(s, [=] {
    if (!runs(st, RUN_FAST)) return;
    const int j = st->j;
    if ((j / 64) % 5 == 2) {  // the Gram-based form falls back: the host resumes the step
      st->flags |= 2;
      st->halt = 1;
      return;
    }
    st->fjb_cmp = kept(j, st->fjb);
    for (int c = 0; c < st->fjb_cmp; c++) tau[j + c] = 0.5;
  });
}
void launch_panel_exact(double*, int, int j, int m, int fjb, double, int, double* tau,
                        DevState* st, hipStream_t s, int mode) {
  submit(s, [=] {
    if (mode == RUN_ALWAYS) {  // host j and fjb, tau from its first entry
      st->fjb_cmp = kept(j, fjb);
      for (int c = 0; c < st->fjb_cmp; c++) tau[c] = 0.5;
      return;
    }
    if (!runs(st, RUN_EXACT)) return;
    const int jj = st->j;
    st->fjb_cmp = kept(jj, st->fjb);
    st->c0 = jj + st->fjb;
    for (int c = 0; c < st->fjb_cmp; c++) tau[jj + c] = 0.5;
    (void)m;
  });
}
void launch_build_v(const double*, int, int, int, int, const DevState*, double*, int, hipStream_t,
                    int) {}
void launch_gram(const double*, int, int, int, const int*, int, const int*, double*, double*,
                 hipStream_t, const DevState*, int) {}
void launch_larft(const double*, const double*, const DevState*, double*, hipStream_t, int, int) {}
void launch_vgram_exact(const double*, int, int, int, const DevState*, const StepBufs&, double*,
                        hipStream_t) {}
void launch_larfb(double*, int, int, int, int, int, const double*, int, const double*, int,
                  double*, double*, int, hipStream_t) {}
void launch_gemm1_prec(int, double*, int, int, int, int, int, const double*, int, double*, int,
                       hipStream_t, const DevState*) {}
void launch_apply_prec(int, double*, int, int, int, int, int, const double*, int, const double*,
                       hipStream_t, const DevState*) {}
void launch_wsum(int, int, int, int, const double*, int, const double*, double*, int, hipStream_t,
                 const DevState*) {}
void launch_fused(double*, int, int, int, int, int, const StepBufs&, double*, int, int, int, int,
                  hipStream_t, const DevState*) {}
void launch_norm_update(const double*, int, int, int, int n, double* work, double*, DevState* st,
                        double, hipStream_t s) {
  submit(s, [=] {
    volatile double x = work[n - 1];
    (void)x;
    put_maxnrm(st, 0.5);
    if (runs(st, RUN_ALWAYS)) st->pend = 1;
  });
}
void launch_wsum_norm(double*, int, int, int, int n, const double*, const double*, int,
                      const StepBufs&, double* work, double*, DevState* st, double,
                      hipStream_t s) {
  submit(s, [=] {
    volatile double x = work[n - 1];
    (void)x;
    if (runs(st, RUN_ALWAYS)) st->pend = 1;
  });
}
void launch_step_end(DevState* st, int* ncols_dev, int m, int n, int, double, hipStream_t s) {
  submit(s, [=] { end_step(st, ncols_dev, m, n); });
}
void launch_resume(DevState* st, hipStream_t s) {
  submit(s, [=] { st->halt = 0; });
}
void launch_set_kk(DevState* st, int kk, hipStream_t s) {
  submit(s, [=] { st->fjb_cmp = kk; });
}
void launch_set_fused(DevState* st, int f, hipStream_t s) {
  submit(s, [=] { st->fused = f; });
}
void launch_zero(void* p, size_t bytes, hipStream_t s) {
  submit(s, [=] { memset(p, 0, bytes); });
}
void launch_swap2(double*, int, int, int, int, hipStream_t) {}
void launch_policy(const double*, int, int, const void*, DevState*, hipStream_t) {}

// wide steps: the kernels' outputs the host reads
void launch_topk_wide(const double* work, int nsub, int kmax, double* val[2], int* idx[2],
                      const double** out_val, hipStream_t s) {
  submit(s, [=] {
    volatile double x = work[nsub - 1];
    (void)x;
    for (int i = 0; i < kmax; i++) {
      val[0][i] = 1.0 - 1e-3 * i;
      idx[0][i] = i;
    }
  });
  *out_val = val[0];
}
const int* wide_topk_idx(double* val[2], int* idx[2], const double* out_val) {
  return out_val == val[0] ? idx[0] : idx[1];
}
void launch_gram_wide(const double*, int, int, int, const int*, int nc, double* G, int ldg,
                      hipStream_t s) {
  submit(s, [=] {
    for (int q = 0; q < nc; q++)
      for (int p = 0; p < nc; p++) G[p + (size_t)q * ldg] = (p == q) ? 1.0 : 0.0;
  });
}
void launch_select_wide(const double*, int, const double*, int kmax, double, double, int* out,
                        hipStream_t s) {
  submit(s, [=] {
    const int fjb = kmax - kmax / 7;
    out[0] = kmax;
    out[1] = fjb;
    for (int q = 0; q < fjb; q++) out[2 + q] = q + (q >= 3 && kmax > fjb ? 1 : 0) * (q % 2);
  });
}
void launch_swap_pairs(double*, int, int, int, double*, int*, const int* pairs, int np,
                       hipStream_t s) {
  submit(s, [=] {
    volatile int x = pairs[2 * np - 1];
    (void)x;
  });
}
void launch_wide_state(DevState* st, int j, int kk, hipStream_t s) {
  submit(s, [=] {
    st->j = j;
    st->fjb_cmp = kk;
    st->maxnrm_bits = 0ull;
    st->done = 0;
    st->halt = 0;
  });
}

// the batch kernel: every factorisation's outputs over their extents
int launch_dm_batch(int waves, int, const void* items, const int* order, int count, int* queue,
                    double*, double* tau, int* jpvt, int* ncols, int* info, int nb, double,
                    double, int, double, double, const void*, float*, unsigned*, int,
                    const float*, int, hipStream_t s) {
  if (waves != 4) return -1;
  submit(s, [=] {
    const qrdm_batch_item* it = static_cast<const qrdm_batch_item*>(items);
    for (int b = 0; b < count; b++) {
      const int i = order[b];
      const qrdm_batch_item& t = it[i];
      const int mn = minmn(t.m, t.n), k = nb < KMAX ? nb : KMAX;
      int j = 0, steps = 0;
      while (j < mn) {
        const int kk = std::max(1, std::min(k, mn - j) - (j % 3 == 0 ? 1 : 0));
        ncols[t.col_off + steps++] = kk;
        for (int c = 0; c < kk; c++) tau[t.tau_off + j + c] = 0.5;
        j += kk;
      }
      for (int c = 0; c < t.n; c++) jpvt[t.col_off + c] = c + 1;
      info[i] = 0;
    }
    queue[0] = queue[1] = 0;
  });
  return 0;
}

// training kernels of the learned proposal: not on this path
void launch_policy_grad(const float*, const float*, const float*, int, float*, float*,
                        hipStream_t) {}
void launch_policy_update(float*, float*, float*, const float*, const float*, float*, float*,
                          int, float, float, float, float, int, hipStream_t) {}
void launch_policy_pack(const float*, void*, float, float, float, float, hipStream_t) {}
void launch_policy_predict(const void*, const float*, int, float*, hipStream_t) {}
}  // namespace qrdm
