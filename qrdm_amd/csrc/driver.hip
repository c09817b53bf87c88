// driver.hip -- host driver of the DM-QR library and its C ABI.
//
// dgeqrdm_work's control flow (src/dgeqrdm_work.c:420-838) on the host, the
// arithmetic in the kernels of kernels.hip.  Per DM step: top-k of the partial
// norms -> candidate Gram -> greedy acceptance -> (host) permute_marked index
// walk -> column permutation -> Householder panel -> T -> trailing update ->
// norm downdate.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/qrdm_amd.h"
#include "qrdm_internal.h"

using namespace qrdm;

namespace {

thread_local std::string g_err;

constexpr double kEps = 1.1102230246251565e-16;  // LAPACK dlamch('E') = 2^-53
// Panels with at least this many rows take the Gram-based path.
constexpr int kFastPanelRows = 128;
constexpr int kBatch = 8;  // DM steps queued between synchronisations
// Layout of the per-step upload: up to 2*KMAX (dst, src) pairs, then KMAX
// candidate ranks of the panel columns.
constexpr int kRankOff = 4 * KMAX;
constexpr int kPermInts = kRankOff + KMAX;

struct HipError {
  hipError_t e;
  const char* what;
};

#define HIPCHK(x)                                      \
  do {                                                 \
    hipError_t e_ = (x);                               \
    if (e_ != hipSuccess) throw HipError{e_, #x};      \
  } while (0)

// Grow-only device workspace, one per device, reused across calls.
struct Workspace {
  size_t bytes = 0;
  char* base = nullptr;
  void* pinned = nullptr;  // host staging for the per-batch state readback (two slots)
  size_t pinned_bytes = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};  // completion of each slot's readback
};

std::mutex g_mu;
Workspace g_ws[64];

// Kernel-time accounting (qrdm_profile_*): HIP events around each class of
// kernels on the launch stream, resolved after the call's final sync.
constexpr int kNumClasses = 8;
enum { PC_LARFB = 0, PC_NORM = 1, PC_SELECT = 2, PC_PANEL = 3, PC_PERM = 4, PC_INIT = 5,
       PC_APPLY = 6, PC_WSUM = 7 };
struct Profiler {
  bool on = false;
  unsigned mask = 0xffffffffu;  // bit c: class c is timed
  double ms[kNumClasses] = {}, launches[kNumClasses] = {}, work[kNumClasses] = {};
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  struct Rec {
    int cls;
    hipEvent_t a, b;
    double work;
  };
  std::vector<Rec> open;
  hipEvent_t ev() {
    if (used == pool.size()) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      pool.push_back(e);
    }
    return pool[used++];
  }
  // begin/end bracket one region; work is its algorithmic flops or bytes
  int begin(int cls, hipStream_t s) {
    if (!on || !((mask >> cls) & 1u)) return -1;
    Rec r{cls, ev(), ev(), 0.0};
    HIPCHK(hipEventRecord(r.a, s));
    open.push_back(r);
    return (int)open.size() - 1;
  }
  void end(int id, hipStream_t s, double w) {
    if (id < 0) return;
    open[id].work = w;
    HIPCHK(hipEventRecord(open[id].b, s));
  }
  void resolve() {  // call after the stream is synchronised
    for (auto& r : open) {
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, r.a, r.b));
      ms[r.cls] += t;
      launches[r.cls] += 1;
      work[r.cls] += r.work;
    }
    open.clear();
    used = 0;
  }
};
Profiler g_prof;

template <typename T>
T* carve(char*& p, size_t count) {
  T* r = reinterpret_cast<T*>(p);
  p += ((count * sizeof(T) + 255) / 256) * 256;
  return r;
}

struct Buffers {
  DevState* st;
  double* work;      // n: partial norms (permuted with the columns)
  double* workold;   // n: norms at last recompute (positional)
  double* topk_val;   // topk_blocks(n) * 64: the top-k stage's per-workgroup lists
  int* topk_idx;
  double* gram_part;  // nslab * 64 * 64
  double* G;          // 64 * 64
  double* S;          // 64 * 64 (V^T V)
  double* T;          // 64 * 64
  double* R1;         // 64 * 64 (fast panel)
  double* R1inv;      // 64 * 64
  double* G2;         // 64 * 64
  double* M;          // 64 * 64
  double* V;          // 2 * mpad * 64: the dense V of two consecutive steps
  int ldv;
  double* Wp;  // nsplit * 64 * n: partial products of the trailing update
  double* W;   // 64 * n: W of the three-kernel trailing update (relative to c0)
  double* Wf;  // 2 * 64 * n: W of two consecutive steps, by absolute column (fused form)
  int* perm;   // kPermInts: (dst, src) pairs, then the panel's candidate ranks
  int* jp;     // n: jpvt on the device during the DM loop
  int* ncols;  // n: fjb_cmp of each DM step
  unsigned* tickets;  // launch tickets (zero between uses): [0] top-k, [1] selection,
                      // [2] reconstruction, [3] T of the Gram-based panel
  int nsplit;
  // wide steps (nb > KMAX) only: sort keys, the candidate Gram, the
  // selection's output, the swap pairs
  double* wval[2] = {nullptr, nullptr};
  int* widx[2] = {nullptr, nullptr};
  double* wG = nullptr;  // KW x KW
  int* wsel = nullptr;   // 2 + KW
  int* wpairs = nullptr;  // 4 KW
};

size_t layout_buffers(int m, int n, char* base, Buffers* b, int nb = KMAX) {
  char* p = base;
  const int mpad = ((m + 63) / 64) * 64;
  const int nslab = (m + 31) / 32 + 1;  // slabs of the Gram and fast-panel passes
  const int nsplit = 16;  // row slabs of the trailing update's partial products
  static_assert(nsplit >= 16, "the fused pass uses up to 16 row slabs");
  Buffers t;
  t.st = carve<DevState>(p, 1);
  t.work = carve<double>(p, n);
  t.workold = carve<double>(p, n);
  t.topk_val = carve<double>(p, (size_t)topk_blocks(n) * KMAX);
  t.topk_idx = carve<int>(p, (size_t)topk_blocks(n) * KMAX);
  t.gram_part = carve<double>(p, (size_t)nslab * KMAX * KMAX);
  t.G = carve<double>(p, KMAX * KMAX);
  t.S = carve<double>(p, KMAX * KMAX);
  t.T = carve<double>(p, KMAX * KMAX);
  t.R1 = carve<double>(p, KMAX * KMAX);
  t.R1inv = carve<double>(p, KMAX * KMAX);
  t.G2 = carve<double>(p, KMAX * KMAX);
  t.M = carve<double>(p, KMAX * KMAX);
  t.ldv = mpad;
  t.V = carve<double>(p, (size_t)2 * mpad * KMAX);
  t.nsplit = nsplit;
  t.Wp = carve<double>(p, (size_t)nsplit * KMAX * n);
  t.W = carve<double>(p, (size_t)KMAX * n);
  t.Wf = carve<double>(p, (size_t)2 * KMAX * n);
  t.perm = carve<int>(p, kPermInts);
  t.jp = carve<int>(p, n);
  t.ncols = carve<int>(p, n);
  t.tickets = carve<unsigned>(p, (size_t)(n + 63) / 64 + 4);
  if (nb > KMAX) {
    for (int q = 0; q < 2; q++) {
      t.wval[q] = carve<double>(p, (size_t)wide_sort_keys(n));
      t.widx[q] = carve<int>(p, (size_t)wide_sort_keys(n));
    }
    t.wG = carve<double>(p, (size_t)KW * KW);
    t.wsel = carve<int>(p, 2 + KW);
    t.wpairs = carve<int>(p, 4 * KW);
  }
  if (b) *b = t;
  return (size_t)(p - base);
}

// Index walk of permute_marked (src/dgeqrdm_work.c:149-262) with nz = 0 and
// marks 0/1: the same two-pointer scan and placement order, recorded as a
// sequence of position swaps.  marked has ncols entries, zero on entry and on
// exit.
void permute_walk(int ncols, std::vector<int>& marked, const int* sel, int nsel,
                  std::vector<std::pair<int, int>>& swaps) {
  swaps.clear();
  for (int s = 0; s < nsel; s++) marked[sel[s]] = 1;
  int jb = 0, jt = ncols - 1;
  auto swp = [&](int p, int q) {
    std::swap(marked[p], marked[q]);
    swaps.emplace_back(p, q);
  };
  for (int s = 0; s < nsel; s++) {
    const int jc = sel[s];
    while (jb < jt && marked[jt] == 1) {
      swp(jt, jb);
      while (jb < ncols && marked[jb] != 0) jb++;
    }
    if (marked[jc] == 1) {
      while (jb < ncols && marked[jb] != 0) jb++;
      if (jc <= jb || jc < nsel) continue;
      if (marked[jb] == 0) {
        swp(jc, jb);
        jb++;
      }
    }
  }
  for (int p = 0; p < nsel; p++) marked[p] = 0;  // the accepted ones now lead
}

struct Args {
  int layout, m, n, lda, nb;
  double* a;
  int* jpvt;
  double* tau;
  int* ncols;
  const double* thres;
  hipStream_t s;
  qrdm_step* trace;
  int ntrace;
  int* nsteps;
  int trail = TRAIL_FP64;  // precision of the trailing update's products
  const void* policy = nullptr;  // learned (tau, delta) network on the device, or null
  // A caller-given device workspace (at least layout_buffers(m, n) bytes,
  // 256-byte aligned): the call then allocates nothing -- the state readback
  // goes to a static host mirror and waits on the stream, regions that must
  // start at zero are cleared by a kernel, and profiling is suspended.
  char* scratch = nullptr;
};

// Host mirror of the device state for calls on a caller-given workspace (two
// readback slots per device).
DevState g_host_state[64][2];

int check_args(const Args& x) {
  const double delta = x.thres[0], tau = x.thres[1];
  int info = 0;
  if (x.layout != 102 && x.layout != 101) info = -1;
  else if (x.m <= 0) info = -2;
  else if (x.n <= 0) info = -3;
  else if (x.lda < (x.m > 1 ? x.m : 1)) info = -5;
  else if (delta < 0.0 || delta > 1.0) info = -9;
  else if (tau < 0.0 || tau > 1.0) info = -9;
  else if (x.nb <= 0) info = -10;
  if (info == 0 && x.nb > KW) {
    fprintf(stderr, "DGEQRDM: nb = %d exceeds the largest DM block of this device build (%d)\n",
            x.nb, KW);
    return -10;
  }
  if (info != 0) {
    // the reference reports through xerbla_ and returns -1 (:577-581)
    fprintf(stderr, " ** On entry to DGEQRDM parameter number %2d had an illegal value\n", -info);
    return -1;
  }
  return 0;
}

Workspace& workspace_for(int m, int n, Buffers* b, int nb = KMAX) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  Workspace& ws = g_ws[dev & 63];
  size_t need = layout_buffers(m, n, nullptr, nullptr, nb) + 256;
  if (ws.bytes < need) {
    if (ws.base) HIPCHK(hipFree(ws.base));
    ws.base = nullptr;
    ws.bytes = 0;
    HIPCHK(hipMalloc(&ws.base, need));
    ws.bytes = need;
  }
  if (!ws.pinned) {
    ws.pinned_bytes = 2 * sizeof(DevState) + 4096;
    HIPCHK(hipHostMalloc(&ws.pinned, ws.pinned_bytes, hipHostMallocDefault));
    for (int q = 0; q < 2; q++) HIPCHK(hipEventCreateWithFlags(&ws.ev[q], hipEventDisableTiming));
  }
  layout_buffers(m, n, ws.base, b, nb);
  // The buffer offsets depend on (m, n): clear the trailing update's ticket
  // counters, which must start every call at zero.
  HIPCHK(hipMemset(b->tickets, 0, sizeof(unsigned) * ((size_t)(n + 63) / 64 + 4)));
  return ws;
}

void blocked_qr(double* A, int lda, int m, int n, int ncols_qr, double* tau, Buffers& B,
                hipStream_t s);

// DM steps wider than KMAX (64 < nb <= KW), host-driven: one synchronisation
// per step (src/dgeqrdm_work.c:694-787 with k_max up to KW; wide.hip).  The
// panel is the column-by-column one on all fjb columns; the kept reflectors
// reach the columns past the block in KMAX-column chunks (H_0 .. H_{k-1} in
// order, as the reference's single dlarfb applies them); the norm downdate
// sums the step's fjb_cmp new rows.  Returns 0 (the results are copied back
// and the stream synchronised).
int run_wide(const Args& x, Buffers& B, int nfxd, double eta, int enable_stop, double tol3z) {
  const int m = x.m, n = x.n, lda = x.lda, nb = x.nb;
  const double delta = x.thres[0], tau_dm = x.thres[1];
  const int minmn = m < n ? m : n;
  hipStream_t s = x.s;
  double* A = x.a;
  HIPCHK(hipMemcpyAsync(B.jp, x.jpvt, sizeof(int) * n, hipMemcpyHostToDevice, s));
  std::vector<int> marked(n, 0), hout(2 + KW), hidx(KW), sel, pairs;
  std::vector<std::pair<int, int>> swaps;
  int j = nfxd, it = 0;
  while (j < minmn) {
    const int msub = m - j, nsub = n - j;
    const int kmax = std::min(std::min(nb, msub), nsub);
    // --- DM_perm (src/dgeqrdm_work.c:280-418): candidates, Gram, greedy
    const double* val = nullptr;
    launch_topk_wide(B.work + j, nsub, kmax, B.wval, B.widx, &val, s);
    const int* idx = wide_topk_idx(B.wval, B.widx, val);
    launch_gram_wide(A, lda, j, m, idx, kmax, B.wG, KW, s);
    launch_select_wide(B.wG, KW, val, kmax, tau_dm, delta, B.wsel, s);
    HIPCHK(hipMemcpyAsync(hout.data(), B.wsel, sizeof(int) * (2 + KW), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(hidx.data(), idx, sizeof(int) * kmax, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const int nc = hout[0], fjb = hout[1];
    sel.resize(fjb);
    for (int q = 0; q < fjb; q++) sel[q] = hidx[hout[2 + q]];
    // permute_marked (:149-262): the walk here, the moves on the device
    permute_walk(nsub, marked, sel.data(), fjb, swaps);
    pairs.resize(2 * swaps.size());
    for (size_t q = 0; q < swaps.size(); q++) {
      pairs[2 * q] = swaps[q].first;
      pairs[2 * q + 1] = swaps[q].second;
    }
    if (!swaps.empty()) {
      HIPCHK(hipMemcpyAsync(B.wpairs, pairs.data(), sizeof(int) * pairs.size(),
                            hipMemcpyHostToDevice, s));
      launch_swap_pairs(A, lda, m, j, B.work, B.jp, B.wpairs, (int)swaps.size(), s);
    }
    // --- dgeqr2_mia (:735-738) on the fjb columns, truncation included
    launch_panel_exact(A, lda, j, m, fjb, tau_dm, 1, x.tau + j, B.st, s);
    int kk = 0;
    HIPCHK(hipMemcpyAsync(&kk, &B.st->fjb_cmp, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // --- dlarft + dlarfb (:740-767) on columns j + fjb .. n, KMAX reflectors
    // at a time
    if (j + fjb < n) {
      for (int i0 = 0; i0 < kk; i0 += KMAX) {
        const int kq = std::min(KMAX, kk - i0);
        launch_set_kk(B.st, kq, s);
        launch_build_v(A, lda, j + i0, m, B.ldv, B.st, B.V, B.ldv, s);
        launch_gram(B.V, B.ldv, j + i0, m, nullptr, KMAX, nullptr, B.gram_part, B.S, s);
        launch_larft(B.S, x.tau + j + i0, B.st, B.T, s);
        launch_larfb(A, lda, j + i0, m, j + fjb, n, B.V, B.ldv, B.T, 1, B.Wp, B.W, B.nsplit, s);
      }
    }
    // --- norm_update (:774-778) over the kk new rows
    launch_wide_state(B.st, j, kk, s);
    launch_norm_update(A, lda, j, m, n, B.work, B.workold, B.st, tol3z, s);
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, &B.st->maxnrm_bits, sizeof(bits), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double maxnrm = 0.0;
    memcpy(&maxnrm, &bits, sizeof(double));
    x.ncols[it] = kk;
    if (x.trace && it < x.ntrace) {
      qrdm_step& r = x.trace[it];
      memset(&r, 0, sizeof(r));
      r.j = j;
      r.kmax = kmax;
      r.nc = nc;
      r.fjb = fjb;
      r.fjb_cmp = kk;
      r.nswap = (int)swaps.size();
      r.panel = 0;
      r.maxnrm = maxnrm;
      r.tau = tau_dm;
      r.delta = delta;
      r.prec = TRAIL_FP64;
      for (int q = 0; q < fjb && q < 64; q++) r.sel[q] = sel[q];
    }
    it++;
    j += kk;
    if (enable_stop && maxnrm * sqrt((double)(nsub - kk)) <= eta) break;  // :782-785
  }
  if (x.nsteps) *x.nsteps = it;
  HIPCHK(hipMemcpyAsync(x.jpvt, B.jp, sizeof(int) * n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int run_dm(const Args& x) {
  const int m = x.m, n = x.n, lda = x.lda, nb = x.nb;
  const double delta = x.thres[0], tau_dm = x.thres[1];
  const double tol3z = sqrt(kEps);
  hipStream_t s = x.s;
  int enable_stop = 0;
  double eta = 0.0;
  if (x.ncols[0] == 1) {
    enable_stop = 1;
    eta = kEps * n;
  } else if (x.ncols[0] == 2) {
    enable_stop = 2;
    eta = kEps * sqrt((double)n);
  } else if (x.ncols[0] == 3) {
    enable_stop = 3;
    eta = x.thres[2];
  }
  if (x.nsteps) *x.nsteps = 0;
  const int minmn = m < n ? m : n;

  Buffers B;
  DevState* hst = nullptr;
  hipEvent_t* evs = nullptr;  // completion of each readback slot (none: wait on the stream)
  const bool own_ws = x.scratch == nullptr;
  if (own_ws) {
    Workspace& ws = workspace_for(m, n, &B, x.nb);
    hst = reinterpret_cast<DevState*>(ws.pinned);
    evs = ws.ev;
  } else {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    layout_buffers(m, n, x.scratch, &B, x.nb);
    hst = g_host_state[dev & 63];
    // the launch tickets must start at zero (the workspace keeps the last
    // call's contents)
    launch_zero(B.tickets, sizeof(unsigned) * ((size_t)(n + 63) / 64 + 4), s);
  }
  auto clear = [&](void* p, size_t bytes) {
    if (own_ws) HIPCHK(hipMemsetAsync(p, 0, bytes, s));
    else launch_zero(p, bytes, s);
  };
  // profiling records events, which a call on a caller-given workspace does
  // not create: suspended for it
  struct ProfPause {
    bool was, pause;
    ProfPause(bool p) : was(g_prof.on), pause(p) { if (pause) g_prof.on = false; }
    ~ProfPause() { if (pause) g_prof.on = was; }
  } prof_pause(!own_ws);
  double* A = x.a;

  // Fixed columns (src/dgeqrdm_work.c:594-610): jpvt != 0 columns go first.
  int nfxd = 0;
  for (int jj = 0; jj < n; jj++) {
    if (x.jpvt[jj] != 0) {
      if (jj != nfxd) {
        launch_swap2(A, lda, m, jj, nfxd, s);
        x.jpvt[jj] = x.jpvt[nfxd];
        x.jpvt[nfxd] = jj + 1;
      } else {
        x.jpvt[jj] = jj + 1;
      }
      nfxd++;
    } else {
      x.jpvt[jj] = jj + 1;
    }
  }
  clear(B.st, sizeof(DevState));
  if (nfxd > 0) {
    // Unpivoted blocked QR of the fixed columns, applied to the rest
    // (LAPACKE_dgeqrf + LAPACKE_dormqr, src/dgeqrdm_work.c:618-635).
    blocked_qr(A, lda, m, n, nfxd < m ? nfxd : m, x.tau, B, s);
  }
  int info = 0;
  if (nfxd < minmn) {
    int pi = g_prof.begin(PC_INIT, s);
    launch_init_norms(A, lda, nfxd, m, n, B.work, B.workold, B.st, s);
    g_prof.end(pi, s, 8.0 * (double)(m - nfxd) * n);
    // columns < nfxd are not DM columns; their norms are never read
    HIPCHK(hipMemcpyAsync(hst, B.st, sizeof(DevState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // A NaN or infinity makes the reference's LAPACKE_dlarfb_mia NaN scan
    // return -13 after parts of A were already overwritten
    // (src/dlarfb.c:54-127); here it is caught by the norm sweep before any
    // write, and A is left as it was.
    if (hst->flags & 1) return -13;
    double maxnrm = 0.0;
    {
      unsigned long long bits = hst->maxnrm_bits;
      memcpy(&maxnrm, &bits, sizeof(double));
    }
    eta = eta * maxnrm;
    if (nb > KMAX) return run_wide(x, B, nfxd, eta, enable_stop, tol3z);
    // --- the DM loop (src/dgeqrdm_work.c:694-787), driven from the device:
    // every kernel of a step reads j and the step's decisions from DevState,
    // so the host queues kBatch steps at a time and synchronises once per
    // batch; kernels of steps past the end return at once.  Host-side sizes
    // use j0, the column reached at the last synchronisation (j >= j0).
    const char* pe = getenv("QRDM_PANEL");
    const bool force_exact = pe && strcmp(pe, "exact") == 0;
    const int fast_rows = force_exact ? INT_MAX : kFastPanelRows;
    // The FP64 trailing update runs deferred and merged into the next step's
    // pass over C (trailing.hip); the reduced-precision products, and
    // QRDM_TRAILING=3pass, apply each step's update in three kernels.
    // The merged pass moves C in 32-byte vectors: lda % 4 == 0, A 32-byte
    // aligned, at least four rows.
    const char* pt = getenv("QRDM_TRAILING");
    bool fused_now = !(pt && strcmp(pt, "3pass") == 0) && lda % 4 == 0 && m >= 4 &&
                     (reinterpret_cast<uintptr_t>(A) & 31) == 0;
    StepBufs bufs{B.V, fused_now ? (size_t)B.ldv * KMAX : 0, B.ldv, B.Wf,
                  fused_now ? (size_t)KMAX * n : 0};
    // Reduced-precision forms (x.trail): the trailing update stays FP64 until
    // the remaining block is small enough for the low precision's rounding to
    // stay below 1e-12 of A -- maxnrm sqrt(n - j) <= 1e-12 maxnrm0 /
    // (64 u_low), maxnrm0 the largest initial column norm (a lower bound of
    // ||A||), 64 the products' accumulation length, u_low 2^-24 (FP32) or
    // 2^-17 (split BF16) -- and from then on runs in x.trail in the
    // three-kernel form.  The host decides between batches (at a batch
    // boundary with nothing queued behind it).
    const int low = (x.trail == QRDM_TRAILING_FP32 || x.trail == QRDM_TRAILING_FP32_ALL)
                        ? TRAIL_FP32 : TRAIL_BF16X3;
    const bool scheduled = x.trail == QRDM_TRAILING_FP32 || x.trail == QRDM_TRAILING_BF16X3;
    const bool always_low = x.trail >= QRDM_TRAILING_FP32_ALL;
    int prec_now = always_low ? low : TRAIL_FP64;
    if (always_low) {
      fused_now = false;
      bufs = StepBufs{B.V, 0, B.ldv, B.Wf, 0};
    }
    bool switch_wanted = false;
    const double low_bound = 1e-12 * maxnrm / (64.0 * (low == TRAIL_FP32 ? 0x1p-24 : 0x1p-17));
    if (fused_now) {
      // the previous step's V and W of the first step: zero, so that its
      // deferred update changes nothing
      clear(B.V, sizeof(double) * 2 * (size_t)B.ldv * KMAX);
      clear(B.Wf, sizeof(double) * 2 * (size_t)KMAX * n);
    }
    HIPCHK(hipMemcpyAsync(B.jp, x.jpvt, sizeof(int) * n, hipMemcpyHostToDevice, s));
    launch_dm_begin(B.st, nfxd, minmn, tau_dm, delta, fused_now ? 1 : 0, s);
    const int batch = x.trace ? 1 : kBatch;
    int j0 = nfxd, it = 0;
    // the exact panel and dlarft of the step at column jl (RUN_EXACT: they
    // run when the step's panel is not the Gram-based one, or that fell back)
    auto queue_exact = [&](int jl) {
      int pq = g_prof.begin(PC_PANEL, s);
      launch_panel_exact(A, lda, jl, m, 0, tau_dm, 1, x.tau, B.st, s, RUN_EXACT);
      launch_vgram_exact(A, lda, jl, m, B.st, bufs, B.gram_part, s);
      launch_larft(B.gram_part, x.tau, B.st, B.T, s, RUN_EXACT, vgram_exact_slabs(jl, m));
      g_prof.end(pq, s, 0.0);
    };
    // --- dlarfb (src/dgeqrdm_work.c:762-767) from c0 (set by the panel),
    // norm_update (src/dgeqrdm_work.c:774-778); the stop test and the next j
    // follow in the step's end
    auto queue_update = [&](int jl) {
      if (fused_now) {
        // C -= V_prev W_prev and the partials V^T C in one pass, then W, the
        // R12 rows and the norm downdate (the update itself stays deferred)
        const int ns = fused_slabs(jl, m, jl + 1, n);
        int pl = g_prof.begin(PC_LARFB, s);
        launch_fused(A, lda, jl, m, jl + 1, n, bufs, B.Wp, ns, 1, 1, 0, s, B.st);
        g_prof.end(pl, s, 0.0);
        int pw = g_prof.begin(PC_WSUM, s);
        launch_wsum_norm(A, lda, jl, m, n, B.T, B.Wp, ns, bufs, B.work, B.workold, B.st, tol3z,
                         s);
        g_prof.end(pw, s, 0.0);
        return;
      }
      int pl = g_prof.begin(PC_LARFB, s);
      launch_gemm1_prec(prec_now, A, lda, jl, m, jl + 1, n, B.V, B.ldv, B.Wp, B.nsplit, s, B.st);
      g_prof.end(pl, s, 0.0);
      int pw = g_prof.begin(PC_WSUM, s);
      launch_wsum(jl, m, jl + 1, n, B.T, 1, B.Wp, B.W, B.nsplit, s, B.st);
      g_prof.end(pw, s, 0.0);
      int pa = g_prof.begin(PC_APPLY, s);
      launch_apply_prec(prec_now, A, lda, jl, m, jl + 1, n, B.V, B.ldv, B.W, s, B.st);
      g_prof.end(pa, s, 0.0);
      int pn = g_prof.begin(PC_NORM, s);
      launch_norm_update(A, lda, jl, m, n, B.work, B.workold, B.st, tol3z, s);
      g_prof.end(pn, s, 0.0);
    };
    // one DM step from column jl (a lower bound of the step's j)
    auto queue_step = [&](int jl, bool tall) {
      // --- DM_perm (src/dgeqrdm_work.c:280-418) ---------------------------
      int ps = g_prof.begin(PC_SELECT, s);
      launch_topk(B.work, jl, m, n, nb, tau_dm, B.topk_val, B.topk_idx, B.tickets, B.st, B.ncols,
                  enable_stop, eta, s);
      // the learned proposal replaces the step's candidate count and delta
      if (x.policy) launch_policy(B.work, m, n, x.policy, B.st, s);
      // candidate Gram over rows j..m (cosines need it only when nc > 1;
      // the kernel reads nc from the state and zero-fills beyond it), then
      // the greedy acceptance and the permutation plan
      launch_gram_select(A, lda, jl, m, n, B.gram_part, B.G, delta, fast_rows, tall ? 0 : 1,
                         B.perm, B.tickets + 1, B.st, bufs, s);
      g_prof.end(ps, s, 0.0);
      // --- dgeqr2_mia (src/dgeqrdm_work.c:735-738) ------------------------
      int pp = g_prof.begin(PC_PERM, s);
      if (tall)
        launch_perm_chol1(A, lda, m, B.work, B.jp, B.perm, B.G, tau_dm, B.R1, B.R1inv, B.st,
                          bufs, s);
      else
        launch_apply_perm(A, lda, m, B.work, B.jp, B.perm, B.st, bufs, s);
      g_prof.end(pp, s, 0.0);
      if (tall) {
        int pq = g_prof.begin(PC_PANEL, s);
        launch_panel_fast(A, lda, jl, m, tau_dm, x.tau, B.gram_part, B.R1, B.R1inv, B.G2, B.T,
                          B.M, bufs, B.tickets + 2, B.st, s);
        g_prof.end(pq, s, 0.0);
      } else {
        queue_exact(jl);
      }
      queue_update(jl);
    };
    // a batch of steps from column jl, the pending step end, and the state
    // read back into slot q
    DevState* hb[2] = {hst, hst + 1};
    auto readback = [&](int q) {
      launch_step_end(B.st, B.ncols, m, n, enable_stop, eta, s);
      HIPCHK(hipMemcpyAsync(hb[q], B.st, sizeof(DevState), hipMemcpyDeviceToHost, s));
      if (evs) HIPCHK(hipEventRecord(evs[q], s));
    };
    auto queue_batch = [&](int jl, int q) {
      // Tall panels: Gram-based Householder reconstruction (panel.hip), which
      // also leaves V and T for the trailing update, its first Cholesky beside
      // the column moves; the exact panel is not queued, and a step that
      // needs it (short panel, fjb < 2, or the Gram-based form falling back)
      // halts until the host resumes it.  Short panels: the column-by-column
      // exact panel, then dlarft.
      const bool tall = m - jl >= kFastPanelRows && !force_exact;
      for (int b = 0; b < batch; b++) queue_step(jl, tall);
      readback(q);
    };
    auto wait_slot = [&](int q) {
      if (evs) HIPCHK(hipEventSynchronize(evs[q]));
      else HIPCHK(hipStreamSynchronize(s));
      return hb[q];
    };
    // While the batch being waited for cannot end the loop (its first column
    // plus a batch of KMAX-wide steps stays below min(m, n)), the next batch
    // is queued before its state is read back, so the device does not idle
    // through the readback; its grids are sized from the same lower bound.
    int q = 0, jend = j0;
    queue_batch(j0, q);
    while (true) {
      const bool ahead = !x.trace && !switch_wanted && j0 + batch * KMAX < minmn;
      if (ahead) queue_batch(j0, q ^ 1);
      DevState* h = wait_slot(q);
      bool in_flight = ahead;
      if (h->halt) {
        // the halted step (at h->j, selection and column moves done)
        // continues with the exact panel, after the batch queued behind it
        // (which the halt stops at once)
        if (in_flight) wait_slot(q ^ 1);
        in_flight = false;
        const int jh = h->j;
        launch_resume(B.st, s);
        queue_exact(jh);
        queue_update(jh);
        readback(q);
        h = wait_slot(q);
        if (h->halt) throw HipError{hipErrorUnknown, "resumed DM step halted again"};
      }
      if (x.trace && h->it > it && h->it - 1 < x.ntrace) {
        const StepRec& l = h->last;
        qrdm_step& r = x.trace[h->it - 1];
        memset(&r, 0, sizeof(r));
        r.j = l.j;
        r.kmax = l.kmax;
        r.nc = l.nc;
        r.fjb = l.fjb;
        r.fjb_cmp = l.fjb_cmp;
        r.nswap = l.nswap;
        r.panel = l.panel;
        r.maxnrm = l.maxnrm;
        r.tau = l.tau;
        r.delta = l.delta;
        r.prec = prec_now;
        for (int t = 0; t < l.fjb && t < 64; t++) r.sel[t] = h->sel[t];
      }
      it = h->it;
      if (h->done) {  // a batch queued ahead returns at once; the final sync drains it
        jend = h->j;
        break;
      }
      j0 = h->j;
      if (scheduled && prec_now == TRAIL_FP64 &&
          h->last.maxnrm * sqrt((double)(n - j0)) <= low_bound)
        switch_wanted = true;
      if (in_flight) {
        q ^= 1;  // the batch queued ahead is the one to wait for next
      } else {
        if (switch_wanted) {
          // the pending merged update reaches rows and columns >= j0 (the
          // R12 rows of the last step are final), then the lower precision
          if (fused_now && it > 0) {
            launch_fused(A, lda, j0, m, j0, n, bufs, B.Wp, fused_slabs(j0, m, j0, n), 1, 0,
                         it & 1, s, nullptr);
            launch_set_fused(B.st, 0, s);
          }
          fused_now = false;
          bufs = StepBufs{B.V, 0, B.ldv, B.Wf, 0};
          prec_now = low;
          switch_wanted = false;
        }
        queue_batch(j0, q);
      }
    }
    if (fused_now && it > 0) {
      // the last step's update is still deferred: it reaches the trailing
      // block when the stop test ended the loop before min(m, n)
      const int je = jend;
      launch_fused(A, lda, je, m, je, n, bufs, B.Wp, fused_slabs(je, m, je, n), 1, 0, it & 1, s,
                   nullptr);
    }
    if (x.nsteps) *x.nsteps = it;
    HIPCHK(hipMemcpyAsync(x.jpvt, B.jp, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    if (it > 0) HIPCHK(hipMemcpyAsync(x.ncols, B.ncols, sizeof(int) * it, hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  if (g_prof.on) {
    g_prof.resolve();
    HIPCHK(hipMemcpy(hst, B.st, sizeof(DevState), hipMemcpyDeviceToHost));
    const StepWork& w = hst->work;
    g_prof.work[PC_SELECT] += w.select;
    g_prof.work[PC_PERM] += w.perm;
    g_prof.work[PC_PANEL] += w.panel;
    g_prof.work[PC_LARFB] += w.gemm1;
    g_prof.work[PC_APPLY] += w.gemm2;
    g_prof.work[PC_WSUM] += w.wsum;
    g_prof.work[PC_NORM] += w.norm + 8.0 * (double)hst->recomp_rows;
  }
  return info;
}

// Unpivoted blocked Householder QR (dgeqrf semantics): exact panels of up to
// KMAX columns, each followed by its block update of the columns to the right.
void blocked_qr(double* A, int lda, int m, int n, int ncols_qr, double* tau, Buffers& B,
                hipStream_t s) {
  for (int j = 0; j < ncols_qr; j += KMAX) {
    const int kb = (ncols_qr - j) < KMAX ? (ncols_qr - j) : KMAX;
    launch_panel_exact(A, lda, j, m, kb, 0.0, 0, tau + j, B.st, s);
    launch_build_v(A, lda, j, m, B.ldv, B.st, B.V, B.ldv, s);
    launch_gram(B.V, B.ldv, j, m, nullptr, KMAX, nullptr, B.gram_part, B.S, s);
    launch_larft(B.S, tau + j, B.st, B.T, s);
    launch_larfb(A, lda, j, m, j + kb, n, B.V, B.ldv, B.T, 1, B.Wp, B.W, B.nsplit, s);
  }
}

int run_geqrf(double* A, int lda, int m, int n, double* tau, hipStream_t s) {
  Buffers B;
  workspace_for(m, n, &B);
  HIPCHK(hipMemsetAsync(B.st, 0, sizeof(DevState), s));
  blocked_qr(A, lda, m, n, m < n ? m : n, tau, B, s);
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

// C := Q C or Q^T C, Q = H(0)...H(k-1) stored in A (m x k) and tau.
int run_ormqr(int trans, int m, int n, int k, const double* A, int lda, const double* tau,
              double* C, int ldc, hipStream_t s) {
  Buffers B;
  workspace_for(m, n, &B);
  HIPCHK(hipMemsetAsync(B.st, 0, sizeof(DevState), s));
  const int nblk = (k + KMAX - 1) / KMAX;
  for (int q = 0; q < nblk; q++) {
    const int b = trans ? q : nblk - 1 - q;  // Q^T: H(0) first; Q: H(k-1) first
    const int j = b * KMAX;
    const int kb = (k - j) < KMAX ? (k - j) : KMAX;
    launch_set_kk(B.st, kb, s);
    launch_build_v(A, lda, j, m, B.ldv, B.st, B.V, B.ldv, s);
    launch_gram(B.V, B.ldv, j, m, nullptr, KMAX, nullptr, B.gram_part, B.S, s);
    launch_larft(B.S, tau + j, B.st, B.T, s);
    launch_larfb(C, ldc, j, m, 0, n, B.V, B.ldv, B.T, trans, B.Wp, B.W, B.nsplit, s);
  }
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

template <typename F>
int guarded_call(F f) {
  std::lock_guard<std::mutex> lk(g_mu);
  try {
    g_err.clear();
    return f();
  } catch (const HipError& e) {
    g_err = std::string(e.what) + ": " + hipGetErrorString(e.e);
    return -1000 - (int)e.e;
  }
}

int guarded(const Args& x) {
  std::lock_guard<std::mutex> lk(g_mu);
  try {
    g_err.clear();
    return run_dm(x);
  } catch (const HipError& e) {
    g_err = std::string(e.what) + ": " + hipGetErrorString(e.e);
    return -1000 - (int)e.e;
  }
}

}  // namespace

namespace {
// Copies a column-major host block into a fresh device block (ld rounded up).
struct DevBlock {
  double* p = nullptr;
  int ld = 0;
  ~DevBlock() {
    if (p) (void)hipFree(p);
  }
};
void to_device(DevBlock& b, const double* h, int ldh, int rows, int cols) {
  b.ld = ((rows + 31) / 32) * 32;
  if (b.ld < 32) b.ld = 32;
  HIPCHK(hipMalloc(&b.p, sizeof(double) * (size_t)b.ld * (cols > 0 ? cols : 1)));
  if (rows > 0 && cols > 0)
    HIPCHK(hipMemcpy2D(b.p, sizeof(double) * b.ld, h, sizeof(double) * ldh,
                       sizeof(double) * rows, cols, hipMemcpyHostToDevice));
}
void to_host(double* h, int ldh, const DevBlock& b, int rows, int cols) {
  if (rows > 0 && cols > 0)
    HIPCHK(hipMemcpy2D(h, sizeof(double) * ldh, b.p, sizeof(double) * b.ld,
                       sizeof(double) * rows, cols, hipMemcpyDeviceToHost));
}
}  // namespace

// ---------------------------------------------------------------------------
// Batched DM-QR: a plan holds the matrices' descriptors on the device, the
// per-workgroup schedule and the launch shape; a run is one launch of
// k_dm_batch (optionally replayed from a captured hipGraph).
struct qrdm_batch_plan {
  int count = 0;
  int nb_req = 0;   // the caller's nb (<= 0: picked by qrdm_batch_autotune)
  int nb_run = 64;  // nb of the runs
  int waves = 4;    // 64-thread waves per workgroup (4 or 8)
  int gram_rows = 128;  // panels of >= this many rows take the Gram-based form (0: none)
  int grid = 0;
  std::vector<qrdm_batch_item> items;
  void* d_items = nullptr;
  int* d_sched = nullptr;  // count: the matrices in the order they are taken
  int* d_queue = nullptr;  // 2: the launch's work counter and finished-workgroup count
  size_t a_ext = 0, tau_ext = 0, col_ext = 0;  // elements the batch spans
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  struct Key {
    const void *a, *tau, *jp, *nc, *info;
    double th[3];
    int stop, waves, nb, gram_rows;
    bool operator==(const Key& o) const {
      return a == o.a && tau == o.tau && jp == o.jp && nc == o.nc && info == o.info &&
             th[0] == o.th[0] && th[1] == o.th[1] && th[2] == o.th[2] && stop == o.stop &&
             waves == o.waves && nb == o.nb && gram_rows == o.gram_rows;
    }
  } key{};
  double tuned_ms[16] = {};
  int tuned_n = 0;
  const void* policy = nullptr;  // learned (tau, delta) network on the device, or null
};

namespace {

// Flop count of one factorisation (2 m n^2 - 2 n^3 / 3 for m >= n).
double qr_flops(int m, int n) {
  const double a = m > n ? m : n, b = m > n ? n : m;
  return 2.0 * a * b * b - 2.0 * b * b * b / 3.0;
}

// The order the workgroups take the matrices in: decreasing flop count
// (ties: lowest index), so that the last matrices to start are the
// shortest (list scheduling, longest first).
std::vector<int> batch_order(const std::vector<qrdm_batch_item>& it) {
  const int cnt = (int)it.size();
  std::vector<int> order(cnt);
  for (int i = 0; i < cnt; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return qr_flops(it[x].m, it[x].n) > qr_flops(it[y].m, it[y].n);
  });
  return order;
}

void batch_release_graph(qrdm_batch_plan* p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
}

void batch_launch(qrdm_batch_plan* p, int waves, int nb, int gram_rows, double* a, double* tau,
                  int* jp, int* nc, int* info, const double* th, int stop, hipStream_t s,
                  float* trace = nullptr, unsigned* trace_count = nullptr, int trace_cap = 0,
                  const float* ranges = nullptr) {
  if (launch_dm_batch(waves, p->grid, p->d_items, p->d_sched, p->count, p->d_queue, a, tau, jp,
                      nc, info, nb, th[0], th[1], stop, th[2], sqrt(kEps), p->policy, trace,
                      trace_count, trace_cap, ranges, gram_rows, s) != 0)
    throw HipError{hipErrorInvalidValue, "batch launch shape"};
  HIPCHK(hipGetLastError());
}

}  // namespace

// ---------------------------------------------------------------------------
// Workspace of qrdm_dgeqrdm_dev_noalloc: a device array of fixed size (every
// device the library runs on holds its own copy until the process ends).  In
// host code the array's name is not a device address: the address is taken
// on each device once, on the first call there, with hipGetSymbolAddress, and
// only that address is used.  Calls that overlap on one device share it and
// race; a call starts by clearing what must start at zero.
static constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) char scratch[kScratchBytes];
static void* g_scratch_ptr[64];  // per device; null before the first call there

extern "C" {

const char* qrdm_version(void) { return "qrdm_amd 0.1 (gfx950)"; }

void qrdm_profile_enable(int on) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_prof.on = on != 0;
}

void qrdm_profile_select(unsigned mask) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_prof.mask = mask;
}

void qrdm_profile_reset(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (int c = 0; c < kNumClasses; c++) g_prof.ms[c] = g_prof.launches[c] = g_prof.work[c] = 0;
}

int qrdm_profile_read(double* out, int nclass) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (int c = 0; c < nclass && c < kNumClasses; c++) {
    out[3 * c] = g_prof.ms[c];
    out[3 * c + 1] = g_prof.launches[c];
    out[3 * c + 2] = g_prof.work[c];
  }
  return kNumClasses;
}

int qrdm_permute_walk(int ncols, const int* sel, int nsel, int* swaps) {
  if (ncols <= 0 || nsel <= 0 || nsel > ncols) return -1;
  std::vector<int> marked(ncols, 0);
  for (int s = 0; s < nsel; s++) {
    if (sel[s] < 0 || sel[s] >= ncols || marked[sel[s]]) return -1;
    marked[sel[s]] = 1;
  }
  for (int s = 0; s < nsel; s++) marked[sel[s]] = 0;
  std::vector<std::pair<int, int>> sw;
  permute_walk(ncols, marked, sel, nsel, sw);
  for (size_t i = 0; i < sw.size(); i++) {
    swaps[2 * i] = sw[i].first;
    swaps[2 * i + 1] = sw[i].second;
  }
  return (int)sw.size();
}

const char* qrdm_last_error(void) { return g_err.c_str(); }

int qrdm_dgeqrdm_dev_trace(int matrix_layout, int m, int n, double* d_a, int lda, int* jpvt,
                           double* d_tau, int* ncols, const double* thres, int nb, void* stream,
                           qrdm_step* trace, int ntrace, int* nsteps) {
  Args x{matrix_layout, m, n, lda, nb, d_a, jpvt, d_tau, ncols, thres,
         (hipStream_t)stream, trace, ntrace, nsteps};
  int bad = check_args(x);
  if (bad) return bad;
  return guarded(x);
}

int qrdm_dgeqrdm_dev_opt(int matrix_layout, int m, int n, double* d_a, int lda, int* jpvt,
                         double* d_tau, int* ncols, const double* thres, int nb,
                         const qrdm_opts* opts, void* stream, qrdm_step* trace, int ntrace,
                         int* nsteps) {
  Args x{matrix_layout, m, n, lda, nb, d_a, jpvt, d_tau, ncols, thres,
         (hipStream_t)stream, trace, ntrace, nsteps};
  int bad = check_args(x);
  if (bad) return bad;
  if (opts) {
    if (opts->trailing < QRDM_TRAILING_FP64 || opts->trailing > QRDM_TRAILING_BF16X3_ALL) {
      fprintf(stderr, "DGEQRDM: unknown trailing-update precision %d\n", opts->trailing);
      return -1;
    }
    x.trail = opts->trailing;
    x.policy = opts->policy;
  }
  return guarded(x);
}

int qrdm_dgeqrdm_dev(int matrix_layout, int m, int n, double* d_a, int lda, int* jpvt,
                     double* d_tau, int* ncols, const double* thres, int nb, void* stream) {
  return qrdm_dgeqrdm_dev_trace(matrix_layout, m, n, d_a, lda, jpvt, d_tau, ncols, thres, nb,
                                stream, nullptr, 0, nullptr);
}

int qrdm_dgeqrdm_dev_noalloc(int matrix_layout, int m, int n, double* d_a, int lda, int* jpvt,
                             double* d_tau, int* ncols, const double* thres, int nb,
                             void* stream) {
  Args x{matrix_layout, m, n, lda, nb, d_a, jpvt, d_tau, ncols, thres,
         (hipStream_t)stream, nullptr, 0, nullptr};
  int bad = check_args(x);
  if (bad) return bad;
  if (nb > KMAX) {
    fprintf(stderr,
            "qrdm_dgeqrdm_dev_noalloc: nb = %d: steps wider than %d are walked by the host with "
            "lists it allocates; use qrdm_dgeqrdm_dev\n",
            nb, KMAX);
    return -10;
  }
  const size_t need = layout_buffers(m, n, nullptr, nullptr);
  if (need > kScratchBytes) {
    fprintf(stderr,
            "qrdm_dgeqrdm_dev_noalloc: m = %d, n = %d needs %zu bytes of workspace, more than "
            "the %zu bytes of its scratch array\n",
            m, n, need, kScratchBytes);
    return -14;
  }
  return guarded_call([&] {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw HipError{hipErrorInvalidDevice, "device index above 63"};
    if (!g_scratch_ptr[dev]) {
      void* ptr = nullptr;
      HIPCHK(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
      g_scratch_ptr[dev] = ptr;
    }
    x.scratch = static_cast<char*>(g_scratch_ptr[dev]);
    return run_dm(x);
  });
}

int qrdm_dgeqrf_dev(int m, int n, double* d_a, int lda, double* d_tau, void* stream) {
  if (m < 0 || n < 0 || lda < (m > 1 ? m : 1)) return -1;
  if (m == 0 || n == 0) return 0;
  return guarded_call([&] { return run_geqrf(d_a, lda, m, n, d_tau, (hipStream_t)stream); });
}

int qrdm_dormqr_dev(int trans, int m, int n, int k, const double* d_a, int lda,
                    const double* d_tau, double* d_c, int ldc, void* stream) {
  if (m < 0 || n < 0 || k < 0 || k > m || lda < (m > 1 ? m : 1) || ldc < (m > 1 ? m : 1))
    return -1;
  if (m == 0 || n == 0 || k == 0) return 0;
  return guarded_call(
      [&] { return run_ormqr(trans, m, n, k, d_a, lda, d_tau, d_c, ldc, (hipStream_t)stream); });
}


int qrdm_dgeqrf(int m, int n, double* a, int lda, double* tau) {
  if (m < 0 || n < 0 || lda < (m > 1 ? m : 1)) return -1;
  if (m == 0 || n == 0) return 0;
  const int k = m < n ? m : n;
  try {
    DevBlock da, dt;
    to_device(da, a, lda, m, n);
    to_device(dt, tau, k, k, 1);
    int info = qrdm_dgeqrf_dev(m, n, da.p, da.ld, dt.p, nullptr);
    if (info != 0) return info;
    to_host(a, lda, da, m, n);
    to_host(tau, k, dt, k, 1);
    return 0;
  } catch (const HipError& e) {
    g_err = std::string(e.what) + ": " + hipGetErrorString(e.e);
    return -1000 - (int)e.e;
  }
}

int qrdm_dormqr(int trans, int m, int n, int k, const double* a, int lda, const double* tau,
                double* c, int ldc) {
  if (m < 0 || n < 0 || k < 0 || k > m || lda < (m > 1 ? m : 1) || ldc < (m > 1 ? m : 1))
    return -1;
  if (m == 0 || n == 0 || k == 0) return 0;
  try {
    DevBlock da, dt, dc;
    to_device(da, a, lda, m, k);
    to_device(dt, tau, k, k, 1);
    to_device(dc, c, ldc, m, n);
    int info = qrdm_dormqr_dev(trans, m, n, k, da.p, da.ld, dt.p, dc.p, dc.ld, nullptr);
    if (info != 0) return info;
    to_host(c, ldc, dc, m, n);
    return 0;
  } catch (const HipError& e) {
    g_err = std::string(e.what) + ": " + hipGetErrorString(e.e);
    return -1000 - (int)e.e;
  }
}

int dgeqrdm(int matrix_layout, int m, int n, double* a, int lda, int* jpvt, double* tau,
            int* ncols, double* thres, int nb) {
  Args x{matrix_layout, m, n, lda, nb, a, jpvt, tau, ncols, thres, nullptr, nullptr, 0, nullptr};
  int bad = check_args(x);
  if (bad) return bad;
  const int minmn = m < n ? m : n;
  double *d_a = nullptr, *d_tau = nullptr;
  int ldd = ((m + 31) / 32) * 32;
  try {
    HIPCHK(hipMalloc(&d_a, sizeof(double) * (size_t)ldd * n));
    HIPCHK(hipMalloc(&d_tau, sizeof(double) * (size_t)minmn));
    HIPCHK(hipMemcpy2D(d_a, sizeof(double) * ldd, a, sizeof(double) * lda, sizeof(double) * m, n,
                       hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_tau, tau, sizeof(double) * minmn, hipMemcpyHostToDevice));
  } catch (const HipError& e) {
    g_err = std::string(e.what) + ": " + hipGetErrorString(e.e);
    if (d_a) (void)hipFree(d_a);
    if (d_tau) (void)hipFree(d_tau);
    return -1000 - (int)e.e;
  }
  x.a = d_a;
  x.lda = ldd;
  x.tau = d_tau;
  int info = guarded(x);
  if (info > -1000) {
    hipError_t e1 = hipMemcpy2D(a, sizeof(double) * lda, d_a, sizeof(double) * ldd,
                                sizeof(double) * m, n, hipMemcpyDeviceToHost);
    hipError_t e2 = hipMemcpy(tau, d_tau, sizeof(double) * minmn, hipMemcpyDeviceToHost);
    if (e1 != hipSuccess || e2 != hipSuccess) {
      g_err = std::string("copy back: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2);
      info = -1000 - (int)(e1 != hipSuccess ? e1 : e2);
    }
  }
  (void)hipFree(d_a);
  (void)hipFree(d_tau);
  return info;
}


qrdm_batch_plan* qrdm_batch_plan_create(int count, const qrdm_batch_item* items, int nb) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_err.clear();
  if (count <= 0 || !items) {
    g_err = "batch: count must be positive";
    return nullptr;
  }
  if (nb > KMAX) {
    g_err = "batch: nb exceeds 64";
    return nullptr;
  }
  auto* p = new qrdm_batch_plan;
  p->count = count;
  p->nb_req = nb;
  p->nb_run = nb > 0 ? nb : KMAX;
  p->items.assign(items, items + count);
  for (int i = 0; i < count; i++) {
    const qrdm_batch_item& t = items[i];
    if (t.m < 1 || t.n < 1 || t.n > BATCH_NMAX || t.lda < t.m || t.a_off < 0 || t.tau_off < 0 ||
        t.col_off < 0) {
      g_err = "batch: item " + std::to_string(i) +
              " needs m >= 1, 1 <= n <= 1024, lda >= m and non-negative offsets";
      delete p;
      return nullptr;
    }
    const int k = t.m < t.n ? t.m : t.n;
    p->a_ext = std::max(p->a_ext, (size_t)t.a_off + (size_t)t.lda * (t.n - 1) + t.m);
    p->tau_ext = std::max(p->tau_ext, (size_t)t.tau_off + k);
    p->col_ext = std::max(p->col_ext, (size_t)t.col_off + t.n);
  }
  try {
    int dev = 0, cus = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    p->grid = std::max(1, std::min(cus, count));
    const std::vector<int> order = batch_order(p->items);
    HIPCHK(hipMalloc(&p->d_items, sizeof(qrdm_batch_item) * count));
    HIPCHK(hipMalloc(&p->d_sched, sizeof(int) * order.size()));
    HIPCHK(hipMalloc(&p->d_queue, sizeof(int) * 2));
    HIPCHK(hipMemcpy(p->d_items, items, sizeof(qrdm_batch_item) * count, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(p->d_sched, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
    // zero once here; every launch leaves it zero again
    HIPCHK(hipMemset(p->d_queue, 0, sizeof(int) * 2));
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking));
  } catch (const HipError& e) {
    g_err = std::string(e.what) + ": " + hipGetErrorString(e.e);
    qrdm_batch_plan_destroy(p);
    return nullptr;
  }
  return p;
}

void qrdm_batch_plan_destroy(qrdm_batch_plan* p) {
  if (!p) return;
  batch_release_graph(p);
  if (p->cap) (void)hipStreamDestroy(p->cap);
  if (p->d_items) (void)hipFree(p->d_items);
  if (p->d_sched) (void)hipFree(p->d_sched);
  if (p->d_queue) (void)hipFree(p->d_queue);
  delete p;
}

int qrdm_batch_plan_query(const qrdm_batch_plan* p, int* out, int nout) {
  if (!p || !out) return -1;
  const double vals[5] = {(double)p->waves, (double)p->grid, (double)p->nb_run,
                          (double)p->tuned_n, (double)p->gram_rows};
  for (int i = 0; i < nout && i < 5; i++) out[i] = (int)vals[i];
  return 5;
}

int qrdm_batch_plan_set_policy(qrdm_batch_plan* p, const void* d_policy) {
  if (!p) return -1;
  p->policy = d_policy;
  batch_release_graph(p);
  return 0;
}

int qrdm_batch_trace(qrdm_batch_plan* p, double* d_a, double* d_tau, int* d_jpvt, int* d_ncols,
                     int* d_info, const double* thres, float* d_records, int cap,
                     unsigned* d_count, const float* ranges, void* stream) {
  if (!p || !d_a || !d_tau || !d_jpvt || !d_ncols || !d_info || !thres || !d_records ||
      !d_count || !ranges || cap < 0)
    return -1;
  return guarded_call([&] {
    hipStream_t s = (hipStream_t)stream;
    const double th[3] = {thres[0], thres[1], 0.0};
    batch_launch(p, p->waves, p->nb_run, p->gram_rows, d_a, d_tau, d_jpvt, d_ncols, d_info, th, 0,
                 s, d_records, d_count, cap, ranges);
    return 0;
  });
}

int qrdm_batch_plan_set_shape(qrdm_batch_plan* p, int waves) {
  if (!p) return -1;
  if (waves != 4) {
    // 8 waves per workgroup (k_dm_batch<8>) is switched off: three runs of it
    // on a 4096-matrix batch ended in a host-heap abort whose cause is not
    // identified.
    g_err = "batch: the launch shape is 4 waves per workgroup (8 is switched off)";
    return -1;
  }
  p->waves = waves;
  batch_release_graph(p);
  return 0;
}

int qrdm_batch_plan_set_panel_rows(qrdm_batch_plan* p, int rows) {
  if (!p) return -1;
  p->gram_rows = rows > 0 ? rows : 0;
  batch_release_graph(p);
  return 0;
}

int qrdm_batch_autotune(qrdm_batch_plan* p, const double* d_a, const double* thres,
                        void* stream) {
  if (!p || !d_a || !thres) return -1;
  return guarded_call([&] {
    hipStream_t s = (hipStream_t)stream;
    double *a = nullptr, *tau = nullptr;
    int *jp = nullptr, *nc = nullptr, *info = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&] {
      if (a) (void)hipFree(a);
      if (tau) (void)hipFree(tau);
      if (jp) (void)hipFree(jp);
      if (nc) (void)hipFree(nc);
      if (info) (void)hipFree(info);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
    };
    try {
      HIPCHK(hipMalloc(&a, sizeof(double) * p->a_ext));
      HIPCHK(hipMalloc(&tau, sizeof(double) * p->tau_ext));
      HIPCHK(hipMalloc(&jp, sizeof(int) * p->col_ext));
      HIPCHK(hipMalloc(&nc, sizeof(int) * p->col_ext));
      HIPCHK(hipMalloc(&info, sizeof(int) * p->count));
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      const int nbs_auto[3] = {16, 32, 64};
      const int rows_c[3] = {0, 128, 256};  // the panel form: column by column, or Gram-based
      const int nnb = p->nb_req > 0 ? 1 : 3;
      float best = 1e30f;
      int bw = p->waves, bnb = p->nb_run, brows = p->gram_rows;
      p->tuned_n = 0;
      for (int ib = 0; ib < nnb; ib++) {
        const int nb = p->nb_req > 0 ? p->nb_req : nbs_auto[ib];
        for (int ir = 0; ir < 3; ir++) {
          const int waves = 4;  // the one launch shape (8 waves is switched off)
          float t = 0.f;
          for (int rep = 0; rep < 2; rep++) {  // the first run warms up
            HIPCHK(hipMemcpyAsync(a, d_a, sizeof(double) * p->a_ext, hipMemcpyDeviceToDevice, s));
            HIPCHK(hipEventRecord(e0, s));
            batch_launch(p, waves, nb, rows_c[ir], a, tau, jp, nc, info, thres, 0, s);
            HIPCHK(hipEventRecord(e1, s));
            HIPCHK(hipEventSynchronize(e1));
            HIPCHK(hipEventElapsedTime(&t, e0, e1));
          }
          if (p->tuned_n < 16) p->tuned_ms[p->tuned_n++] = t;
          if (t < best) {
            best = t;
            bw = waves;
            bnb = nb;
            brows = rows_c[ir];
          }
        }
      }
      p->waves = bw;
      p->nb_run = bnb;
      p->gram_rows = brows;
      batch_release_graph(p);
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
    return 0;
  });
}

int qrdm_dgeqrdm_batched(qrdm_batch_plan* p, double* d_a, double* d_tau, int* d_jpvt,
                         int* d_ncols, int* d_info, const double* thres, int stop_mode,
                         int use_graph, void* stream) {
  if (!p || !d_a || !d_tau || !d_jpvt || !d_ncols || !d_info || !thres) return -1;
  if (thres[0] < 0.0 || thres[0] > 1.0 || thres[1] < 0.0 || thres[1] > 1.0 || stop_mode < 0 ||
      stop_mode > 3)
    return -1;
  return guarded_call([&] {
    hipStream_t s = (hipStream_t)stream;
    const double th[3] = {thres[0], thres[1], stop_mode == 3 ? thres[2] : 0.0};
    if (!use_graph) {
      batch_launch(p, p->waves, p->nb_run, p->gram_rows, d_a, d_tau, d_jpvt, d_ncols, d_info, th,
                   stop_mode, s);
      return 0;
    }
    qrdm_batch_plan::Key k{d_a, d_tau, d_jpvt, d_ncols, d_info, {th[0], th[1], th[2]},
                           stop_mode, p->waves, p->nb_run, p->gram_rows};
    if (!p->exec || !(k == p->key)) {
      batch_release_graph(p);
      HIPCHK(hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal));
      const int lr = launch_dm_batch(p->waves, p->grid, p->d_items, p->d_sched, p->count,
                                     p->d_queue, d_a, d_tau, d_jpvt, d_ncols, d_info, p->nb_run,
                                     th[0], th[1],
                                     stop_mode, th[2], sqrt(kEps), p->policy, nullptr, nullptr, 0,
                                     nullptr, p->gram_rows, p->cap);
      HIPCHK(hipStreamEndCapture(p->cap, &p->graph));
      if (lr != 0) {
        batch_release_graph(p);
        throw HipError{hipErrorInvalidValue, "batch launch shape"};
      }
      HIPCHK(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
      p->key = k;
    }
    HIPCHK(hipGraphLaunch(p->exec, s));
    return 0;
  });
}

int qrdm_policy_sizes(int* out, int nout) {
  const int v[3] = {policy_nparam(), (int)policy_net_bytes(), POLICY_REC};
  for (int i = 0; i < nout && i < 3; i++) out[i] = v[i];
  return 3;
}

static float* policy_partials_buffer() {
  static float* buf[64] = {};
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (!buf[dev & 63])
    HIPCHK(hipMalloc(&buf[dev & 63], sizeof(float) * ((size_t)policy_partials() *
                                                      (policy_nparam() + 1))));
  return buf[dev & 63];
}

int qrdm_policy_grad(const float* d_theta, const float* d_x, const float* d_y, int ns,
                     float* d_grad, float* d_loss, void* stream) {
  if (!d_theta || !d_x || !d_y || !d_grad || !d_loss || ns <= 0) return -1;
  return guarded_call([&] {
    hipStream_t s = (hipStream_t)stream;
    float* part = policy_partials_buffer();
    float* lpart = part + (size_t)policy_partials() * policy_nparam();
    launch_policy_grad(d_theta, d_x, d_y, ns, part, lpart, s);
    launch_policy_update(nullptr, nullptr, nullptr, part, lpart, d_grad, d_loss, 0, 0.f, 0.f,
                         0.f, 0.f, 1, s);
    HIPCHK(hipGetLastError());
    return 0;
  });
}

int qrdm_policy_adam(float* d_theta, float* d_m1, float* d_m2, const float* d_x,
                     const float* d_y, int ns, float lr, int t, float* d_loss, void* stream) {
  if (!d_theta || !d_m1 || !d_m2 || !d_x || !d_y || !d_loss || ns <= 0 || t < 1) return -1;
  return guarded_call([&] {
    hipStream_t s = (hipStream_t)stream;
    float* part = policy_partials_buffer();
    float* lpart = part + (size_t)policy_partials() * policy_nparam();
    launch_policy_grad(d_theta, d_x, d_y, ns, part, lpart, s);
    launch_policy_update(d_theta, d_m1, d_m2, part, lpart, nullptr, d_loss, 1, lr, 0.9f,
                         0.999f, 1e-8f, t, s);
    HIPCHK(hipGetLastError());
    return 0;
  });
}

int qrdm_policy_pack(const float* d_theta, const float* ranges, void* d_net, void* stream) {
  if (!d_theta || !ranges || !d_net) return -1;
  return guarded_call([&] {
    launch_policy_pack(d_theta, d_net, ranges[0], ranges[1], ranges[2], ranges[3],
                       (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
  });
}

int qrdm_policy_predict(const void* d_net, const float* d_x, int ns, float* d_out,
                        void* stream) {
  if (!d_net || !d_x || !d_out || ns <= 0) return -1;
  return guarded_call([&] {
    launch_policy_predict(d_net, d_x, ns, d_out, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return 0;
  });
}
}  // extern "C"
