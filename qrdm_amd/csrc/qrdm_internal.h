// Internal interface between the host driver (driver.hip) and the kernels
// (kernels.hip) of libqrdm_amd.so.  Not installed; not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qrdm {

constexpr int KMAX = 64;          // largest DM block the kernels handle
// Widest DM step (nb) a call takes; steps wider than KMAX run host-driven.
constexpr int KW = 256;
constexpr int GRAM_ROWS = 64;     // rows per workgroup of the Gram kernel

// Record of one completed DM step (kept for tracing).
struct StepRec {
  int j, kmax, nc, fjb, fjb_cmp, nswap, panel;
  double maxnrm;
  double tau, delta;  // the DM parameters the step used
};

// Algorithmic work of the steps run so far, per kernel class (flops or HBM
// bytes), accumulated on the device for the kernel-time accounting.  With the
// merged trailing update (fused), gemm1 holds the fused passes' flops (the
// previous step's update and this step's V^T C) and wsum those of W = T^T
// sum Wp and the R12 rows; gemm2 and norm stay zero.
struct StepWork {
  double select, perm, panel, gemm1, gemm2, norm, wsum;
};

// Per-factorisation state shared by the kernels.  Lives in device memory.
// The DM loop is driven from the device: every kernel of a step reads the
// step's column j (and what earlier kernels of the step decided) from here,
// so the host queues steps without waiting for them.
struct DevState {
  int nc;          // number of norm candidates
  int fjb;         // accepted columns
  int fjb_cmp;     // columns reduced by the panel
  int flags;       // bit 0: non-finite value met; bit 1: fast panel fell back
  int k1;          // columns the fast panel's first Cholesky factored
  unsigned long long maxnrm_bits;  // max partial norm (bit pattern, >= 0)
  unsigned long long recomp_rows;  // rows read by norm recomputes (profiling)
  int sel[KMAX];        // accepted columns (relative to j), acceptance order
  int selpos[KMAX];     // their candidate ranks
  int cand_idx[KMAX];   // candidates (relative to j), sorted
  double cand_val[KMAX];
  int j;       // first column of the current step
  int it;      // steps completed
  int done;    // 1: the loop has ended; kernels of later steps return at once
  int kmax;    // min(nb, m - j, n - j, KMAX) of the current step
  int fast;    // 1: the current step's panel is the Gram-based one
  int c0;      // first column of the current step's trailing update
  int npair;   // (dst, src) column pairs of the current step's permutation
  int nswap;   // swaps made by the current step's permutation walk
  int pend;    // 1: the current step's end (k_step_end's work) has not run yet
  int halt;    // the step needs the exact panel, which was not queued; the DM
               // kernels return until the host resumes the step.  2: the step's
               // column moves still run (set before them); the next step's
               // top-k turns it into 1, which stops the moves too
  int fused;   // 1: the trailing update is deferred and merged into the next step's pass
               // (trailing.hip); 0: applied within its own step
  int kk_prev;  // fjb_cmp of the previous step (0 before the first): the width of the
                // update the fused pass applies
  int policy_on;      // 1: the step's tau and delta come from the learned policy
  double tau_step;    // tau and delta of the current step (the defaults, or the
  double delta_step;  // policy's proposal)
  StepRec last;
  StepWork work;
};

// The DM loop's reflector buffers: the dense V (ldv x KMAX, zero above the
// panel) and W = T^T V^T C (KMAX x n, by absolute column, column-major: W(i, c)
// at W[c * KMAX + i]) of the current step and of the previous one,
// alternating with the step's parity: step it writes buffer it & 1.  With
// st->fused the previous step's update C -= V W is still to be applied
// (trailing.hip).  vstride = 0 / wstride = 0: one buffer.
struct StepBufs {
  double* V;
  size_t vstride;
  int ldv;
  double* W;
  size_t wstride;
};
__device__ __forceinline__ double* v_cur(const StepBufs& b, const DevState* st) {
  return b.V + (size_t)(st->it & 1) * b.vstride;
}
__device__ __forceinline__ double* v_prev(const StepBufs& b, const DevState* st) {
  return b.V + (size_t)((st->it & 1) ^ 1) * b.vstride;
}
__device__ __forceinline__ double* w_cur(const StepBufs& b, const DevState* st) {
  return b.W + (size_t)(st->it & 1) * b.wstride;
}
__device__ __forceinline__ double* w_prev(const StepBufs& b, const DevState* st) {
  return b.W + (size_t)((st->it & 1) ^ 1) * b.wstride;
}
// Whether the previous step's update is still to be applied.
__device__ __forceinline__ bool update_pending(const DevState* st) {
  return st->fused && st->it > 0;
}

// The workitem id through an empty asm: what a helper derives from it (LDS
// offsets, lane masks) is computed where the helper runs, not hoisted to the
// top of a persistent kernel's loop and held in registers across its other
// phases.
__device__ __forceinline__ int local_tid() {
  int t = threadIdx.x;
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(t));
#endif
  return t;
}

// Value of x in lane l of the wave (l a compile-time or wave-uniform index).
__device__ __forceinline__ double rdlane(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Value of x in lane i (0 <= i < 16, a compile-time index once the caller's
// loop is unrolled) of each 16-lane row, in every lane of that row: the DPP
// row broadcast (row_newbcast:i), register to register, no scalar registers.
__device__ __forceinline__ int row_bcast_i32(int x, int i) {
#define QRDM_RB(I) \
  case I:          \
    return __builtin_amdgcn_mov_dpp(x, 0x150 + I, 0xF, 0xF, false);
  switch (i) {
    QRDM_RB(0) QRDM_RB(1) QRDM_RB(2) QRDM_RB(3) QRDM_RB(4) QRDM_RB(5) QRDM_RB(6) QRDM_RB(7)
    QRDM_RB(8) QRDM_RB(9) QRDM_RB(10) QRDM_RB(11) QRDM_RB(12) QRDM_RB(13) QRDM_RB(14)
    default:
      return __builtin_amdgcn_mov_dpp(x, 0x150 + 15, 0xF, 0xF, false);
  }
#undef QRDM_RB
}
__device__ __forceinline__ double row_bcast(double x, int i) {
  const long long b = __double_as_longlong(x);
  const int lo = row_bcast_i32((int)b, i);
  const int hi = row_bcast_i32((int)(b >> 32), i);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Orders a wave's LDS accesses: its earlier stores are visible to its later
// loads (one wave's LDS operations complete in order).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Data handed from the workgroups of a launch to its last workgroup (atomic
// ticket): stores and loads coherent at device scope (performed at the
// device's coherence point, past every XCD's L2), so the ticket needs no
// cache-wide write-back; each wave waits for its own stores to complete
// (handoff_wait) before the barrier that precedes the ticket.
__device__ __forceinline__ void st_dev(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_dev(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_dev(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_dev(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void handoff_wait() { __builtin_amdgcn_s_waitcnt(0); }
// Called by one thread after every wave's handoff_wait and a barrier: whether
// this workgroup is the last of n to arrive; the last one resets the ticket
// for the next launch.
__device__ __forceinline__ bool last_to_arrive(unsigned* ticket, unsigned n) {
  const bool last =
      __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == n - 1;
  if (last) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return last;
}

// Which kernels of a step run: host-driven callers (RUN_ALWAYS) pass their
// own geometry; the DM loop's kernels (RUN_DM) take j from DevState and
// return once the loop is done; RUN_EXACT / RUN_FAST kernels belong to one
// of the two panel forms and return unless the step uses that form.
enum RunMode { RUN_ALWAYS = 0, RUN_DM = 1, RUN_EXACT = 2, RUN_FAST = 3 };

// The DM loop has ended, or waits for the host to resume a halted step.
__device__ __forceinline__ bool dm_stopped(const DevState* st) { return st->done || st->halt; }

__device__ __forceinline__ bool step_runs(const DevState* st, int mode) {
  if (mode == RUN_ALWAYS) return true;
  if (dm_stopped(st)) return false;
  if (mode == RUN_EXACT) return !st->fast || (st->flags & 2);
  if (mode == RUN_FAST) return st->fast && !(st->flags & 2);
  return true;
}

// ---- launchers (kernels.hip) ------------------------------------------------
// Kernels that take a RunMode read j (and the other per-step values named
// below) from st when mode != RUN_ALWAYS; the host arguments are then upper
// bounds used only to size the grid.
void launch_init_norms(const double* A, int lda, int r0, int m, int n, double* work,
                       double* workold, DevState* st, hipStream_t s);
// The pending end of the previous step (see launch_step_end), then the
// sorted top-kmax of work[j..n) (value descending, index ascending) and the
// candidate count (> tau * max), into st.  kmax = min(nb, m-j, n-j, KMAX).
// j0 <= j sizes the grid; tmp_val/tmp_idx hold topk_blocks(n) * KMAX
// entries; *ticket is 0 between launches.
int topk_blocks(int n);
void launch_topk(const double* work, int j0, int m, int n, int nb, double tau, double* tmp_val,
                 int* tmp_idx, unsigned* ticket, DevState* st, int* ncols_dev, int enable_stop,
                 double eta, hipStream_t s);
// Gram of ncols (<= 64) columns over rows [r0, r1): columns are
// base[(cols ? cols[c] : c) * ld + r].  partial needs ceil((r1-r0)/GRAM_ROWS)
// * 64 * 64 doubles.  G is 64 x 64 column-major; entries beyond ncols are 0.
// mode RUN_DM: base and r0 move by st->j columns and rows (the candidates of
// the step); RUN_EXACT: r0 = st->j (the step's V), run only with the exact
// panel.  r0 is then the lower bound j0.
void launch_gram(const double* base, int ld, int r0, int r1, const int* cols, int ncols_max,
                 const int* ncols_dev, double* partial, double* G, hipStream_t s,
                 const DevState* st = nullptr, int mode = RUN_ALWAYS);
// The step's candidate Gram over rows j..m (RUN_DM geometry; j0 <= j sizes
// the grid), then, in the last workgroup of the reduce (*ticket is 0 between
// launches), the greedy cosine acceptance and the step's plan: permute_marked's
// walk, the (dst, src) column pairs into perm[0..4*KMAX), the panel columns'
// candidate ranks into perm[4*KMAX..5*KMAX), the panel form.
// exact_queued = 0: the step's exact-panel kernels are not queued, and a step
// that needs them (short panel, fjb < 2) halts for the host to resume it.
// With st->fused and a previous step, the candidate columns first receive
// that step's deferred update (all rows from j, written back), and their W
// columns are cleared.
void launch_gram_select(double* A, int lda, int j0, int m, int n, double* partial,
                        double* G, double delta, int fast_rows, int exact_queued, int* perm,
                        unsigned* ticket, DevState* st, const StepBufs& bufs, hipStream_t s);
// Moves the step's columns (st->npair pairs in perm), their work entries and
// jpvt entries (jp may be null).
// With a pending update the moved columns' W entries move with them.
void launch_apply_perm(double* A, int lda, int m, double* work, int* jp, const int* perm,
                       DevState* st, const StepBufs& bufs, hipStream_t s);
// Column-by-column panel on columns [j, j + fjb) (RUN_ALWAYS: host j, fjb;
// RUN_EXACT: from st, tau indexed from j).  Sets st->fjb_cmp and st->c0.
void launch_panel_exact(double* A, int lda, int j, int m, int fjb, double tau_dm,
                        int truncate, double* tau, DevState* st, hipStream_t s,
                        int mode = RUN_ALWAYS);
void launch_build_v(const double* A, int lda, int j, int m, int mpad, const DevState* st,
                    double* V, int ldv, hipStream_t s, int mode = RUN_ALWAYS);
// T from S = V^T V; with nslab > 0, S is summed from that many partial Grams.
void launch_larft(const double* S, const double* tau, const DevState* st, double* T,
                  hipStream_t s, int mode = RUN_ALWAYS, int nslab = 0);
// The DM loop's exact-panel V and its partial Grams (RUN_EXACT; j0 <= j
// sizes the grid: vgram_exact_slabs(j0, m) partials).
int vgram_exact_slabs(int j0, int m);
void launch_vgram_exact(const double* A, int lda, int j0, int m, const DevState* st,
                        const StepBufs& bufs, double* partial, hipStream_t s);
// Block reflector from the left on rows [j, m), columns [c0, n):
// trans = 1: C := (I - V T V^T)^T C (dgeqrdm's trailing update);
// trans = 0: C := (I - V T V^T) C.
void launch_larfb(double* A, int lda, int j, int m, int c0, int n, const double* V, int ldv,
                  const double* T, int trans, double* Wp, double* W, int nsplit, hipStream_t s);
// The two halves of launch_larfb: W = op(T) V^T C, then C -= V W.  With st
// (the DM loop) j and c0 come from st and the host values are lower bounds.
void launch_larfb_w(double* A, int lda, int j, int m, int c0, int n, const double* V, int ldv,
                    const double* T, int trans, double* Wp, double* W, int nsplit, hipStream_t s,
                    const DevState* st = nullptr);
void launch_larfb_apply(double* A, int lda, int j, int m, int c0, int n, const double* V, int ldv,
                        const double* W, hipStream_t s, const DevState* st = nullptr);
// Precision of the trailing update's two products (mixed.hip): FP64 runs
// launch_gemm1 / launch_larfb_apply; FP32 and BF16X3 the matrix-core forms
// with FP32 accumulation (C stays FP64).
enum { TRAIL_FP64 = 0, TRAIL_FP32 = 1, TRAIL_BF16X3 = 2 };
void launch_gemm1_prec(int prec, double* A, int lda, int j, int m, int c0, int n,
                       const double* V, int ldv, double* Wp, int nsplit, hipStream_t s,
                       const DevState* st);
void launch_apply_prec(int prec, double* A, int lda, int j, int m, int c0, int n,
                       const double* V, int ldv, const double* W, hipStream_t s,
                       const DevState* st);
// Batched DM-QR (batched.hip): one persistent launch of grid workgroups of
// 64 * waves threads; items points to device BatchItem records
// (qrdm_batch_item); the workgroups take the matrices order[0..count) in
// turn from the counter queue[0] (queue: two ints, zero before the first
// launch; the last workgroup out zeroes them again, so launches of one
// queue must not overlap).
// policy: a PolicyNet (learned.hip) proposing each step's (tau, delta), or
// nullptr; trace: per-step training records, or nullptr (ranges: the label
// normalisation tau_lo, tau_hi, delta_lo, delta_hi).
constexpr int BATCH_NMAX = 1024;
// gram_rows: panels of at least that many rows (and two or more columns) take
// the Gram-based form (panel.hip's, inside the workgroup), the others the
// column-by-column one; <= 0: column by column only.
// Returns -1 (nothing launched) for a launch shape other than 4 waves.
int launch_dm_batch(int waves, int grid, const void* items, const int* order, int count,
                     int* queue, double* A, double* tau, int* jpvt, int* ncols,
                     int* info, int nb, double delta, double tau_dm, int stop_mode, double eta3,
                     double tol3z, const void* policy, float* trace, unsigned* trace_count,
                     int trace_cap, const float* ranges, int gram_rows, hipStream_t s);
// The learned (tau, delta) proposal (learned.hip).
int policy_nparam();
size_t policy_net_bytes();
int policy_partials();
constexpr int POLICY_REC = 40;  // floats per training record
void launch_policy(const double* work, int m, int n, const void* net, DevState* st,
                   hipStream_t s);
void launch_policy_grad(const float* theta, const float* x, const float* y, int ns,
                        float* partial, float* loss_partial, hipStream_t s);
void launch_policy_update(float* theta, float* m1, float* m2, const float* partial,
                          const float* loss_partial, float* grad, float* loss_out, int adam,
                          float lr, float beta1, float beta2, float eps, int t, hipStream_t s);
void launch_policy_pack(const float* theta, void* net, float tau_lo, float tau_hi,
                        float delta_lo, float delta_hi, hipStream_t s);
void launch_policy_predict(const void* net, const float* x, int ns, float* out, hipStream_t s);
// launch_larfb_w's two launches: the partial products Wp = V^T C per row
// slab, then W = op(T) sum Wp.
void launch_gemm1(double* A, int lda, int j, int m, int c0, int n, const double* V, int ldv,
                  double* Wp, int nsplit, hipStream_t s, const DevState* st);
void launch_wsum(int j, int m, int c0, int n, const double* T, int trans, const double* Wp,
                 double* W, int nsplit, hipStream_t s, const DevState* st);
// Norm downdate of columns [j + fjb_cmp, n) (j, fjb_cmp from st; j0 <= j
// sizes the grid).
void launch_norm_update(const double* A, int lda, int j0, int m, int n, double* work,
                        double* workold, DevState* st, double tol3z, hipStream_t s);
// End of a DM step: ncols, the step record and work, the stop test
// (enable_stop, eta absolute), the next j, done.  k_norm_update marks it
// pending; the next step's k_topk runs it, and this launch runs it after the
// last step queued.
void launch_step_end(DevState* st, int* ncols_dev, int m, int n, int enable_stop, double eta,
                     hipStream_t s);
// Start of the DM loop at column j.
void launch_dm_begin(DevState* st, int j, int minmn, double tau, double delta, int fused,
                     hipStream_t s);

void launch_set_kk(DevState* st, int kk, hipStream_t s);
// ---- wide DM steps, 64 < nb <= KW (wide.hip): host-driven, one sync per step.
// The top kmax (<= KW) of work[0..nsub) (value descending, index ascending)
// by sort passes through the two key buffers (wide_sort_keys(n) entries
// each); *out_val gets the buffer holding them, wide_topk_idx its indices.
int wide_sort_keys(int n);
void launch_topk_wide(const double* work, int nsub, int kmax, double* val[2], int* idx[2],
                      const double** out_val, hipStream_t s);
const int* wide_topk_idx(double* val[2], int* idx[2], const double* out_val);
// Gram of the nc candidate columns j + idx[c] over rows [j, m) (tile pairs
// on or above the diagonal of 64 x 64 tiles; G column-major, ld ldg).
void launch_gram_wide(const double* A, int lda, int j, int m, const int* idx, int nc, double* G,
                      int ldg, hipStream_t s);
// out[0] = nc, out[1] = fjb, out[2 + s] = rank of the s-th accepted candidate.
void launch_select_wide(const double* G, int ldg, const double* val, int kmax, double tau,
                        double delta, int* out, hipStream_t s);
// np position swaps (pairs relative to column j) of A's columns, work, jpvt.
void launch_swap_pairs(double* A, int lda, int m, int j, double* work, int* jp, const int* pairs,
                       int np, hipStream_t s);
// st->j = j, st->fjb_cmp = kk, the running max cleared.
void launch_wide_state(DevState* st, int j, int kk, hipStream_t s);
// Tall-panel DM QR without per-column grid reductions (panel.hip), RUN_FAST
// kernels: j, fjb from st.  launch_perm_chol1 moves the step's columns (as
// launch_apply_perm) and factors the panel's Gram, gathered from the candidate
// Gram Gc (64 x 64, candidate order) by the panel columns' candidate ranks
// perm[4*KMAX..]; launch_panel_fast then runs the rest of the panel.  j0 <= j
// sizes the grids.  Sets st->fjb_cmp and st->c0, or the fallback flag.
void launch_perm_chol1(double* A, int lda, int m, double* work, int* jp, const int* perm,
                       const double* Gc, double tau_dm, double* R1, double* R1inv, DevState* st,
                       const StepBufs& bufs, hipStream_t s);
void launch_panel_fast(double* A, int lda, int j0, int m, double tau_dm, double* tau,
                       double* gram_part, double* R1, double* R1inv, double* G2, double* T,
                       double* M, const StepBufs& bufs, unsigned* ticket, DevState* st,
                       hipStream_t s);
int apply_perm_blocks(int m);
// Clears the halt of a step whose exact panel the host queues next.
void launch_resume(DevState* st, hipStream_t s);
void launch_swap2(double* A, int lda, int m, int p, int q, hipStream_t s);
// Zeroes bytes (a multiple of 4) at p with a kernel on s.
void launch_zero(void* p, size_t bytes, hipStream_t s);
// Sets DevState::fused (the form of the following steps' trailing update).
void launch_set_fused(DevState* st, int fused, hipStream_t s);

// ---- the merged trailing update (trailing.hip) -------------------------------
// Row slabs of a fused pass over rows [r_lo, m), columns [c0, n) (at most
// fused_max_slabs()); the partials Wp hold that many KMAX x n blocks.
int fused_slabs(int r_lo, int m, int c0, int n);
int fused_max_slabs();
// One pass over C (rows [r_lo, m), columns [c0, n)): apply: C -= V_prev
// W_prev; accum: Wp[slab] = V_cur^T C.  With st (the DM loop) r_lo and c0
// come from st (host values: lower bounds sizing the grid) and the buffers
// from the step's parity; without, buffer `cur` is the current one.
void launch_fused(double* A, int lda, int r_lo, int m, int c0, int n, const StepBufs& bufs,
                  double* Wp, int ns, int apply, int accum, int cur, hipStream_t s,
                  const DevState* st);
// The end of a fused step: W = T^T sum_slab Wp for the columns [j + kk, n)
// (zero below c0), the step's R12 rows C_top - V_top W (written), the norm
// downdate and recompute of those columns (the recomputed ones get the whole
// update now and a zero W column), into the current W buffer.  j0 <= j sizes
// the grid; ns as the fused pass.
void launch_wsum_norm(double* A, int lda, int j0, int m, int n, const double* T,
                      const double* Wp, int ns, const StepBufs& bufs, double* work,
                      double* workold, DevState* st, double tol3z, hipStream_t s);

}  // namespace qrdm
