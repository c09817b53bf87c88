// Internal interface between the host driver (driver.hip) and the kernels
// (kernels.hip) of libqrdm_amd.so.  Not installed; not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qrdm {

constexpr int KMAX = 64;          // largest DM block the kernels handle
constexpr int TOPK_CHUNK = 1024;  // keys sorted per workgroup by the top-k stage
constexpr int GRAM_ROWS = 64;     // rows per workgroup of the Gram kernel

// Small per-factorisation state shared by the kernels of one step.  Lives in
// device memory; the host reads it back between kernels where it must.
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
};

// ---- launchers (kernels.hip) ------------------------------------------------
void launch_init_norms(const double* A, int lda, int r0, int m, int n, double* work,
                       double* workold, DevState* st, hipStream_t s);
// Sorted top-k of keys[0..n) (value descending, index ascending).  tmp_val/
// tmp_idx hold 2 * half entries, half >= ceil(n/TOPK_CHUNK)*KMAX.
void launch_topk(const double* keys, int n, int kmax, double tau, double* tmp_val,
                 int* tmp_idx, int half, DevState* st, hipStream_t s);
// Gram of ncols (<= 64) columns over rows [r0, r1): columns are
// base[(cols ? cols[c] : c) * ld + r].  partial needs ceil((r1-r0)/GRAM_ROWS)
// * 64 * 64 doubles.  G is 64 x 64 column-major; entries beyond ncols are 0.
void launch_gram(const double* base, int ld, int r0, int r1, const int* cols, int ncols_max,
                 const int* ncols_dev, double* partial, double* G, hipStream_t s);
void launch_greedy(const double* G, double delta, int kmax, DevState* st, hipStream_t s);
void launch_apply_perm(double* A, int lda, int m, double* work, const int* perm, int npair,
                       hipStream_t s);
void launch_panel_exact(double* A, int lda, int j, int msub, int fjb, double tau_dm,
                        int truncate, double* tau, DevState* st, hipStream_t s);
void launch_build_v(const double* A, int lda, int j, int m, int mpad, const DevState* st,
                    double* V, int ldv, hipStream_t s);
void launch_larft(const double* S, const double* tau, const DevState* st, double* T,
                  hipStream_t s);
// Block reflector from the left on rows [j, m), columns [c0, n):
// trans = 1: C := (I - V T V^T)^T C (dgeqrdm's trailing update);
// trans = 0: C := (I - V T V^T) C.
void launch_larfb(double* A, int lda, int j, int m, int c0, int n, const double* V, int ldv,
                  const double* T, int trans, double* Wp, double* W, int nsplit,
                  unsigned* tickets, hipStream_t s);
// the two halves of launch_larfb: W = op(T) V^T C, then C -= V W
void launch_larfb_w(double* A, int lda, int j, int m, int c0, int n, const double* V, int ldv,
                    const double* T, int trans, double* Wp, double* W, int nsplit,
                    unsigned* tickets, hipStream_t s);
void launch_larfb_apply(double* A, int lda, int j, int m, int c0, int n, const double* V, int ldv,
                        const double* W, hipStream_t s);
void launch_norm_update(const double* A, int lda, int j, int m, int n, double* work,
                        double* workold, DevState* st, double tol3z, hipStream_t s);

void launch_set_kk(DevState* st, int kk, hipStream_t s);
void launch_gram_reduce(const double* partial, int nslab, double* G, hipStream_t s);
// Tall-panel DM QR without per-column grid reductions (panel.hip).  Gc is
// the step's candidate Gram (64 x 64, candidate order) and rank[p] the
// candidate rank of panel column p.
void launch_panel_fast(double* A, int lda, int j, int m, int fjb, double tau_dm, double* tau,
                       const double* Gc, const int* rank, double* gram_part, double* R1,
                       double* R1inv, double* G2, double* T, double* M, double* V, int ldv,
                       DevState* st, hipStream_t s);
void launch_clear_flag(DevState* st, int bits, hipStream_t s);
void launch_swap2(double* A, int lda, int m, int p, int q, hipStream_t s);

}  // namespace qrdm
