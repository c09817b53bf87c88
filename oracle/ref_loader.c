/*
 * ref_loader.c -- lazily binds the reference compiled from its own sources
 * (oracle/_ref/libqrdm_ref.so, made by oracle/build_ref.sh).  TEST
 * INFRASTRUCTURE ONLY: used by the tests and by bench.py's cpu_baseline leg.
 *
 * The reference object keeps two references that resolve nowhere in its tree
 * (DM_perm2, apply_jpvt; src/dgeqrdm_work.c:721,808), both on branches that
 * never execute.  It therefore has to be opened with RTLD_LAZY, which Python's
 * ctypes does not do; this loader does it and forwards the calls.
 */
#include <dlfcn.h>
#include <stdio.h>

typedef int (*ref_dgeqrdm_fn)(int, int, int, double *, int, int *, double *, int *,
                              double *, int);
typedef int (*ref_dgeqp3_fn)(int, int, int, double *, int, int *, double *);
typedef void (*ref_setthreads_fn)(int);

static void *g_ref = NULL, *g_blas = NULL;
static ref_dgeqrdm_fn g_dgeqrdm = NULL;
static ref_dgeqp3_fn g_dgeqp3 = NULL;
static ref_setthreads_fn g_setthreads = NULL;

/* Returns 0 on success; the paths are the compiled reference and the BLAS it
 * links (for dgeqp3 and the thread count). */
int oracle_ref_open(const char *ref_path, const char *blas_path) {
  if (!g_ref) {
    g_ref = dlopen(ref_path, RTLD_LAZY | RTLD_LOCAL);
    if (!g_ref) {
      fprintf(stderr, "oracle_ref_open: %s\n", dlerror());
      return -1;
    }
    g_dgeqrdm = (ref_dgeqrdm_fn)dlsym(g_ref, "dgeqrdm");
  }
  if (!g_blas && blas_path) {
    g_blas = dlopen(blas_path, RTLD_LAZY | RTLD_LOCAL);
    if (g_blas) {
      g_dgeqp3 = (ref_dgeqp3_fn)dlsym(g_blas, "scipy_LAPACKE_dgeqp3");
      g_setthreads = (ref_setthreads_fn)dlsym(g_blas, "scipy_openblas_set_num_threads");
    }
  }
  return g_dgeqrdm ? 0 : -2;
}

int oracle_ref_dgeqrdm(int layout, int m, int n, double *a, int lda, int *jpvt, double *tau,
                       int *ncols, double *thres, int nb) {
  if (!g_dgeqrdm) return -1000;
  return g_dgeqrdm(layout, m, n, a, lda, jpvt, tau, ncols, thres, nb);
}

int oracle_ref_dgeqp3(int layout, int m, int n, double *a, int lda, int *jpvt, double *tau) {
  if (!g_dgeqp3) return -1000;
  return g_dgeqp3(layout, m, n, a, lda, jpvt, tau);
}

int oracle_ref_set_threads(int n) {
  if (!g_setthreads) return -1;
  g_setthreads(n);
  return 0;
}
