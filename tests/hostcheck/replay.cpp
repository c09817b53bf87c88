// Replays, on the host-only build of the library (stub_runtime.cpp), the C-ABI
// call sequence of the batched measurement: a 4096-matrix plan (launch shape
// requests, autotune, hipGraph replays and an eager launch), then one
// qrdm_dgeqrdm_dev call per matrix with host jpvt/ncols arrays of exactly n
// entries, then the plan's destruction; and the library's other host paths
// (traced and option calls, fixed columns, wide steps, the allocation-free
// entry point, the host-pointer entry point).  Run under AddressSanitizer: any
// host-side overrun or use after free in those paths ends it with a report.
//
// usage: replay <shapes file: one "m n" per line>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/qrdm_amd.h"

namespace {
int fails = 0;
void check(bool ok, const char* what, int v) {
  if (!ok) {
    fprintf(stderr, "FAIL %s (%d): %s\n", what, v, qrdm_last_error());
    fails++;
  }
}
double* dev_doubles(size_t n) {  // the stub's device memory is host memory
  double* p = (double*)calloc(n ? n : 1, sizeof(double));
  return p;
}
int* dev_ints(size_t n) { return (int*)calloc(n ? n : 1, sizeof(int)); }

// one qrdm_dgeqrdm_dev call as the per-matrix loop makes it: jpvt and ncols
// host arrays of n entries (their only extent), stop mode in ncols[0]
int one_call(int m, int n, int lda, int nb, int stop, double* d_a, double* d_tau,
             const int* fixed = nullptr) {
  int* jp = new int[n];
  int* nc = new int[n];
  for (int c = 0; c < n; c++) jp[c] = fixed ? fixed[c] : 0;
  memset(nc, 0, sizeof(int) * n);
  nc[0] = stop;
  const double th[3] = {0.9, 0.15, 1e-10};
  const int info = qrdm_dgeqrdm_dev(102, m, n, d_a, lda, jp, d_tau, nc, th, nb, nullptr);
  // the DM steps cover the columns after the fixed ones
  int sum = 0, k = 0, nfxd = 0;
  for (int c = 0; c < n; c++) nfxd += (fixed && fixed[c] != 0) ? 1 : 0;
  const int mn = (m < n ? m : n) - nfxd;
  while (k < n && sum < mn && nc[k] > 0) sum += nc[k++];
  delete[] nc;
  delete[] jp;
  return info != 0 ? info : (sum == mn ? 0 : -99);
}
}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  FILE* f = fopen(argv[1], "r");
  if (!f) return 2;
  std::vector<qrdm_batch_item> items;
  long long a = 0, t = 0, c = 0;
  int m, n;
  while (fscanf(f, "%d %d", &m, &n) == 2) {
    qrdm_batch_item it{m, n, m, 0, a, t, c};
    items.push_back(it);
    a += (long long)m * n;
    t += m < n ? m : n;
    c += n;
  }
  fclose(f);
  const int cnt = (int)items.size();
  printf("batch of %d matrices\n", cnt);

  // ---- the batch: plan, shape requests, autotune, graph replays, eager run
  double* src = dev_doubles(a);
  double* work = dev_doubles(a);
  double* tau = dev_doubles(t);
  int* jp = dev_ints(c);
  int* ncb = dev_ints(c);
  int* info = dev_ints(cnt);
  qrdm_batch_plan* p = qrdm_batch_plan_create(cnt, items.data(), 64);
  check(p != nullptr, "plan", 0);
  check(qrdm_batch_plan_set_shape(p, 8) == -1, "8 waves refused", 0);
  check(qrdm_batch_plan_set_shape(p, 4) == 0, "4 waves", 0);
  const double th[3] = {0.9, 0.15, 0.0};
  check(qrdm_batch_plan_set_panel_rows(p, 128) == 0, "panel rows", 0);
  check(qrdm_batch_autotune(p, src, th, nullptr) == 0, "autotune", 0);
  for (int rep = 0; rep < 4; rep++) {
    memcpy(work, src, sizeof(double) * a);
    check(qrdm_dgeqrdm_batched(p, work, tau, jp, ncb, info, th, 0, 1, nullptr) == 0, "graph", rep);
  }
  const double th2[3] = {0.9, 0.15, 1e-9};
  check(qrdm_dgeqrdm_batched(p, work, tau, jp, ncb, info, th2, 3, 1, nullptr) == 0, "regraph", 0);
  check(qrdm_dgeqrdm_batched(p, work, tau, jp, ncb, info, th, 0, 0, nullptr) == 0, "eager", 0);
  for (int i = 0; i < cnt; i++) {
    int s = 0, k = 0;
    const int mn = items[i].m < items[i].n ? items[i].m : items[i].n;
    while (k < items[i].n && s < mn) s += ncb[items[i].col_off + k++];
    check(s == mn && info[i] == 0, "batch item", i);
  }
  int q[5];
  check(qrdm_batch_plan_query(p, q, 5) == 5 && q[0] == 4, "query", q[0]);

  // ---- the per-matrix loop over the same matrices
  for (int i = 0; i < cnt; i++) {
    const qrdm_batch_item& it = items[i];
    const int r = one_call(it.m, it.n, it.m, 64, i % 4, work + it.a_off, tau + it.tau_off);
    check(r == 0, "loop call", i);
  }
  qrdm_batch_plan_destroy(p);
  printf("batch and per-matrix loop done\n");

  // ---- the other host paths
  {
    const int M = 3000, N = 1700;
    double* d_a = dev_doubles((size_t)(M + 5) * N);
    double* d_tau = dev_doubles(M);  // min(m, n) of every call below
    std::vector<int> fixed(N, 0);
    fixed[3] = fixed[40] = fixed[41] = fixed[900] = 1;
    check(one_call(M, N, M + 5, 64, 0, d_a, d_tau, fixed.data()) == 0, "fixed", 0);
    for (int nb : {65, 128, 200, 256})
      check(one_call(M, N, M, nb, nb & 3, d_a, d_tau) == 0, "wide", nb);
    check(one_call(N, M, N, 64, 0, d_a, d_tau) == 0, "wide matrix", 0);
    check(one_call(90, 40, 90, 64, 0, d_a, d_tau) == 0, "short panels", 0);
    // traced and option calls: trace arrays of min(m, n) and fewer records
    for (int trail = 0; trail <= 4; trail++) {
      for (int cap : {N, 7}) {
        std::vector<qrdm_step> tr(cap);
        int* jpv = new int[N]();
        int* nc = new int[N]();
        int ns = 0;
        qrdm_opts o{trail, nullptr};
        const double th3[3] = {0.9, 0.15, 0.0};
        const int r = qrdm_dgeqrdm_dev_opt(102, M, N, d_a, M, jpv, d_tau, nc, th3, 64, &o,
                                           nullptr, tr.data(), cap, &ns);
        check(r == 0 && ns > 0, "opt", trail);
        delete[] nc;
        delete[] jpv;
      }
    }
    // the allocation-free entry point, twice, the second smaller
    for (int k = 0; k < 2; k++) {
      const int mm = k ? 700 : 2048, nn = k ? 500 : 2048;
      int* jpv = new int[nn]();
      int* nc = new int[nn]();
      const double th3[3] = {0.9, 0.15, 0.0};
      check(qrdm_dgeqrdm_dev_noalloc(102, mm, nn, d_a, mm + 3, jpv, d_tau, nc, th3, 64,
                                     nullptr) == 0, "noalloc", k);
      delete[] nc;
      delete[] jpv;
    }
    // the host-pointer entry point
    {
      const int mm = 500, nn = 300;
      double* ha = new double[(size_t)mm * nn]();
      double* ht = new double[nn];
      int* jpv = new int[nn]();
      int* nc = new int[nn]();
      double th3[3] = {0.9, 0.15, 0.0};
      check(dgeqrdm(102, mm, nn, ha, mm, jpv, ht, nc, th3, 64) == 0, "host dgeqrdm", 0);
      delete[] nc;
      delete[] jpv;
      delete[] ht;
      delete[] ha;
    }
    free(d_tau);
    free(d_a);
  }
  free(info);
  free(ncb);
  free(jp);
  free(tau);
  free(work);
  free(src);
  printf("%s (%d failures)\n", fails ? "FAILED" : "clean", fails);
  return fails ? 1 : 0;
}
