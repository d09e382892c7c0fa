/* A plain C caller of the library's C ABI (include/qradient_amd.h): what a
 * binding of the reference's McClean.grad_run (qradient/circuit_logic/
 * mc_clean.py:47-78) from a language other than Python would do.
 *
 *   c_caller <circuit file> <dtype: 0 complex128, 1 complex64>
 *
 * The circuit file holds "n L", then L*n axes (0 X, 1 Y, 2 Z), L*n angles
 * and n-1 weights w_i of the observable sum_i w_i Z_i Z_{i+1}, whitespace
 * separated.  The program runs the circuit's grad run twice: directly
 * (qg_ctx_create, qg_mcclean_plan, qg_mcclean_grad_run) and through the
 * serving loop (qg_server_create, qg_submit, qg_wait_any, qg_poll), and
 * prints both as lines "direct E g[0] ... g[L*n-1]" and "served E g..." with
 * 17 significant digits.  Exit status 0, or 1 with qg_last_error() on stderr. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "qradient_amd.h"

#define CHECK(call)                                                        \
  do {                                                                     \
    int rc_ = (call);                                                      \
    if (rc_ != QG_OK) {                                                    \
      fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, qg_last_error()); \
      return 1;                                                            \
    }                                                                      \
  } while (0)

static void print_run(const char* tag, double E, const double* g, int count) {
  printf("%s %.17g", tag, E);
  for (int k = 0; k < count; ++k) printf(" %.17g", g[k]);
  printf("\n");
}

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s <circuit file> <dtype 0|1>\n", argv[0]);
    return 1;
  }
  const int dtype = atoi(argv[2]);
  FILE* f = fopen(argv[1], "r");
  if (!f) {
    fprintf(stderr, "cannot open %s\n", argv[1]);
    return 1;
  }
  int n = 0, L = 0;
  if (fscanf(f, "%d %d", &n, &L) != 2 || n < 2 || L < 1) {
    fprintf(stderr, "bad header\n");
    return 1;
  }
  const int np = n * L;
  int32_t* axes = malloc(sizeof(int32_t) * np);
  double* angles = malloc(sizeof(double) * np);
  qg_term* terms = malloc(sizeof(qg_term) * (n - 1));
  double* grad = malloc(sizeof(double) * np);
  double* grad_served = malloc(sizeof(double) * np);
  if (!axes || !angles || !terms || !grad || !grad_served) return 1;
  for (int k = 0; k < np; ++k)
    if (fscanf(f, "%d", &axes[k]) != 1) return 1;
  for (int k = 0; k < np; ++k)
    if (fscanf(f, "%lf", &angles[k]) != 1) return 1;
  for (int i = 0; i < n - 1; ++i) {
    terms[i].kind = QG_TERM_ZZ;
    terms[i].i = i;
    terms[i].j = i + 1;
    terms[i].pad = 0;
    if (fscanf(f, "%lf", &terms[i].w) != 1) return 1;
  }
  fclose(f);

  /* direct: one context, one plan for the shape, one grad run */
  qg_ctx* ctx = NULL;
  qg_plan* plan = NULL;
  double E = 0.0;
  CHECK(qg_ctx_create(0, NULL, &ctx));
  CHECK(qg_mcclean_plan(ctx, n, L, dtype, QG_ENT_CNOT, 0, &plan));
  CHECK(qg_mcclean_grad_run(plan, axes, angles, terms, n - 1, NULL, &E, grad));
  print_run("direct", E, grad, np);

  /* served: a two-slot server with a 1 GiB arena, shortest-first admission */
  qg_server* srv = NULL;
  int64_t ticket = 0, done = 0;
  double E_served = 0.0;
  qg_request_info info;
  CHECK(qg_server_create(0, 2, dtype, (size_t)1 << 30, (size_t)1 << 30, 4, 8, QG_SERVER_SHORTEST_FIRST, &srv));
  CHECK(qg_submit(srv, n, L, axes, angles, terms, n - 1, &ticket));
  int rc = QG_ERR_PENDING;
  for (int tries = 0; tries < 120 && rc == QG_ERR_PENDING; ++tries) rc = qg_wait_any(srv, 1000, &done);
  CHECK(rc);
  if (done != ticket) {
    fprintf(stderr, "qg_wait_any returned ticket %lld, expected %lld\n", (long long)done, (long long)ticket);
    return 1;
  }
  CHECK(qg_poll(srv, ticket, 0, &E_served, grad_served, &info));
  print_run("served", E_served, grad_served, np);
  printf("info queue_ms %.3f run_ms %.3f ckpt_every %d\n", info.queue_ms, info.run_ms, (int)info.ckpt_every);

  CHECK(qg_server_destroy(srv));
  CHECK(qg_plan_destroy(plan));
  CHECK(qg_ctx_destroy(ctx));
  free(axes);
  free(angles);
  free(terms);
  free(grad);
  free(grad_served);
  return 0;
}
