/*
 * C restatement of the reference algorithm for McClean.grad_run and
 * Qaoa.grad_run -- TEST INFRASTRUCTURE ONLY (the checker and the timed CPU
 * baseline; the product library never links it).
 *
 * It follows the reference's operation sequence, including its costs:
 *   McClean (qradient/circuit_logic/mc_clean.py:47-78): RY(pi/4) on every
 *     qubit; per layer the CNOT ladder (odd bonds first, state.py:229-241),
 *     a stored copy of the state (state_history), n rotations; lambda = O psi;
 *     E = Re<psi_L|lambda>; per layer backwards: n inverse rotations, a copy
 *     into tmp, then per qubit rot(+a), drot(-a), grad = -2 Re<hist_i|v>
 *     and a restore from tmp (mc_clean.py:69-76), then the inverse ladder.
 *   QAOA (qradient/circuit_logic/qaoa.py:40-70): |+>, per layer history,
 *     exp(-i g C), history, RX(b) on all qubits; lambda = C psi; backwards
 *     the summed-X and -iC generators with tmp copies.
 * Gate formulas: state.py:90-97 (x), 142-149 (y), 168-175 (z); qubit q is
 * index bit n-1-q (state.py:84-87).  Single-threaded, complex128.
 *
 * Built with -DORACLE_REAL='long double' it runs the same algorithm in x87
 * extended precision (64-bit mantissa), which pins the FP64 results: at the
 * benchmark's depth two FP64 operation orders of the same algorithm differ by
 * about 2e-12, and the extended-precision value tells which is closer.
 * Inputs are the same doubles either way; outputs are rounded to double.
 */
#include <complex.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>

#ifndef ORACLE_REAL
#define ORACLE_REAL double
#endif
typedef ORACLE_REAL real_t;
typedef ORACLE_REAL _Complex cd; /* a typedef name cannot take _Complex */

static void rot(cd* v, int n, int q, int axis, double a) {
  const uint64_t N = 1ull << n, bit = 1ull << (n - 1 - q);
  const real_t h = (real_t)a / 2;
  const real_t c = cos(h), s = sin(h);
  if (axis == 2) {
    const cd p0 = cexp(-I * h), p1 = cexp(I * h);
    for (uint64_t x = 0; x < N; ++x) v[x] *= (x & bit) ? p1 : p0;
    return;
  }
  for (uint64_t x = 0; x < N; ++x) {
    if (x & bit) continue;
    const cd a0 = v[x], a1 = v[x | bit];
    if (axis == 0) { /* sin(a/2)(-iX)v + cos(a/2)v */
      v[x] = c * a0 - I * s * a1;
      v[x | bit] = c * a1 - I * s * a0;
    } else { /* -iY = [[0,-1],[1,0]] */
      v[x] = c * a0 - s * a1;
      v[x | bit] = c * a1 + s * a0;
    }
  }
}

static void drot(cd* v, int n, int q, int axis, double a) {
  const uint64_t N = 1ull << n, bit = 1ull << (n - 1 - q);
  const real_t h = (real_t)a / 2;
  const real_t c = cos(h) / 2, s = sin(h) / 2;
  if (axis == 2) {
    const cd p0 = (-I / 2) * cexp(-I * h), p1 = (I / 2) * cexp(I * h);
    for (uint64_t x = 0; x < N; ++x) v[x] *= (x & bit) ? p1 : p0;
    return;
  }
  for (uint64_t x = 0; x < N; ++x) {
    if (x & bit) continue;
    const cd a0 = v[x], a1 = v[x | bit];
    if (axis == 0) {
      v[x] = c * (-I * a1) - s * a0;
      v[x | bit] = c * (-I * a0) - s * a1;
    } else {
      v[x] = c * (-a1) - s * a0;
      v[x | bit] = c * a0 - s * a1;
    }
  }
}

static void cnot(cd* v, int n, int ctl, int tgt) {
  const uint64_t N = 1ull << n, bc = 1ull << (n - 1 - ctl), bt = 1ull << (n - 1 - tgt);
  for (uint64_t x = 0; x < N; ++x) {
    if ((x & bc) && !(x & bt)) {
      const cd t = v[x];
      v[x] = v[x | bt];
      v[x | bt] = t;
    }
  }
}

/* stacking 0: odd bonds act first (state.py:235-239); 1: even first (237-241) */
static void ladder(cd* v, int n, int stacking) {
  for (int grp = 0; grp < 2; ++grp) {
    const int first = stacking == 0 ? (grp == 0 ? 1 : 0) : (grp == 0 ? 0 : 1);
    for (int i = first; i + 1 < n; i += 2) cnot(v, n, i, i + 1);
  }
}

/* kinds: 0 x, 1 y, 2 z, 3 zz (observable.py:44-73) */
static void apply_obs(const cd* v, cd* out, int n, int nt, const int* kind, const int* ii, const int* jj,
                      const double* w) {
  const uint64_t N = 1ull << n;
  for (uint64_t x = 0; x < N; ++x) {
    real_t d = 0.0;
    for (int k = 0; k < nt; ++k) {
      const int zi = (x >> (n - 1 - ii[k])) & 1 ? -1 : 1;
      if (kind[k] == 2) d += (real_t)w[k] * zi;
      if (kind[k] == 3) d += (real_t)w[k] * zi * (((x >> (n - 1 - jj[k])) & 1) ? -1 : 1);
    }
    out[x] = d * v[x];
  }
  for (int k = 0; k < nt; ++k) {
    if (kind[k] > 1) continue;
    const uint64_t bit = 1ull << (n - 1 - ii[k]);
    for (uint64_t x = 0; x < N; ++x) {
      const cd p = v[x ^ bit];
      if (kind[k] == 0) out[x] += (real_t)w[k] * p;
      else out[x] += (real_t)w[k] * ((x & bit) ? I * p : -I * p);
    }
  }
}

static real_t re_dot(const cd* a, const cd* b, uint64_t N) {
  real_t s = 0.0;
  for (uint64_t x = 0; x < N; ++x) s += creal(conj(a[x]) * b[x]);
  return s;
}

/* Runs the first Ls layers' forward and backward sweep; Ls = L for the full
 * grad_run.  grad receives Ls*n values (layer-major).  Returns 0 or -1 (OOM). */
int oracle_mcclean_grad_run(int n, int L, int Ls, const int* axes, const double* angles, int nt, const int* kind,
                            const int* ii, const int* jj, const double* w, double* E, double* grad) {
  (void)L;
  const uint64_t N = 1ull << n;
  cd* hist = (cd*)malloc(sizeof(cd) * N * (size_t)(Ls + 1));
  cd* v = (cd*)malloc(sizeof(cd) * N);
  cd* lam = (cd*)malloc(sizeof(cd) * N);
  cd* tmp = (cd*)malloc(sizeof(cd) * N);
  if (!hist || !v || !lam || !tmp) {
    free(hist); free(v); free(lam); free(tmp);
    return -1;
  }
  memset(v, 0, sizeof(cd) * N);
  v[0] = 1.0;
  for (int q = 0; q < n; ++q) rot(v, n, q, 1, M_PI / 4);
  for (int i = 0; i < Ls; ++i) {
    ladder(v, n, 0);
    memcpy(hist + (size_t)i * N, v, sizeof(cd) * N);
    for (int q = 0; q < n; ++q) rot(v, n, q, axes[i * n + q], angles[i * n + q]);
  }
  memcpy(hist + (size_t)Ls * N, v, sizeof(cd) * N);
  apply_obs(v, lam, n, nt, kind, ii, jj, w);
  *E = (double)re_dot(hist + (size_t)Ls * N, lam, N);
  for (int i = Ls - 1; i >= 0; --i) {
    for (int q = 0; q < n; ++q) rot(lam, n, q, axes[i * n + q], -angles[i * n + q]);
    memcpy(tmp, lam, sizeof(cd) * N);
    for (int q = 0; q < n; ++q) {
      rot(lam, n, q, axes[i * n + q], angles[i * n + q]);
      drot(lam, n, q, axes[i * n + q], -angles[i * n + q]);
      grad[i * n + q] = (double)(-2 * re_dot(hist + (size_t)i * N, lam, N));
      memcpy(lam, tmp, sizeof(cd) * N);
    }
    ladder(lam, n, 1);
  }
  free(hist); free(v); free(lam); free(tmp);
  return 0;
}

static void diag_C(real_t* C, int n, int nt, const int* kind, const int* ii, const int* jj, const double* w) {
  const uint64_t N = 1ull << n;
  for (uint64_t x = 0; x < N; ++x) {
    real_t d = 0.0;
    for (int k = 0; k < nt; ++k) {
      const int zi = (x >> (n - 1 - ii[k])) & 1 ? -1 : 1;
      if (kind[k] == 2) d += (real_t)w[k] * zi;
      if (kind[k] == 3) d += (real_t)w[k] * zi * (((x >> (n - 1 - jj[k])) & 1) ? -1 : 1);
    }
    C[x] = d;
  }
}

/* Qaoa.grad_run (qaoa.py:40-70); grad[2*i] = dE/dbeta_i, grad[2*i+1] = dE/dgamma_i. */
int oracle_qaoa_grad_run(int n, int p, const double* betas, const double* gammas, int nt, const int* kind,
                         const int* ii, const int* jj, const double* w, double* E, double* grad) {
  const uint64_t N = 1ull << n;
  cd* hist = (cd*)malloc(sizeof(cd) * N * (size_t)(2 * p + 1));
  cd* v = (cd*)malloc(sizeof(cd) * N);
  cd* tmp = (cd*)malloc(sizeof(cd) * N);
  cd* u = (cd*)malloc(sizeof(cd) * N);
  real_t* C = (real_t*)malloc(sizeof(real_t) * N);
  if (!hist || !v || !tmp || !u || !C) {
    free(hist); free(v); free(tmp); free(u); free(C);
    return -1;
  }
  diag_C(C, n, nt, kind, ii, jj, w);
  for (uint64_t x = 0; x < N; ++x) v[x] = pow((real_t)2, -(real_t)n / 2);
  for (int i = 0; i < p; ++i) {
    memcpy(hist + (size_t)(2 * i) * N, v, sizeof(cd) * N);
    for (uint64_t x = 0; x < N; ++x) v[x] *= cexp(-I * ((real_t)gammas[i] * C[x]));
    memcpy(hist + (size_t)(2 * i + 1) * N, v, sizeof(cd) * N);
    for (int q = 0; q < n; ++q) rot(v, n, q, 0, betas[i]);
  }
  memcpy(hist + (size_t)(2 * p) * N, v, sizeof(cd) * N);
  for (uint64_t x = 0; x < N; ++x) v[x] *= C[x];
  *E = (double)re_dot(hist + (size_t)(2 * p) * N, v, N);
  for (int i = p - 1; i >= 0; --i) {
    for (int q = 0; q < n; ++q) rot(v, n, q, 0, -betas[i]);
    memcpy(tmp, v, sizeof(cd) * N);
    /* x_summed: .5 * sum_q (-i X_q) v (state.py:107-123) */
    memset(u, 0, sizeof(cd) * N);
    for (int q = 0; q < n; ++q) {
      const uint64_t bit = 1ull << (n - 1 - q);
      for (uint64_t x = 0; x < N; ++x) u[x] += (-I / 2) * v[x ^ bit];
    }
    grad[2 * i] = (double)(-2 * re_dot(hist + (size_t)(2 * i + 1) * N, u, N));
    memcpy(v, tmp, sizeof(cd) * N);
    for (uint64_t x = 0; x < N; ++x) v[x] *= cexp(I * ((real_t)gammas[i] * C[x]));
    memcpy(tmp, v, sizeof(cd) * N);
    for (uint64_t x = 0; x < N; ++x) u[x] = v[x] * (-I * C[x]);
    grad[2 * i + 1] = (double)(-2 * re_dot(hist + (size_t)(2 * i) * N, u, N));
    memcpy(v, tmp, sizeof(cd) * N);
  }
  free(hist); free(v); free(tmp); free(u); free(C);
  return 0;
}
