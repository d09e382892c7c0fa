// Elementwise HIP kernels behind the device-resident State API and the
// generic (non-diagonal) observable path.  One thread per amplitude or pair;
// these are not the grad_run hot loop (qg_kernels.hip is), but they keep every
// State method of the reference on the GPU.
#include <hip/hip_runtime.h>
#include "qg_internal.h"

namespace qg {

template <class S>
struct V2;
template <>
struct V2<float> {
  typedef float2 T;
};
template <>
struct V2<double> {
  typedef double2 T;
};

// x0' = c x0 - i s (P x)_0 ; x1' = c x1 - i s (P x)_1 on global bit b
// (every 2x2 of the reference's rotations and their derivatives has this form).
template <class S>
__global__ void k_rot1(typename V2<S>::T* v, uint32_t npairs, int b, int type, double c, double s) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  const uint32_t lo = i & ((1u << b) - 1u);
  const uint32_t i0 = ((i >> b) << (b + 1)) | lo;
  const uint32_t i1 = i0 | (1u << b);
  const double ar = v[i0].x, ai = v[i0].y, br = v[i1].x, bi = v[i1].y;
  double r0, j0, r1, j1;
  if (type == G_X) {
    r0 = c * ar + s * bi; j0 = c * ai - s * br;
    r1 = c * br + s * ai; j1 = c * bi - s * ar;
  } else if (type == G_Y) {
    r0 = c * ar - s * br; j0 = c * ai - s * bi;
    r1 = c * br + s * ar; j1 = c * bi + s * ai;
  } else {
    r0 = c * ar + s * ai; j0 = c * ai - s * ar;
    r1 = c * br - s * bi; j1 = c * bi + s * br;
  }
  v[i0].x = (S)r0; v[i0].y = (S)j0;
  v[i1].x = (S)r1; v[i1].y = (S)j1;
}

// out[y] = in[map(y)] for a GF(2)-linear index map given by its columns.
template <class S>
__global__ void k_permute(const typename V2<S>::T* __restrict__ in, typename V2<S>::T* __restrict__ out,
                          uint32_t N, const uint32_t* __restrict__ map, int n) {
  const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= N) return;
  uint32_t x = 0;
  for (int k = 0; k < n; ++k) x ^= map[k] & (0u - ((y >> k) & 1u));
  out[y] = in[x];
}

__device__ __forceinline__ double diag_at(const DTerm* __restrict__ t, int nt, uint32_t x) {
  double D = 0.0;
  for (int k = 0; k < nt; ++k) {
    double z = ((x >> t[k].a) & 1u) ? -1.0 : 1.0;
    if (t[k].b >= 0) z *= ((x >> t[k].b) & 1u) ? -1.0 : 1.0;
    D += t[k].w * z;
  }
  return D;
}

// v *= exp(-i * angle * D(x))   (and CZ-ladder sign when cz != 0 and nt == 0)
template <class S>
__global__ void k_diag_phase(typename V2<S>::T* v, uint32_t N, const DTerm* __restrict__ t, int nt, double angle,
                             int cz, int n) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  double re = v[x].x, im = v[x].y;
  if (cz) {
    const uint32_t chain = (n >= 2) ? (uint32_t)((1ull << (n - 1)) - 1ull) : 0u;
    if (__popc(x & (x >> 1) & chain) & 1) { re = -re; im = -im; }
  } else {
    double sn, cs;
    sincos(angle * diag_at(t, nt, x), &sn, &cs);
    const double r = cs * re + sn * im, i = cs * im - sn * re;
    re = r; im = i;
  }
  v[x].x = (S)re;
  v[x].y = (S)im;
}

// out = O in, O = D(x) + sum_k w_k P_k (P in {X, Y} on global bit b_k).
template <class S>
__global__ void k_obs_apply(const typename V2<S>::T* __restrict__ in, typename V2<S>::T* __restrict__ out,
                            uint32_t N, const DTerm* __restrict__ dt, int ndt, const DTerm* __restrict__ xy, int nxy) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  const double D = diag_at(dt, ndt, x);
  double re = D * (double)in[x].x, im = D * (double)in[x].y;
  for (int k = 0; k < nxy; ++k) {
    const int b = xy[k].a;
    const double w = xy[k].w;
    const uint32_t y = x ^ (1u << b);
    const double pr = in[y].x, pi = in[y].y;
    if (xy[k].b == 0) {  // X
      re += w * pr; im += w * pi;
    } else {             // Y: (Y v)_x = -i v_y if bit 0, +i v_y if bit 1
      const double sg = ((x >> b) & 1u) ? 1.0 : -1.0;
      re += w * (-sg * pi); im += w * (sg * pr);
    }
  }
  out[x].x = (S)re;
  out[x].y = (S)im;
}

// part[block] = sum over this block's elements of Re(conj(a) b) (mode 0),
// |a|^2 (mode 1), in FP64, fixed-order tree.
template <class S>
__global__ void k_dot(const typename V2<S>::T* __restrict__ a, const typename V2<S>::T* __restrict__ b, uint32_t N,
                      int mode, double* __restrict__ part) {
  __shared__ double red[256];
  double acc = 0.0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const double ar = a[i].x, ai = a[i].y;
    if (mode == 0) acc += ar * (double)b[i].x + ai * (double)b[i].y;
    else acc += ar * ar + ai * ai;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// part[k * gridDim.x + block] = this block's share of <P_k> for the terms
// t[0, nt), nt <= kTermsPerLaunch, in FP64:
//   Z / ZZ: |v_x|^2 z(x);  X: Re(conj(v_x) v_{x^m});  Y: Re(conj(v_x) (Y v)_x);
//   identity (kind 3): |v_x|^2
// with (Y v)_x = -i v_{x^m} on bit 0 and +i v_{x^m} on bit 1.  Fixed order:
// grid-stride sums, wave butterflies, then the block's waves in turn.
template <class S>
__global__ void __launch_bounds__(256) k_term_expec(const typename V2<S>::T* __restrict__ v, uint32_t N,
                                                    const PTerm* __restrict__ t, int nt, double* __restrict__ part) {
  __shared__ double red[4][kTermsPerLaunch];
  double acc[kTermsPerLaunch];
#pragma unroll
  for (int k = 0; k < kTermsPerLaunch; ++k) acc[k] = 0.0;
  for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < N; x += gridDim.x * blockDim.x) {
    const double ar = v[x].x, ai = v[x].y;
    const double pr = ar * ar + ai * ai;
#pragma unroll
    for (int k = 0; k < kTermsPerLaunch; ++k) {
      if (k < nt) {
        const PTerm tk = t[k];
        if (tk.kind == 3) {
          acc[k] += pr;
        } else if (tk.kind == 0) {
          uint32_t par = (x >> tk.a) & 1u;
          if (tk.b >= 0) par ^= (x >> tk.b) & 1u;
          acc[k] += par ? -pr : pr;
        } else {
          const uint32_t bit = (x >> tk.a) & 1u;
          const uint32_t y = x ^ (1u << tk.a);
          const double cr = v[y].x, ci = v[y].y;
          if (tk.kind == 1) {
            acc[k] += ar * cr + ai * ci;
          } else {
            const double im = ar * ci - ai * cr;  // Im(conj(v_x) v_{x^m})
            acc[k] += bit ? -im : im;
          }
        }
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kTermsPerLaunch; ++k) {
    double s = acc[k];
    for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m, 64);
    if (lane == 0) red[wave][k] = s;
  }
  __syncthreads();
  if ((int)threadIdx.x < nt) {
    const int k = threadIdx.x;
    part[(size_t)k * gridDim.x + blockIdx.x] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
  }
}

// The same sums for diagonal terms only (Z / ZZ, and the identity, kind 3):
// <P_k> sums |v_x|^2 (-1)^popc(x & m_k).  The grid stride is a power of two
// (2^sb amplitudes), so x = g + j 2^sb with g the thread's global index and j
// the iteration: the sign splits into (-1)^popc(g & m_k), fixed per thread,
// and (-1)^popc(j & (m_k >> sb)), the same for the whole grid at iteration j.
// Each thread sums (-1)^popc(j & mh_k) |v_x|^2 over its iterations (one FMA
// per term and amplitude) and applies its own sign at the end.
// Up to kDiagTermsPerLaunch terms per launch: the per-thread sign of a term
// is recomputed from its mask after the loop, so a term costs one FP64
// accumulator in the loop.
template <class S>
__global__ void __launch_bounds__(256) k_term_expec_diag(const typename V2<S>::T* __restrict__ v, uint32_t N,
                                                         const PTerm* __restrict__ t, int nt, int sb,
                                                         double* __restrict__ part) {
  constexpr int NT = kDiagTermsPerLaunch;
  __shared__ double red[4][NT];
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  auto mask_of = [&](int k) -> uint32_t {
    return (k < nt && t[k].kind == 0) ? (1u << t[k].a) | (t[k].b >= 0 ? (1u << t[k].b) : 0u) : 0u;
  };
  uint32_t mh[NT];
  double acc[NT];
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    const uint32_t m = mask_of(k);
    mh[k] = sb < 32 ? m >> sb : 0u;
    acc[k] = 0.0;
  }
  const uint32_t iters = N >> sb;  // N and the stride are powers of two
  for (uint32_t j = 0; j < iters; j += 2) {
    const typename V2<S>::T a0 = v[g + (j << sb)];
    const double p0 = (double)a0.x * a0.x + (double)a0.y * a0.y;
    double p1 = 0.0;
    if (j + 1 < iters) {
      const typename V2<S>::T a1 = v[g + ((j + 1) << sb)];
      p1 = (double)a1.x * a1.x + (double)a1.y * a1.y;
    }
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      const double s0 = (__popc(j & mh[k]) & 1u) ? -1.0 : 1.0;
      const double s1 = (__popc((j + 1) & mh[k]) & 1u) ? -1.0 : 1.0;
      acc[k] = __builtin_fma(s1, p1, __builtin_fma(s0, p0, acc[k]));
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NT; ++k) {
    const double sl = (__popc(g & mask_of(k)) & 1u) ? -1.0 : 1.0;
    double s = sl * acc[k];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) red[wave][k] = s;
  }
  __syncthreads();
  if ((int)threadIdx.x < nt) {
    const int k = threadIdx.x;
    part[(size_t)k * gridDim.x + blockIdx.x] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
  }
}

// out[k] = sum of part[k * np + 0 .. np): one 256-thread workgroup per term,
// coalesced strided partial sums, then a fixed-order tree.
__global__ void __launch_bounds__(256) k_sum_terms(const double* __restrict__ part, int np, int nt,
                                                   double* __restrict__ out) {
  __shared__ double red[256];
  const int k = blockIdx.x;
  if (k >= nt) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) s += part[(size_t)k * np + i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = red[0];
}

// *out = sum of part[0, np): one 256-thread workgroup, fixed-order tree.
__global__ void __launch_bounds__(256) k_sum(const double* __restrict__ part, int np, double* __restrict__ out) {
  __shared__ double red[256];
  if (blockIdx.x != 0) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// amp(x) = tab[popcount(x)]  (product initial states)
template <class S>
__global__ void k_init(typename V2<S>::T* v, uint32_t N, const double* __restrict__ tab) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  v[x].x = (S)tab[__popc(x)];
  v[x].y = (S)0.0;
}

template <class S>
__global__ void k_convert_in(const double2* __restrict__ in, typename V2<S>::T* __restrict__ out, uint32_t N) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  out[x].x = (S)in[x].x;
  out[x].y = (S)in[x].y;
}

template <class S>
__global__ void k_convert_out(const typename V2<S>::T* __restrict__ in, double2* __restrict__ out, uint32_t N) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  out[x].x = (double)in[x].x;
  out[x].y = (double)in[x].y;
}

template <class S>
__global__ void k_scale(typename V2<S>::T* v, uint32_t N, double re, double im) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= N) return;
  const double a = v[x].x, b = v[x].y;
  v[x].x = (S)(re * a - im * b);
  v[x].y = (S)(re * b + im * a);
}

// ---------------------------------------------------------------------------
// The elementwise kernels count amplitudes in 32 bits: states of up to 2^31.
constexpr int kSimpleMaxN = 31;

// ---------------------------------------------------------------------------
// Basis sampling: the inverse CDF of the distribution |v_x|^2 over positions
// k = 0 .. N-1, position k holding amplitude order[k] (or k without an
// order).  This is scipy's rv_discrete(values=(arange(N), dist)).rvs as the
// reference's dense sampled gradients call it (mc_clean.py:246-254,
// qaoa.py:196-198): a draw u maps to the first position whose cumulative
// probability reaches u, and to position 0 when none does
// (argmax(qvals >= u)).  Two levels: masses of 1024-position blocks and
// their prefix, then a binary search over blocks and a walk inside one.
// ---------------------------------------------------------------------------
template <class S>
__device__ __forceinline__ double prob_at(const typename V2<S>::T* __restrict__ v, const int32_t* __restrict__ order,
                                          uint32_t k) {
  const uint32_t x = order ? (uint32_t)order[k] : k;
  const double re = v[x].x, im = v[x].y;
  return re * re + im * im;
}

template <class S>
__global__ void __launch_bounds__(256) k_block_mass(const typename V2<S>::T* __restrict__ v, uint32_t N,
                                                    const int32_t* __restrict__ order, double* __restrict__ mass) {
  __shared__ double red[256];
  const uint32_t base = blockIdx.x * (uint32_t)kSampleBlock + threadIdx.x * 4u;
  double s = 0.0;
  for (uint32_t j = 0; j < 4; ++j)
    if (base + j < N) s += prob_at<S>(v, order, base + j);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) mass[blockIdx.x] = red[0];
}

// Inclusive prefix sums of mass[0, nb) in place, one workgroup of 256 threads.
__global__ void __launch_bounds__(256) k_mass_prefix(double* __restrict__ mass, int nb) {
  __shared__ double tot[256];
  const int chunk = (nb + 255) / 256;
  const int lo = (int)threadIdx.x * chunk, hi = min(nb, lo + chunk);
  double s = 0.0;
  for (int b = lo; b < hi; ++b) s += mass[b];
  tot[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0;
    for (int t = 0; t < 256; ++t) {
      a += tot[t];
      tot[t] = a;
    }
  }
  __syncthreads();
  double a = threadIdx.x ? tot[threadIdx.x - 1] : 0.0;
  for (int b = lo; b < hi; ++b) {
    a += mass[b];
    mass[b] = a;
  }
}

template <class S>
__global__ void k_basis_search(const typename V2<S>::T* __restrict__ v, uint32_t N, const int32_t* __restrict__ order,
                               const double* __restrict__ pref, int nb, const double* __restrict__ u, int shots,
                               int32_t* __restrict__ idx) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= shots) return;
  const double us = u[s];
  if (!(pref[nb - 1] >= us)) {
    idx[s] = 0;
    return;
  }
  int lo = 0, hi = nb - 1;  // the first block whose prefix reaches us
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (pref[mid] >= us) hi = mid;
    else lo = mid + 1;
  }
  double acc = lo ? pref[lo - 1] : 0.0;
  const uint32_t k0 = (uint32_t)lo * (uint32_t)kSampleBlock;
  const uint32_t k1 = min(N, k0 + (uint32_t)kSampleBlock);
  uint32_t k = k0;
  for (; k < k1; ++k) {
    acc += prob_at<S>(v, order, k);
    if (acc >= us) break;
  }
  idx[s] = (int32_t)(k < k1 ? k : k1 - 1);  // rounding between the two levels: the block's last position
}

static inline dim3 grid_for(uint64_t n, int threads) { return dim3((unsigned)((n + threads - 1) / threads)); }

#define QG_DT_SWITCH(dtype, CALL) \
  do {                            \
    if ((dtype) == 1) {           \
      typedef float S;            \
      CALL;                       \
    } else {                      \
      typedef double S;           \
      CALL;                       \
    }                             \
  } while (0)

hipError_t simple_rot1(int dtype, void* v, int n, int b, int type, double c, double s, hipStream_t st) {
  if (n < 1 || n > kSimpleMaxN) return hipErrorInvalidValue;
  const uint32_t np = (uint32_t)(1ull << (n - 1));
  QG_DT_SWITCH(dtype, hipLaunchKernelGGL(k_rot1<S>, grid_for(np, 256), dim3(256), 0, st,
                                         (typename V2<S>::T*)v, np, b, type, c, s));
  return hipGetLastError();
}

hipError_t simple_permute(int dtype, const void* in, void* out, int n, const uint32_t* dmap, hipStream_t st) {
  if (n < 1 || n > kSimpleMaxN) return hipErrorInvalidValue;
  const uint32_t N = (uint32_t)(1ull << n);
  QG_DT_SWITCH(dtype, hipLaunchKernelGGL(k_permute<S>, grid_for(N, 256), dim3(256), 0, st,
                                         (const typename V2<S>::T*)in, (typename V2<S>::T*)out, N, dmap, n));
  return hipGetLastError();
}

hipError_t simple_diag_phase(int dtype, void* v, int n, const DTerm* dt, int nt, double angle, int cz,
                             hipStream_t st) {
  if (n < 1 || n > kSimpleMaxN) return hipErrorInvalidValue;
  const uint32_t N = (uint32_t)(1ull << n);
  QG_DT_SWITCH(dtype, hipLaunchKernelGGL(k_diag_phase<S>, grid_for(N, 256), dim3(256), 0, st,
                                         (typename V2<S>::T*)v, N, dt, nt, angle, cz, n));
  return hipGetLastError();
}

hipError_t simple_obs_apply(int dtype, const void* in, void* out, int n, const DTerm* dt, int ndt, const DTerm* xy,
                            int nxy, hipStream_t st) {
  if (n < 1 || n > kSimpleMaxN) return hipErrorInvalidValue;
  const uint32_t N = (uint32_t)(1ull << n);
  QG_DT_SWITCH(dtype, hipLaunchKernelGGL(k_obs_apply<S>, grid_for(N, 256), dim3(256), 0, st,
                                         (const typename V2<S>::T*)in, (typename V2<S>::T*)out, N, dt, ndt, xy, nxy));
  return hipGetLastError();
}

// out = sum: mode 0 Re<a|b>, mode 1 <a|a>; part needs `kDotBlocks` doubles.
constexpr int kDotBlocks = 512;
hipError_t simple_dot(int dtype, const void* a, const void* b, int n, int mode, double* part, double* out,
                      hipStream_t st) {
  if (n < 1 || n > kSimpleMaxN) return hipErrorInvalidValue;
  const uint32_t N = (uint32_t)(1ull << n);
  const int blocks = (int)((N + 255) / 256) < kDotBlocks ? (int)((N + 255) / 256) : kDotBlocks;
  QG_DT_SWITCH(dtype, hipLaunchKernelGGL(k_dot<S>, dim3(blocks), dim3(256), 0, st, (const typename V2<S>::T*)a,
                                         (const typename V2<S>::T*)b, N, mode, part));
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, st, (const double*)part, blocks, out);
  return hipGetLastError();
}

// out[k] = <P_k> for the nt terms t (device arrays); part needs
// kTermBlocks * kTermPartRows doubles.  h_t: a host copy of t, or NULL; with
// it, a run of up to kDiagTermsPerLaunch diagonal terms takes one launch of
// the mask kernel, other terms go kTermsPerLaunch at a time to the general
// kernel.  Each term's sum does not depend on the launch it is in.
hipError_t simple_term_expec(int dtype, const void* v, int n, const PTerm* t, int nt, double* part, double* out,
                             hipStream_t st, const PTerm* h_t) {
  if (n < 1 || n > kSimpleMaxN) return hipErrorInvalidValue;
  const uint32_t N = (uint32_t)(1ull << n);
  const int blocks = (int)((N + 255) / 256) < kTermBlocks ? (int)((N + 255) / 256) : kTermBlocks;
  auto is_diag = [&](int k) { return h_t[k].kind == 0 || h_t[k].kind == 3; };
  for (int k0 = 0, m = 0; k0 < nt; k0 += m) {
    // the diagonal kernel needs one thread per amplitude or fewer: N >= 256
    m = 0;
    if (h_t && N >= 256)
      while (m < kDiagTermsPerLaunch && k0 + m < nt && is_diag(k0 + m)) ++m;
    const bool diag = m > 0;
    if (!diag) m = nt - k0 < kTermsPerLaunch ? nt - k0 : kTermsPerLaunch;
    if (diag) {
      int sb = 0;
      while ((1u << sb) < (uint32_t)blocks * 256u) ++sb;
      QG_DT_SWITCH(dtype, hipLaunchKernelGGL(k_term_expec_diag<S>, dim3(blocks), dim3(256), 0, st,
                                             (const typename V2<S>::T*)v, N, t + k0, m, sb, part));
    } else {
      QG_DT_SWITCH(dtype, hipLaunchKernelGGL(k_term_expec<S>, dim3(blocks), dim3(256), 0, st,
                                             (const typename V2<S>::T*)v, N, t + k0, m, part));
    }
    hipLaunchKernelGGL(k_sum_terms, dim3(m), dim3(256), 0, st, (const double*)part, blocks, m, out + k0);
  }
  return hipGetLastError();
}

hipError_t simple_init(int dtype, void* v, int n, const double* tab, hipStream_t st) {
  if (n < 1 || n > kSimpleMaxN) return hipErrorInvalidValue;
  const uint32_t N = (uint32_t)(1ull << n);
  QG_DT_SWITCH(dtype, hipLaunchKernelGGL(k_init<S>, grid_for(N, 256), dim3(256), 0, st, (typename V2<S>::T*)v, N, tab));
  return hipGetLastError();
}

hipError_t simple_scale(int dtype, void* v, int n, double re, double im, hipStream_t st) {
  if (n < 1 || n > kSimpleMaxN) return hipErrorInvalidValue;
  const uint32_t N = (uint32_t)(1ull << n);
  QG_DT_SWITCH(dtype, hipLaunchKernelGGL(k_scale<S>, grid_for(N, 256), dim3(256), 0, st, (typename V2<S>::T*)v, N, re, im));
  return hipGetLastError();
}

hipError_t simple_convert_in(int dtype, const void* in_c128, void* out, int n, hipStream_t st) {
  if (n < 1 || n > kSimpleMaxN) return hipErrorInvalidValue;
  const uint32_t N = (uint32_t)(1ull << n);
  QG_DT_SWITCH(dtype, hipLaunchKernelGGL(k_convert_in<S>, grid_for(N, 256), dim3(256), 0, st, (const double2*)in_c128,
                                         (typename V2<S>::T*)out, N));
  return hipGetLastError();
}

hipError_t simple_convert_out(int dtype, const void* in, void* out_c128, int n, hipStream_t st) {
  if (n < 1 || n > kSimpleMaxN) return hipErrorInvalidValue;
  const uint32_t N = (uint32_t)(1ull << n);
  QG_DT_SWITCH(dtype, hipLaunchKernelGGL(k_convert_out<S>, grid_for(N, 256), dim3(256), 0, st,
                                         (const typename V2<S>::T*)in, (double2*)out_c128, N));
  return hipGetLastError();
}

// mass: (2^n + kSampleBlock - 1) / kSampleBlock doubles of scratch.
hipError_t simple_sample_basis(int dtype, const void* v, int n, const int32_t* order, const double* u, int shots,
                               int32_t* idx, double* mass, hipStream_t st) {
  if (n < 1 || n > kSimpleMaxN || shots < 1) return hipErrorInvalidValue;
  const uint32_t N = (uint32_t)(1ull << n);
  const int nb = (int)((N + kSampleBlock - 1) / kSampleBlock);
  QG_DT_SWITCH(dtype, hipLaunchKernelGGL(k_block_mass<S>, dim3(nb), dim3(256), 0, st, (const typename V2<S>::T*)v, N,
                                         order, mass));
  hipLaunchKernelGGL(k_mass_prefix, dim3(1), dim3(256), 0, st, mass, nb);
  QG_DT_SWITCH(dtype, hipLaunchKernelGGL(k_basis_search<S>, grid_for((uint64_t)shots, 256), dim3(256), 0, st,
                                         (const typename V2<S>::T*)v, N, order, (const double*)mass, nb, u, shots,
                                         idx));
  return hipGetLastError();
}

}  // namespace qg
