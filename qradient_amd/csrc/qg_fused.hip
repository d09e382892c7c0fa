// Whole-circuit grad_run for states that fit one workgroup (n <= 13).
//
// One workgroup runs one circuit end to end: the 2^n amplitudes stay in
// registers (2^R per thread) and LDS for every layer of the forward sweep,
// the observable and the adjoint sweep, so nothing but the parameters and the
// results crosses HBM.  psi is uncomputed alongside lambda in FP64 (no
// history: unitary FP64 uncompute drifts by ~1e-15 per layer), so complex64
// and complex128 requests give the same FP64-quality result here.  A batch of
// circuits is one launch with one workgroup per circuit.
//
// McClean (reference qradient/circuit_logic/mc_clean.py:47-78): RY(pi/4)
// initial layer (table by popcount), per layer the CNOT ladder as an LDS
// permutation (or the CZ sign) then one rotation per qubit.  QAOA
// (qaoa.py:40-70): |+>, per layer exp(-i gamma C) then RX(beta) on all qubits.
#include "qg_pass.hpp"

namespace qg {

struct FusedDesc {
  int32_t n, L, kind;     // kind 0 McClean, 1 QAOA
  int32_t ent;            // McClean: 0 CNOT ladder, 1 CZ ladder
  const int32_t* gtype;   // [L][n] gate type per layer and qubit (McClean)
  const double* gcs;      // [L][n][2] cos(a/2), sin(a/2)          (McClean)
  const double* params;   // QAOA: [L][2] (beta, gamma)
  const DTerm* dterms;    // diagonal observable / classical Hamiltonian
  int32_t ndterms;
  const DTerm* xy;        // x / y terms: a = bit, b = 0 (X) or 1 (Y)
  int32_t nxy;
  uint32_t lad1, lad2;    // CNOT ladder gather map in shift form (ladder_of)
  const double* init_tab; // n + 1 entries
  double* out;            // [1 + L*n] (McClean) or [1 + 2L] (QAOA)
  int32_t grad;           // 0: forward and E only
  int32_t psi_c64;        // psi_out element type: 0 complex128, 1 complex64
  void* psi_out;          // final forward state (natural order) or NULL
};

template <int R>
__device__ __forceinline__ void load_state_natural(Cd (&x)[1 << R], const FusedDesc& d, uint32_t tid, int t) {
#pragma unroll
  for (int a = 0; a < (1 << R); ++a) x[a] = {d.init_tab[__popc(tid | ((uint32_t)a << t))], 0.0};
}

// x(y) <- x(map(y)) over the whole state, through LDS (natural layout in and out).
template <int R>
__device__ __forceinline__ void permute_gather(Cd (&x)[1 << R], void* lds, uint32_t m1, uint32_t m2, uint32_t tid,
                                               int t) {
#pragma unroll
  for (int a = 0; a < (1 << R); ++a) Stor<double>::lds_st(lds, swz<double>(tid | ((uint32_t)a << t)), x[a]);
  __syncthreads();
#pragma unroll
  for (int a = 0; a < (1 << R); ++a) {
    const uint32_t src = ladder_of(tid | ((uint32_t)a << t), m1, m2);
    x[a] = Stor<double>::template lds_ld<double>(lds, swz<double>(src));
  }
  __syncthreads();
}

// x(map(y)) <- x(y): the inverse of permute_gather with the same map.
template <int R>
__device__ __forceinline__ void permute_scatter(Cd (&x)[1 << R], void* lds, uint32_t m1, uint32_t m2, uint32_t tid,
                                                int t) {
#pragma unroll
  for (int a = 0; a < (1 << R); ++a) {
    const uint32_t dst = ladder_of(tid | ((uint32_t)a << t), m1, m2);
    Stor<double>::lds_st(lds, swz<double>(dst), x[a]);
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < (1 << R); ++a) x[a] = Stor<double>::template lds_ld<double>(lds, swz<double>(tid | ((uint32_t)a << t)));
  __syncthreads();
}

template <int R>
__device__ __forceinline__ double diag_at_reg(const FusedDesc& d, uint32_t x) {
  double D = 0.0;
  for (int k = 0; k < d.ndterms; ++k) {
    double z = ((x >> d.dterms[k].a) & 1u) ? -1.0 : 1.0;
    if (d.dterms[k].b >= 0) z *= ((x >> d.dterms[k].b) & 1u) ? -1.0 : 1.0;
    D += d.dterms[k].w * z;
  }
  return D;
}

__device__ __forceinline__ double block_sum(double v, double* red, uint32_t tid, int nthr) {
  const int width = nthr < 64 ? nthr : 64;
  v = wave_sum(v, width);
  const int nwaves = (nthr + 63) / 64;
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < nwaves; ++w) s += red[w];
  return s;
}

// Applies (forward = true) or undoes the single-qubit layer: for McClean one
// gate per qubit, for QAOA RX(beta) on every qubit.  With reduce, the
// reductions Im<lam|P_q|psi> of every qubit are added to acc_out[q] (McClean)
// or summed into acc_out[0] (QAOA) before the gates.
template <int R>
__device__ __forceinline__ void layer_rot(Cd (&psi)[1 << R], Cd (&lam)[1 << R], bool two, const FusedDesc& d, int i,
                                          double sign, void* lds, uint32_t tid, int t, double* red, double* acc_out,
                                          bool reduce) {
  const int n = d.n;
  int los[kMaxStages];
  int ns = 1;
  los[0] = t;
#pragma unroll
  for (int s = 1; s < kMaxStages; ++s) {
    los[s] = 0;
    if (ns == s && los[s - 1] > 0) {
      const int lo = t - R * s;
      los[s] = lo < 0 ? 0 : lo;
      ns = s + 1;
    }
  }
  const int nthr = (int)blockDim.x;
#pragma unroll 1
  for (int s = 0; s < ns; ++s) {
    if (s > 0) {
      lds_transpose<double, R>(psi, lds, tid, los[s - 1], los[s]);
      if (two) lds_transpose<double, R>(lam, lds, tid, los[s - 1], los[s]);
    }
    const int lo = los[s];
    const int nnew = nnew_of(los, s, R);
    static_for<R>([&](auto kc) __attribute__((always_inline)) {
      constexpr int kk = decltype(kc)::value;
      if (kk < nnew) {
        const int b = lo + kk;        // global bit == local bit (one tile)
        const int q = n - 1 - b;
        int ty;
        double c, sn;
        if (d.kind == 0) {
          ty = d.gtype[i * n + q];
          c = d.gcs[2 * (i * n + q)];
          sn = d.gcs[2 * (i * n + q) + 1];
        } else {
          ty = G_X;
          c = cos(0.5 * d.params[2 * i]);
          sn = sin(0.5 * d.params[2 * i]);
        }
        if (reduce) {
          const double v = block_sum(red_bit<R, kk>(lam, psi, ty), red, tid, nthr);
          if (tid == 0) acc_out[d.kind == 0 ? q : 0] += v;
        }
        gate_bit<R, kk>(psi, ty, c, sign * sn);
        if (two) gate_bit<R, kk>(lam, ty, c, sign * sn);
      }
    });
  }
  if (ns > 1) {
    lds_transpose<double, R>(psi, lds, tid, los[ns - 1], t);
    if (two) lds_transpose<double, R>(lam, lds, tid, los[ns - 1], t);
  }
}

template <int R>
__global__ __launch_bounds__(512) void k_fused(const FusedDesc* __restrict__ descs) {
  extern __shared__ __align__(16) char lds[];
  const FusedDesc d = descs[blockIdx.x];
  const uint32_t tid = threadIdx.x;
  const int n = d.n, L = d.L, t = n - R;
  const int nthr = (int)blockDim.x;
  double* red = reinterpret_cast<double*>(lds + ((size_t)16 << n));
  double* acc = red + 16;   // per-qubit accumulators (thread 0 only writes)
  Cd psi[1 << R], lam[1 << R];
  load_state_natural<R>(psi, d, tid, t);
  // ---- forward sweep ----
  for (int i = 0; i < L; ++i) {
    if (d.kind == 0) {
      if (d.ent == 0) {
        permute_gather<R>(psi, lds, d.lad1, d.lad2, tid, t);
      } else {
#pragma unroll
        for (int a = 0; a < (1 << R); ++a) {
          const double sg = cz_sign(tid | ((uint32_t)a << t), n);
          psi[a] = {sg * psi[a].re, sg * psi[a].im};
        }
      }
    } else {
      const double g = d.params[2 * i + 1];
#pragma unroll
      for (int a = 0; a < (1 << R); ++a) {
        double sn, cs;
        sincos(g * diag_at_reg<R>(d, tid | ((uint32_t)a << t)), &sn, &cs);
        psi[a] = cmul(Cd{cs, -sn}, psi[a]);
      }
    }
    layer_rot<R>(psi, lam, false, d, i, 1.0, lds, tid, t, red, acc, false);
  }
  // ---- lambda = O psi, E = <psi|O|psi> ----
  double e = 0.0;
#pragma unroll
  for (int a = 0; a < (1 << R); ++a) {
    const double D = diag_at_reg<R>(d, tid | ((uint32_t)a << t));
    lam[a] = {D * psi[a].re, D * psi[a].im};
  }
  if (d.nxy > 0) {
#pragma unroll
    for (int a = 0; a < (1 << R); ++a) Stor<double>::lds_st(lds, swz<double>(tid | ((uint32_t)a << t)), psi[a]);
    __syncthreads();
    for (int k = 0; k < d.nxy; ++k) {
      const int b = d.xy[k].a;
      const double w = d.xy[k].w;
#pragma unroll
      for (int a = 0; a < (1 << R); ++a) {
        const uint32_t x = tid | ((uint32_t)a << t);
        const Cd p = Stor<double>::template lds_ld<double>(lds, swz<double>(x ^ (1u << b)));
        if (d.xy[k].b == 0) {
          lam[a].re += w * p.re;
          lam[a].im += w * p.im;
        } else {
          const double sg = ((x >> b) & 1u) ? 1.0 : -1.0;
          lam[a].re += w * (-sg * p.im);
          lam[a].im += w * (sg * p.re);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < (1 << R); ++a) e += psi[a].re * lam[a].re + psi[a].im * lam[a].im;
  e = block_sum(e, red, tid, nthr);
  if (tid == 0) d.out[0] = e;
  if (d.psi_out) {
#pragma unroll
    for (int a = 0; a < (1 << R); ++a) {
      const uint32_t x = tid | ((uint32_t)a << t);
      if (d.psi_c64) Stor<float>::st(d.psi_out, x, psi[a]);
      else Stor<double>::st(d.psi_out, x, psi[a]);
    }
  }
  if (!d.grad) return;
  // ---- adjoint sweep ----
  for (int i = L - 1; i >= 0; --i) {
    if (tid == 0)
      for (int q = 0; q < (d.kind == 0 ? n : 1); ++q) acc[q] = 0.0;
    layer_rot<R>(psi, lam, true, d, i, -1.0, lds, tid, t, red, acc, true);
    if (d.kind == 0) {
      if (d.ent == 0) {
        permute_scatter<R>(psi, lds, d.lad1, d.lad2, tid, t);
        permute_scatter<R>(lam, lds, d.lad1, d.lad2, tid, t);
      } else {
#pragma unroll
        for (int a = 0; a < (1 << R); ++a) {
          const double sg = cz_sign(tid | ((uint32_t)a << t), n);
          psi[a] = {sg * psi[a].re, sg * psi[a].im};
          lam[a] = {sg * lam[a].re, sg * lam[a].im};
        }
      }
      if (tid == 0)
        for (int q = 0; q < n; ++q) d.out[1 + i * n + q] = acc[q];
    } else {
      const double g = d.params[2 * i + 1];
      double cr = 0.0;
#pragma unroll
      for (int a = 0; a < (1 << R); ++a) {
        const double D = diag_at_reg<R>(d, tid | ((uint32_t)a << t));
        cr += 2.0 * D * imcj(lam[a], psi[a]);
        double sn, cs;
        sincos(-g * D, &sn, &cs);
        psi[a] = cmul(Cd{cs, -sn}, psi[a]);
        lam[a] = cmul(Cd{cs, -sn}, lam[a]);
      }
      cr = block_sum(cr, red, tid, nthr);
      if (tid == 0) {
        d.out[1 + 2 * i] = acc[0];
        d.out[2 + 2 * i] = cr;
      }
    }
  }
}

size_t fused_lds_bytes(int n) { return ((size_t)16 << n) + (16 + 64) * sizeof(double); }

hipError_t launch_fused(const void* descs, int batch, int n, hipStream_t st) {
  const int R = n < kRegBits ? n : kRegBits;
  const int threads = 1 << (n - R);
  const size_t lds = fused_lds_bytes(n);
  const FusedDesc* d = (const FusedDesc*)descs;
  switch (R) {
    case 4: hipLaunchKernelGGL(k_fused<4>, dim3(batch), dim3(threads), lds, st, d); break;
    case 3: hipLaunchKernelGGL(k_fused<3>, dim3(batch), dim3(threads), lds, st, d); break;
    case 2: hipLaunchKernelGGL(k_fused<2>, dim3(batch), dim3(threads), lds, st, d); break;
    case 1: hipLaunchKernelGGL(k_fused<1>, dim3(batch), dim3(threads), lds, st, d); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t set_fused_attributes() {
  const int maxlds = (int)fused_lds_bytes(kFusedMaxN);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fused<4>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, maxlds);
  return e;
}

size_t fused_desc_size() { return sizeof(FusedDesc); }

}  // namespace qg
