// Tile-pass engine for the statevector grad_run path on CDNA4 (gfx950).
//
// A statevector of 2^n amplitudes is processed in "passes".  Each pass views
// the index bits as K local bits and n-K tile bits; a workgroup handles one
// tile (2^K amplitudes) at a time:
//
//   natural layout : thread bits = local bits [0, t), register bits = [t, K)
//                    (t = K - R, each thread holds 2^R amplitudes in VGPRs)
//   stage s >= 1   : register bits = local [lo_s, lo_s + R) with
//                    lo_s = max(0, t - R*s); reached by an LDS transpose
//
// Every single-qubit gate of the pass is applied to the register bits of the
// stage in which its bit first appears, so all K gates of a tile cost
// ceil(t/R) + 1 LDS round trips and one HBM read + write per state.
// Arithmetic is FP64 in registers for the propagated state; LDS holds the
// storage type, so a complex64 state is rounded at each LDS transpose and at
// the store (a few times per pass, never per gate).  In an adjoint pass the
// read-only companion state (psi from the history) may be held in FP32
// registers: it only enters the FP64-accumulated reductions.
//
// The CNOT ladder of the McClean circuit (reference
// qradient/physical_components/state.py:209-251) is a GF(2)-linear map of the
// index whose high output bits depend only on high input bits, so it is
// applied as an address map on the global load (forward) or store (reverse)
// of the low-bit pass: no pass of its own.
//
// Reductions Im<B|P_q|A> (the reference's gradient entries,
// mc_clean.py:68-76) are accumulated in FP64, summed per wave and per
// workgroup in a fixed order and written to partial rows; k_finalize sums the
// rows in a fixed order, so results are bitwise reproducible for a plan.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "qg_gate_asm.hpp"
#include "qg_internal.h"

namespace qg {

template <class T>
struct Cx {
  T re, im;
};
typedef Cx<double> Cd;

typedef float f2v __attribute__((ext_vector_type(2)));
typedef double d2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) f2v gf2v;
typedef __attribute__((address_space(1))) d2v gd2v;
typedef __attribute__((address_space(3))) f2v lf2v;
typedef __attribute__((address_space(3))) d2v ld2v;

// Global (address space 1) loads/stores compile to global_load/global_store.
template <class S>
struct Stor;
template <>
struct Stor<float> {
  typedef f2v V;
  // Amplitude i of the state at p.  For states under 4 GiB the byte offset is
  // formed in 32 bits, so the address is a scalar base plus one VGPR.
  template <class T>
  static __device__ __forceinline__ Cx<T> ld(const void* p, uint32_t i) {
    const f2v v = *(const gf2v*)((const char*)p + (uint32_t)(i << 3));
    return {(T)v.x, (T)v.y};
  }
  template <class T>
  static __device__ __forceinline__ Cx<T> ld_big(const void* p, uint32_t i) {
    const f2v v = ((const gf2v*)p)[i];
    return {(T)v.x, (T)v.y};
  }
  template <class T>
  static __device__ __forceinline__ void st(void* p, uint32_t i, Cx<T> v) {
    f2v w;
    w.x = (float)v.re;
    w.y = (float)v.im;
    *(gf2v*)((char*)p + (uint32_t)(i << 3)) = w;
  }
  template <class T>
  static __device__ __forceinline__ void st_big(void* p, uint32_t i, Cx<T> v) {
    f2v w;
    w.x = (float)v.re;
    w.y = (float)v.im;
    ((gf2v*)p)[i] = w;
  }
  template <class T>
  static __device__ __forceinline__ void lds_st(void* lds, uint32_t i, Cx<T> v) {
    f2v w;
    w.x = (float)v.re;
    w.y = (float)v.im;
    ((lf2v*)lds)[i] = w;
  }
  template <class T>
  static __device__ __forceinline__ Cx<T> lds_ld(const void* lds, uint32_t i) {
    const f2v v = ((const lf2v*)lds)[i];
    return {(T)v.x, (T)v.y};
  }
  // byte-offset forms
  template <class T>
  static __device__ __forceinline__ Cx<T> ld_o(const void* p, uint32_t off) {
    const f2v v = *(const gf2v*)((const char*)p + off);
    return {(T)v.x, (T)v.y};
  }
  // non-temporal: a read-once stream that should not displace cached data
  template <class T>
  static __device__ __forceinline__ Cx<T> ld_o_nt(const void* p, uint32_t off) {
    const f2v v = __builtin_nontemporal_load((const gf2v*)((const char*)p + off));
    return {(T)v.x, (T)v.y};
  }
  template <class T>
  static __device__ __forceinline__ void st_o(void* p, uint32_t off, Cx<T> v) {
    f2v w;
    w.x = (float)v.re;
    w.y = (float)v.im;
    *(gf2v*)((char*)p + off) = w;
  }
  // by absolute LDS address
  template <class T>
  static __device__ __forceinline__ void lds_st_a(uint32_t addr, Cx<T> v) {
    f2v w;
    w.x = (float)v.re;
    w.y = (float)v.im;
    *(lf2v*)(uintptr_t)addr = w;
  }
  template <class T>
  static __device__ __forceinline__ Cx<T> lds_ld_a(uint32_t addr) {
    const f2v v = *(const lf2v*)(uintptr_t)addr;
    return {(T)v.x, (T)v.y};
  }
  template <class T>
  static __device__ __forceinline__ void lds_st_b(void* lds, uint32_t off, Cx<T> v) {
    f2v w;
    w.x = (float)v.re;
    w.y = (float)v.im;
    *(lf2v*)((char*)lds + off) = w;
  }
  template <class T>
  static __device__ __forceinline__ Cx<T> lds_ld_b(const void* lds, uint32_t off) {
    const f2v v = *(const lf2v*)((const char*)lds + off);
    return {(T)v.x, (T)v.y};
  }
};
template <>
struct Stor<double> {
  typedef d2v V;
  template <class T>
  static __device__ __forceinline__ Cx<T> ld(const void* p, uint32_t i) {
    const d2v v = *(const gd2v*)((const char*)p + (uint32_t)(i << 4));
    return {(T)v.x, (T)v.y};
  }
  template <class T>
  static __device__ __forceinline__ Cx<T> ld_big(const void* p, uint32_t i) {
    const d2v v = ((const gd2v*)p)[i];
    return {(T)v.x, (T)v.y};
  }
  template <class T>
  static __device__ __forceinline__ void st(void* p, uint32_t i, Cx<T> v) {
    d2v w;
    w.x = (double)v.re;
    w.y = (double)v.im;
    *(gd2v*)((char*)p + (uint32_t)(i << 4)) = w;
  }
  template <class T>
  static __device__ __forceinline__ void st_big(void* p, uint32_t i, Cx<T> v) {
    d2v w;
    w.x = (double)v.re;
    w.y = (double)v.im;
    ((gd2v*)p)[i] = w;
  }
  template <class T>
  static __device__ __forceinline__ void lds_st(void* lds, uint32_t i, Cx<T> v) {
    d2v w;
    w.x = (double)v.re;
    w.y = (double)v.im;
    ((ld2v*)lds)[i] = w;
  }
  template <class T>
  static __device__ __forceinline__ Cx<T> lds_ld(const void* lds, uint32_t i) {
    const d2v v = ((const ld2v*)lds)[i];
    return {(T)v.x, (T)v.y};
  }
  template <class T>
  static __device__ __forceinline__ Cx<T> ld_o(const void* p, uint32_t off) {
    const d2v v = *(const gd2v*)((const char*)p + off);
    return {(T)v.x, (T)v.y};
  }
  template <class T>
  static __device__ __forceinline__ Cx<T> ld_o_nt(const void* p, uint32_t off) {
    const d2v v = __builtin_nontemporal_load((const gd2v*)((const char*)p + off));
    return {(T)v.x, (T)v.y};
  }
  template <class T>
  static __device__ __forceinline__ void st_o(void* p, uint32_t off, Cx<T> v) {
    d2v w;
    w.x = (double)v.re;
    w.y = (double)v.im;
    *(gd2v*)((char*)p + off) = w;
  }
  // by absolute LDS address
  template <class T>
  static __device__ __forceinline__ void lds_st_a(uint32_t addr, Cx<T> v) {
    d2v w;
    w.x = (double)v.re;
    w.y = (double)v.im;
    *(ld2v*)(uintptr_t)addr = w;
  }
  template <class T>
  static __device__ __forceinline__ Cx<T> lds_ld_a(uint32_t addr) {
    const d2v v = *(const ld2v*)(uintptr_t)addr;
    return {(T)v.x, (T)v.y};
  }
  template <class T>
  static __device__ __forceinline__ void lds_st_b(void* lds, uint32_t off, Cx<T> v) {
    d2v w;
    w.x = (double)v.re;
    w.y = (double)v.im;
    *(ld2v*)((char*)lds + off) = w;
  }
  template <class T>
  static __device__ __forceinline__ Cx<T> lds_ld_b(const void* lds, uint32_t off) {
    const d2v v = *(const ld2v*)((const char*)lds + off);
    return {(T)v.x, (T)v.y};
  }
};

// Storage of a pass kernel: the element type of the states in HBM (G) and
// of the LDS tiles (L).  float / double: complex64 / complex128 throughout;
// F32L64: complex64 states with FP64 LDS tiles, so a complex64 state is
// rounded only when it is stored (once per pass) instead of at every
// transpose as well.
struct F32L64 {};
template <class S>
struct Store {
  typedef S G;
  typedef S L;
};
template <>
struct Store<F32L64> {
  typedef float G;
  typedef double L;
};

// Uniform-index read of a read-only table through the constant address space:
// compiles to a scalar load (lgkmcnt only), so it never waits on the vector
// memory counter that tracks the tile's global loads and stores.
template <class T>
__device__ __forceinline__ T ldc(const T* p, int i) {
#if defined(__HIP_DEVICE_COMPILE__)
  return ((const __attribute__((address_space(4))) T*)p)[i];
#else
  return p[i];  // host pass: the body is parsed, never run
#endif
}

// LDS slot of local index j: a GF(2)-linear XOR swizzle (a bijection on
// [0, 2^K)) under which every stage layout's 8-byte (complex64) or 16-byte
// (complex128) accesses are free of bank conflicts, for 32 or 64 banks of 4 B.
template <class S>
__device__ __forceinline__ uint32_t swz(uint32_t j) {
  if constexpr (sizeof(S) == 4) return j ^ ((j >> 4) & 31u);
  else return j ^ ((j >> 4) & 15u);
}

// Local index held by (thread tid, register a) when register bits sit at [lo, lo+R).
template <int R>
__device__ __forceinline__ uint32_t lidx(uint32_t tid, uint32_t a, int lo) {
  const uint32_t lowmask = (1u << lo) - 1u;
  return (tid & lowmask) | (a << lo) | ((tid >> lo) << (lo + R));
}

// Global bit offset of a local index (local bits [0,c) U [p, p+K-c)).
__device__ __forceinline__ uint32_t gl_local(uint32_t j, int c, int p) {
  const uint32_t cm = (1u << c) - 1u;
  return (j & cm) | ((j >> c) << p);
}

// Global bit offset of tile index tt (tile bits [c,p) U [p+K-c, n)).
__device__ __forceinline__ uint32_t gl_tile(uint32_t tt, int c, int p, int K) {
  const int w1 = p - c;
  const uint32_t m1 = (w1 >= 32) ? ~0u : ((1u << w1) - 1u);
  const uint32_t hi = (w1 >= 32) ? 0u : (tt >> w1);
  const int sh = p + K - c;
  return ((tt & m1) << c) | (sh >= 32 ? 0u : (hi << sh));
}

__device__ __forceinline__ int gpos_of_local(int l, int c, int p) { return l < c ? l : p + (l - c); }

// The CNOT-ladder index map in shift form.  Each CNOT's target is the bit just
// below its control, and the ladder acts in two groups of disjoint bonds, so
// the image of bit k is bit k plus at most bits k-1 and k-2: the GF(2)-linear
// map is v ^ ((v & m1) >> 1) ^ ((v & m2) >> 2), with the masks derived on the
// host from the ladder's columns.  Three ALU operations instead of a loop
// over the n columns.
__device__ __forceinline__ uint32_t ladder_of(uint32_t v, uint32_t m1, uint32_t m2) {
  return v ^ ((v & m1) >> 1) ^ ((v & m2) >> 2);
}

__device__ __forceinline__ double cz_sign(uint32_t x, int n) {
  const uint32_t chain = (n >= 2) ? (uint32_t)((1ull << (n - 1)) - 1ull) : 0u;
  return (__popc(x & (x >> 1) & chain) & 1) ? -1.0 : 1.0;
}

// ---------------------------------------------------------------------------
// 2x2 gates on register bit KK: U = c*I - i*s*P for P = X, Y, Z (reference
// state.py:90-92, 142-144, 168-170).  The planner negates s for inverses.
// ---------------------------------------------------------------------------
template <int R, int KK, class T>
__device__ __forceinline__ void gate_bit(Cx<T> (&x)[1 << R], int type, double cd, double sd) {
  constexpr int A = 1 << R;
  const T c = (T)cd, s = (T)sd;
  if (type == G_X) {
#pragma unroll
    for (int a = 0; a < A; ++a) {
      if (a & (1 << KK)) continue;
      const Cx<T> x0 = x[a], x1 = x[a | (1 << KK)];
      // x0' = c x0 - i s x1 ; x1' = c x1 - i s x0
      x[a] = {c * x0.re + s * x1.im, c * x0.im - s * x1.re};
      x[a | (1 << KK)] = {c * x1.re + s * x0.im, c * x1.im - s * x0.re};
    }
  } else if (type == G_Y) {
#pragma unroll
    for (int a = 0; a < A; ++a) {
      if (a & (1 << KK)) continue;
      const Cx<T> x0 = x[a], x1 = x[a | (1 << KK)];
      // -iY = [[0,-1],[1,0]]: x0' = c x0 - s x1 ; x1' = c x1 + s x0
      x[a] = {c * x0.re - s * x1.re, c * x0.im - s * x1.im};
      x[a | (1 << KK)] = {c * x1.re + s * x0.re, c * x1.im + s * x0.im};
    }
  } else if (type == G_Z) {
#pragma unroll
    for (int a = 0; a < A; ++a) {
      const Cx<T> v = x[a];
      // bit 0: (c - i s) v ; bit 1: (c + i s) v
      const T ss = (a & (1 << KK)) ? s : -s;
      x[a] = {c * v.re - ss * v.im, c * v.im + ss * v.re};
    }
  }
}

template <class T>
__device__ __forceinline__ double imcj(const Cd& l, const Cx<T>& p) {  // Im(conj(l) p)
  return l.re * (double)p.im - l.im * (double)p.re;
}
template <class T>
__device__ __forceinline__ double recj(const Cd& l, const Cx<T>& p) {  // Re(conj(l) p)
  return l.re * (double)p.re + l.im * (double)p.im;
}

// sum over the register pairs of bit KK of Im<l|P|p>, in FP64: FMA chains
// into two partial sums (even / odd pairs) added at the end, a fixed order.
__device__ __forceinline__ double fmd(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <int R, int KK, class T>
__device__ __forceinline__ double red_bit(const Cd (&l)[1 << R], const Cx<T> (&p)[1 << R], int type) {
  constexpr int A = 1 << R;
  constexpr int M = 1 << KK;
  double s0 = 0.0, s1 = 0.0;
  if (type == G_Z) {
    // sum_a sign(a) Im(conj(l_a) p_a)
#pragma unroll
    for (int a = 0; a < A; ++a) {
      const double pr = (double)p[a].re, pi = (double)p[a].im;
      double& acc = (a & 1) ? s1 : s0;
      if (a & M) {
        acc = fmd(-l[a].re, pi, acc);
        acc = fmd(l[a].im, pr, acc);
      } else {
        acc = fmd(l[a].re, pi, acc);
        acc = fmd(-l[a].im, pr, acc);
      }
    }
  } else if (type == G_X) {
    // Im(conj(l0) p1) + Im(conj(l1) p0)
    int k = 0;
#pragma unroll
    for (int a = 0; a < A; ++a) {
      if (a & M) continue;
      const int b = a | M;
      double& acc = (k++ & 1) ? s1 : s0;
      acc = fmd(l[a].re, (double)p[b].im, acc);
      acc = fmd(-l[a].im, (double)p[b].re, acc);
      acc = fmd(l[b].re, (double)p[a].im, acc);
      acc = fmd(-l[b].im, (double)p[a].re, acc);
    }
  } else if (type == G_Y) {
    // Re(conj(l1) p0) - Re(conj(l0) p1)
    int k = 0;
#pragma unroll
    for (int a = 0; a < A; ++a) {
      if (a & M) continue;
      const int b = a | M;
      double& acc = (k++ & 1) ? s1 : s0;
      acc = fmd(l[b].re, (double)p[a].re, acc);
      acc = fmd(l[b].im, (double)p[a].im, acc);
      acc = fmd(-l[a].re, (double)p[b].re, acc);
      acc = fmd(-l[a].im, (double)p[b].im, acc);
    }
  }
  return s0 + s1;
}

// LDS address of a pointer into the kernel's dynamic LDS.  k_pass has no
// static LDS, so its dynamic LDS starts at address 0: a per-thread base that
// includes it can be XORed with offsets (XOR equals addition on bits the base
// leaves clear), and the accesses need no separate base add.
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) char*)p;
}

// The uniform term of an LDS byte offset, in a scalar register: computed by
// the scalar unit, and kept from being folded back into per-lane swizzles.
template <class S>
__device__ __forceinline__ uint32_t swz_u(uint32_t j) {
  constexpr int SH = sizeof(S) == 4 ? 3 : 4;
  return __builtin_amdgcn_readfirstlane(swz<S>(j) << SH);
}

// The swizzle is GF(2)-linear and the thread and register parts of a local
// index have disjoint bits, so slot(lidx) = swz(thread part) ^ swz(a << lo):
// one per-thread base per layout and a uniform term per register.
// Byte offsets are formed directly (the slot shift is linear too), so each
// access costs one VALU XOR with a scalar term.
template <class S, int R, class T>
__device__ __forceinline__ void lds_transpose(Cx<T> (&x)[1 << R], void* lds, uint32_t tid, int lo_from, int lo_to) {
  constexpr int A = 1 << R;
  constexpr int SH = sizeof(S) == 4 ? 3 : 4;
  lo_from = __builtin_amdgcn_readfirstlane(lo_from);
  lo_to = __builtin_amdgcn_readfirstlane(lo_to);
  uint32_t bw = lds_addr(lds) + (swz<S>(lidx<R>(tid, 0, lo_from)) << SH);
  uint32_t br = lds_addr(lds) + (swz<S>(lidx<R>(tid, 0, lo_to)) << SH);
  // opaque bases: otherwise the shift common to base and scalar term is
  // factored out of the XOR, costing a second VALU op per address
  asm volatile("" : "+v"(bw), "+v"(br));
#pragma unroll
  for (int a = 0; a < A; ++a) Stor<S>::lds_st_a(bw ^ swz_u<S>((uint32_t)a << lo_from), x[a]);
  __syncthreads();
#pragma unroll
  for (int a = 0; a < A; ++a) x[a] = Stor<S>::template lds_ld_a<T>(br ^ swz_u<S>((uint32_t)a << lo_to));
  __syncthreads();
}

// D(x) for the 2^R register values of one thread, with the thread's other
// bits fixed: D(a) = c0 + sum_k s_k(a) h_k + pwc[a], s_k(a) = -1 where
// register bit k of a is set.  The host splits the terms of D by the pass's
// register bits (PassDesc::cterms): c0 sums the terms on other bits, h_k the
// terms pairing register bit k with another bit, and pwc[a] (a table in the
// descriptor) the terms on register bits only.  A term is (mask, w) with
// value w * (-1)^popcount(x & mask) over the thread's other bits.
template <int R>
struct DiagParts {
  double c0;
  double h[R];
};

// w * (-1)^popcount(m): the parity flips the sign bit of w
__device__ __forceinline__ double flipw(double w, uint32_t m) {
  const uint64_t sb = (uint64_t)(__popc(m) & 1) << 63;
  return __longlong_as_double(__double_as_longlong(w) ^ (long long)sb);
}

template <int R>
__device__ __forceinline__ void diag_prepare(const PassDesc* __restrict__ d, uint32_t gthr, DiagParts<R>& o) {
  const DTerm* __restrict__ ct = d->cterms;
  const int nc = ldc(&d->nct, 0);
  double c0 = 0.0;
  for (int i = 0; i < nc; ++i) {
    const DTerm t = ldc(ct, i);
    c0 += flipw(t.w, gthr & (uint32_t)t.a);
  }
  int off = nc;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int nh = ldc(d->nht, k);
    double h = 0.0;
    for (int i = 0; i < nh; ++i) {
      const DTerm t = ldc(ct, off + i);
      h += flipw(t.w, gthr & (uint32_t)t.a);
    }
    o.h[k] = h;
    off += nh;
  }
  o.c0 = c0;
}

template <int R>
__device__ __forceinline__ double diag_value(const PassDesc* __restrict__ d, const DiagParts<R>& o, int a) {
  double v = o.c0 + ldc(d->pwc, a);
#pragma unroll
  for (int k = 0; k < R; ++k) v += ((a >> k) & 1) ? -o.h[k] : o.h[k];
  return v;
}

// exp(-i * gamma * D): PH = 1 reads a table (integer-valued D), PH = 2 evaluates sincos.
// exp(-i gamma D): PH 1 reads the table staged in LDS (tab), PH 2 computes it.
template <int PH>
__device__ __forceinline__ Cd phase_of(const PassDesc* __restrict__ d, const double* tab, double D) {
  if constexpr (PH == 1) {
    const int idx = (int)rint(D) + d->lut_off;
    return {tab[2 * idx], tab[2 * idx + 1]};
  } else {
    double s, c;
    sincos(d->phase_gamma * D, &s, &c);
    return {c, -s};
  }
}

template <class T>
__device__ __forceinline__ Cx<T> cmul(Cd a, Cx<T> b) {
  return {(T)(a.re * b.re - a.im * b.im), (T)(a.re * b.im + a.im * b.re)};
}

template <int I>
using ic = std::integral_constant<int, I>;

template <class F, int... Is>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
  (f(ic<Is>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// *p += v for an LDS slot that only this thread updates, as an LDS atomic
// add without return: the wave does not wait for the slot's old value (such
// a wait would also wait for every LDS read issued before it).  The adds of
// one thread reach the slot in program order, so the slot holds the same
// sum as a read-modify-write in that order.
__device__ __forceinline__ void lds_add(double* p, double v) {
  __builtin_amdgcn_ds_atomic_fadd_f64((__attribute__((address_space(3))) double*)p, v);
}

// x moved across lanes by the DPP control CTRL (each 32-bit half with one
// v_mov_b32_dpp); lanes of rows outside ROWM get 0.
template <int CTRL, int ROWM>
__device__ __forceinline__ double dpp_d(double x) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROWM, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROWM, 0xF, false);
  return __builtin_bit_cast(double, (unsigned long long)(unsigned)lo | ((unsigned long long)(unsigned)hi << 32));
}

// The sum of v over the wave (width lanes), returned in every lane, in a fixed
// order.  A full wave reduces with DPP moves, which stay in the vector ALU:
// pairs (quad_perm 1,0,3,2), quads (2,3,0,1), groups of 8 (row_half_mirror)
// and 16 (row_mirror) in every lane, then row_bcast:15 into rows 1 and 3 and
// row_bcast:31 into rows 2 and 3, which leaves the total in lane 63.  The
// shuffle tree it replaces (a chain of six ds_bpermute pairs per sum) had the
// complex128 adjoint passes at 24 qubits at 197 / 136 / 138 us against 167 /
// 131 / 134 us, and the complex64 one at 100.6 against 97.2 us per pass.
// Narrower waves (passes of fewer than 64 threads) keep the shuffle tree.
__device__ __forceinline__ double wave_sum(double v, int width) {
  if (width == 64) {
    v += dpp_d<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
    v += dpp_d<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
    v += dpp_d<0x141, 0xF>(v);  // row_half_mirror
    v += dpp_d<0x140, 0xF>(v);  // row_mirror
    v += dpp_d<0x142, 0xA>(v);  // row_bcast:15
    v += dpp_d<0x143, 0xC>(v);  // row_bcast:31
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
    return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
  }
  for (int off = width >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, width);
  return v;
}

__device__ __forceinline__ int nnew_of(const int* los, int s, int R) {
  if (s == 0) return R;
  const int d = los[s - 1] - los[s];
  return d < R ? d : R;
}

__device__ __forceinline__ double fmx(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmx(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// In-place rotation of the plane (u, v) by phi: [u; v] <- [[cos, -sin], [sin, cos]] [u; v]
// as three shears with a = tan(phi / 2), b = sin(phi) (|phi| <= pi / 2, so
// |a|, |b| <= 1): one FMA each, every update in place.
//
// The FMAs are written as inline asm with the accumulator tied to the output
// register: every gate type then updates the same registers, and the join
// after the per-bit gate-type branch needs no register copies.
__device__ __forceinline__ void shear3(double& u, double& v, double a, double b) {
  asm("v_fma_f64 %0, -%1, %2, %0" : "+v"(u) : "s"(a), "v"(v));
  asm("v_fma_f64 %0, %1, %2, %0" : "+v"(v) : "s"(b), "v"(u));
  asm("v_fma_f64 %0, -%1, %2, %0" : "+v"(u) : "s"(a), "v"(v));
}
__device__ __forceinline__ void shear3(float& u, float& v, float a, float b) {
  asm("v_fma_f32 %0, -%1, %2, %0" : "+v"(u) : "v"(a), "v"(v));
  asm("v_fma_f32 %0, %1, %2, %0" : "+v"(v) : "v"(b), "v"(u));
  asm("v_fma_f32 %0, -%1, %2, %0" : "+v"(u) : "v"(a), "v"(v));
}

// U = cos(phi) I - i sin(phi) P on register bit KK (see PassDesc), as two
// plane rotations per amplitude pair: X rotates (re0, im1) by -phi and
// (im0, re1) by +phi, Y rotates (re0, re1) and (im0, im1) by +phi, Z rotates
// (re0, im0) by -phi and (re1, im1) by +phi.
// Amplitude order of the pairs of register bit KK (pair p = entries 2p, 2p+1).
struct PairOrder {
  int m[16];
};
__host__ __device__ constexpr PairOrder pair_order(int kk) {
  PairOrder o{};
  int p = 0;
  for (int i = 0; i < 16; ++i)
    if (!(i & (1 << kk))) {
      o.m[2 * p] = i;
      o.m[2 * p + 1] = i | (1 << kk);
      ++p;
    }
  return o;
}

template <int R, int KK, class T>
__device__ __forceinline__ void gate_bit_s(Cx<T> (&x)[1 << R], int type, double ad, double bd) {
  constexpr int A = 1 << R;
  constexpr int M = 1 << KK;
  if constexpr (R == 4) {
    constexpr PairOrder po = pair_order(KK);
    if constexpr (std::is_same<T, double>::value) {
      asm(QG_GATE_ASM("v_fma_f64") : QG_GATE_OPERANDS(x, po.m) : "s"(type), "s"(ad), "s"(bd) : "scc");
    } else {
      const float a = (float)ad, b = (float)bd;
      asm(QG_GATE_ASM("v_fma_f32") : QG_GATE_OPERANDS(x, po.m) : "s"(type), "v"(a), "v"(b) : "scc");
    }
    return;
  }
  const T a = (T)ad, b = (T)bd, na = (T)(-ad), nb = (T)(-bd);
  if (type == G_X) {
#pragma unroll
    for (int i = 0; i < A; ++i) {
      if (i & M) continue;
      shear3(x[i].re, x[i | M].im, na, nb);
      shear3(x[i].im, x[i | M].re, a, b);
    }
  } else if (type == G_Y) {
#pragma unroll
    for (int i = 0; i < A; ++i) {
      if (i & M) continue;
      shear3(x[i].re, x[i | M].re, a, b);
      shear3(x[i].im, x[i | M].im, a, b);
    }
  } else if (type == G_Z) {
#pragma unroll
    for (int i = 0; i < A; ++i) {
      if (i & M) shear3(x[i].re, x[i].im, a, b);
      else shear3(x[i].re, x[i].im, na, nb);
    }
  }
}

// Global addresses (amplitude indices) of a thread's 2^R amplitudes in a tile
// whose thread-0 register-0 index is g: plain, or through the ladder map.
// SH > 0: byte offsets (index << SH) instead of indices.  The register part
// of an address is uniform: it is formed by the scalar unit, so each address
// costs one VALU OR / XOR with the thread's base.
template <int R, int SH = 0>
__device__ __forceinline__ void tile_addresses(const PassDesc* __restrict__ d, bool gather, uint32_t g,
                                               const int (&rpos)[R], const uint32_t (&rbit)[R], int n,
                                               uint32_t (&adr)[1 << R]) {
  if (gather) {
    // the map is linear and the register bits are disjoint from g's bits:
    // map(g | u) = map(g) ^ map(u), with map(u) uniform
    const uint32_t m1 = ldc(&d->lad1, 0), m2 = ldc(&d->lad2, 0);
    uint32_t mbase = ladder_of(g, m1, m2) << SH;
    asm volatile("" : "+v"(mbase));  // opaque (see lds_transpose)
#pragma unroll
    for (int a = 0; a < (1 << R); ++a) {
      uint32_t u = 0;
#pragma unroll
      for (int k = 0; k < R; ++k)
        if (a & (1 << k)) u |= rbit[k];
      adr[a] = mbase ^ __builtin_amdgcn_readfirstlane(ladder_of(u, m1, m2) << SH);
    }
  } else {
    uint32_t gb = g << SH;
    asm volatile("" : "+v"(gb));  // opaque (see lds_transpose)
#pragma unroll
    for (int a = 0; a < (1 << R); ++a) {
      uint32_t u = 0;
#pragma unroll
      for (int k = 0; k < R; ++k)
        if (a & (1 << k)) u |= rbit[k];
      adr[a] = gb | __builtin_amdgcn_readfirstlane(u << SH);
    }
  }
}

// ---------------------------------------------------------------------------
// One pass over the tiles assigned to this workgroup.
//   S  : storage (float: complex64, double: complex128, F32L64: complex64
//        states with FP64 LDS tiles; see Store)
//   R  : register bits per thread (2^R amplitudes)
//   TWO: an adjoint pass with the companion state A and the propagated B
//   PH : diagonal phase method (0 none, 1 table, 2 sincos)
//   POST: (adjoint passes) the general two-state form, with the diagonal ops
//        after the gates (phase, d/dgamma reduction, observable) compiled in:
//        A and B are both rotated and every reduction is taken in the stage
//        that holds its bit.  Without POST an adjoint pass takes the
//        reduce-first form: A is the pass output psi (read-only), every
//        reduction Im<B|P_q|A> is taken in the natural layout before any gate
//        is undone (each P_q commutes with every gate of the pass, so the pass
//        output is a valid frame for all of them), and only B is rotated.
//        With a phase (PH != 0) the reduce-first form is the adjoint of a
//        forward pass that ended with the phase: d/dgamma is taken at the pass
//        output, both states are taken back through the phase in the natural
//        layout, and the pass continues as above.
//   Forward passes (TWO false) with a phase apply it before or after the gates.
//
// The next tile's loads are issued before the current tile is processed (the
// raw storage-type values wait in registers), except in complex128 adjoint
// passes and adjoint passes with post-gate diagonal ops, where the registers
// are needed for the states and the diagonal terms.  In complex64
// adjoint passes the reductions are accumulated per thread across all tiles in
// LDS slots [local bit][thread] and reduced across the workgroup once, at the
// end; complex128 adjoint passes reduce per tile into per-wave slots.  Both
// use a fixed order, so results are bitwise reproducible.
// ---------------------------------------------------------------------------
template <class S, int R, bool TWO, int PH, bool POST, bool BIG, bool WIDE = false>
__device__ __forceinline__ void run_pass(const PassDesc* __restrict__ d, void* lds, uint32_t tid_in, uint32_t blk,
                                         uint32_t nblk) {
  constexpr int A = 1 << R;
  typedef typename Store<S>::G G;  // HBM element type
  typedef typename Store<S>::L LT; // LDS tile element type
  constexpr int GSH = sizeof(G) == 4 ? 3 : 4;  // log2 bytes per stored amplitude
  // WIDE: a 1024-thread workgroup (14 local bits), which must fit 128 VGPRs
  // and leaves no LDS for per-thread accumulators: no prefetch in adjoint
  // passes, per-tile wave sums
  constexpr bool PF = !(TWO && (sizeof(G) == 8 || POST || PH != 0 || WIDE));
  constexpr bool ACCL = TWO && sizeof(LT) == 4 && !WIDE;
  // diagonal ops after the gates: adjoint passes of the general form, and
  // forward passes with a phase (which may sit before or after the gates)
  constexpr bool POST_DIAG = POST || (!TWO && PH != 0);
  // the workgroup's FP64 sum of the diagonal reductions (d/dgamma, E)
  constexpr bool DSUM = TWO && (POST || PH != 0);
  // Register type of A: FP64, except in a reduce-first complex64 pass with a
  // phase or with 1024 threads, which only reads A in its FP64 reductions and
  // keeps it in FP32 registers (exact, or rounded once after the phase) to
  // stay within the register file.
  typedef typename std::conditional<TWO && !POST && (PH != 0 || WIDE) && sizeof(G) == 4, float, double>::type TAr;
  const int n = d->n, K = d->K, t = d->t, c = d->c, p = d->p;
  const uint32_t flags = d->flags;
  const uint32_t ntiles = 1u << (n - K);
  const uint32_t nthr = blockDim.x;
  const int width = nthr < 64 ? (int)nthr : 64;
  const int nwaves = (int)((nthr + 63) / 64);

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
  // trailing stages that only hold spectator bits are skipped
  if (d->nstage < ns) ns = d->nstage > 0 ? d->nstage : 1;

  int rpos[R];
  uint32_t rbit[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    rpos[k] = gpos_of_local(t + k, c, p);
    rbit[k] = 1u << rpos[k];
  }

  // LDS after the tile buffer: per-wave FP64 slots [wave][K + 1] (lane 0 of
  // each wave is the only writer of its row), for ACCL the per-thread
  // accumulators [local bit][thread], then the staged table (the initial
  // state's popcount table or the phase table).
  double* wacc = reinterpret_cast<double*>(reinterpret_cast<char*>(lds) + (sizeof(LT) * 2u << K));
  double* tacc = wacc + nwaves * (kMaxK + 1);
  double* tab = tacc + (ACCL ? K * (int)nthr : 0);
  const bool reduce = TWO && (flags & (F_REDUCE | F_DRED_POST | F_DRED_PRE | F_OBS_DIAG)) && d->row_base >= 0;
  const bool reduce_bits = reduce && (flags & F_REDUCE);
  const uint32_t wave = tid_in >> 6;
  const bool lane0 = (tid_in & 63u) == 0u;
  // staged tables: the phase table (PH 1) first, then the popcount table
  const bool use_lut = PH == 1 && (flags & (F_PHASE_PRE | F_PHASE_POST));
  const int nlut = use_lut ? 2 * (2 * d->lut_off + 1) : 0;
  const int ninit = (flags & F_INIT_TAB) ? n + 1 : 0;
  double* itab = tab + nlut;
  if (reduce || nlut + ninit > 0) {
    if (reduce) {
      for (uint32_t i = tid_in; i < (uint32_t)(nwaves * (K + 1)); i += nthr) wacc[i] = 0.0;
      if constexpr (ACCL) {
        for (int lb = 0; lb < K; ++lb) tacc[lb * nthr + tid_in] = 0.0;
      }
    }
    for (int i = (int)tid_in; i < nlut; i += (int)nthr) tab[i] = d->lut[i];
    for (int i = (int)tid_in; i < ninit; i += (int)nthr) itab[i] = d->init_tab[i];
    __syncthreads();
  }
  double dacc = 0.0;
  const bool gather = (flags & F_GATHER) != 0;
  // reduce-first adjoint: A is a loaded pass output, only read by the
  // reductions (never generated, rotated, sign-flipped after them or stored)
  constexpr bool RF = TWO && !POST;
  const bool load_a = RF || !(flags & F_INIT_TAB);
  const bool load_b = TWO && (flags & F_LOAD_B);
  // (general-form adjoints, which rotate A and store it back, read it normally:
  // the hint is compiled out of them, which keeps their register count)
  const bool nt_a = !POST && (flags & F_NT_A) != 0;
  // A's gates only matter for later reductions or for A's own use after the gates
  const bool a_used_after = !TWO || (flags & (F_STORE_A | F_DRED_POST | F_OBS_DIAG));

  // Tile order, chosen by the planner (F_XCD_ORDER): XCD-aware, workgroup blk
  // runs on XCD blk % 8 and each XCD gets a contiguous range of tiles (tiles
  // that share cache lines share an L2), or round-robin, tile blk + r * nblk
  // (the machine's accesses stay in one moving window of memory).
  const bool xcd = (flags & F_XCD_ORDER) && (nblk >= 8) && (nblk % 8 == 0) && (ntiles % nblk == 0);
  const uint32_t per_x = ntiles / 8, slots = nblk / 8;
  const uint32_t rounds = xcd ? ntiles / nblk : (blk < ntiles ? (ntiles - 1 - blk) / nblk + 1 : 0u);
  auto tile_of = [&](uint32_t r) -> uint32_t {
    return xcd ? (blk % 8) * per_x + r * slots + blk / 8 : blk + r * nblk;
  };

  Cx<G> ra[A];
  Cx<G> rb[TWO ? A : 1];
  // Issues the loads of tile tt into ra / rb (addresses derived from a
  // laundered thread index so they are not kept live across the tile).
  auto fetch = [&](uint32_t tt) __attribute__((always_inline)) {
    uint32_t tid = tid_in;
    asm volatile("" : "+v"(tid));
    uint32_t adr[A];
    tile_addresses<R, BIG ? 0 : GSH>(d, gather, gl_tile(tt, c, p, K) | gl_local(tid, c, p), rpos, rbit, n, adr);
    // every path defines every register (zeros when a state is not loaded),
    // so no value is carried from one tile to the next.  A is read with
    // non-temporal loads where the planner marks it as read once in this
    // sweep (F_NT_A: a history-mode forward pass's input, an adjoint pass's
    // history state), so that the lines the pass writes keep the caches
    // (measured at 24 qubits, complex64: 57.5 -> 51.4 us per forward pass).
    // Passes that work in place or in two ping-pong buffers, and passes over
    // states too large for the caches to keep, read normally.
    if (load_a) {
      if (nt_a) {
        // the empty asm statements keep the two load sequences apart: without
        // them the compiler merges the branches into one sequence of plain
        // loads, dropping the non-temporal hint
        asm volatile("");
#pragma unroll
        for (int a = 0; a < A; ++a)
          ra[a] = BIG ? Stor<G>::template ld_big<G>(d->a_in, adr[a]) : Stor<G>::template ld_o_nt<G>(d->a_in, adr[a]);
        asm volatile("");
      } else {
#pragma unroll
        for (int a = 0; a < A; ++a)
          ra[a] = BIG ? Stor<G>::template ld_big<G>(d->a_in, adr[a]) : Stor<G>::template ld_o<G>(d->a_in, adr[a]);
      }
    } else {
#pragma unroll
      for (int a = 0; a < A; ++a) ra[a] = {(G)0, (G)0};
    }
    if constexpr (TWO) {
      if (load_b) {
#pragma unroll
        for (int a = 0; a < A; ++a)
          rb[a] = BIG ? Stor<G>::template ld_big<G>(d->b_in, adr[a]) : Stor<G>::template ld_o<G>(d->b_in, adr[a]);
      } else {
#pragma unroll
        for (int a = 0; a < A; ++a) rb[a] = {(G)0, (G)0};
      }
    }
  };

  if (PF && rounds > 0) fetch(tile_of(0));
  for (uint32_t r = 0; r < rounds; ++r) {
    const uint32_t tt = tile_of(r);
    if (!PF) {
      // complex128 phase-free passes issue the tile's loads at raised wave
      // priority too (24 qubits: the high adjoint passes 147-148 -> 136-138
      // us; the phase passes measured no gain and keep the default)
      constexpr bool HI = sizeof(G) == 8 && PH == 0 && !BIG;
      if constexpr (HI) __builtin_amdgcn_s_setprio(2);
      fetch(tt);
      if constexpr (HI) __builtin_amdgcn_s_setprio(0);
    }
    uint32_t tid = tid_in;
    asm volatile("" : "+v"(tid));
    const uint32_t gthr = gl_tile(tt, c, p, K) | gl_local(tid, c, p);

    Cx<TAr> xa[A];
    Cd xb[TWO ? A : 1];
    if (load_a) {
#pragma unroll
      for (int a = 0; a < A; ++a) xa[a] = {(TAr)ra[a].re, (TAr)ra[a].im};
    } else {
      uint32_t adr[A];
      tile_addresses<R>(d, gather, gthr, rpos, rbit, n, adr);
#pragma unroll
      for (int a = 0; a < A; ++a) xa[a] = {(TAr)itab[__popc(adr[a])], (TAr)0};
    }
    if constexpr (TWO) {
#pragma unroll
      for (int a = 0; a < A; ++a) xb[a] = Cd{(double)rb[a].re, (double)rb[a].im};
    }

    // ---- diagonal ops in the natural layout, before the gates --------------
    if (flags & F_CZ_PRE) {
#pragma unroll
      for (int a = 0; a < A; ++a) {
        uint32_t x = gthr;
#pragma unroll
        for (int k = 0; k < R; ++k)
          if (a & (1 << k)) x |= rbit[k];
        const TAr sg = (TAr)cz_sign(x, n);
        xa[a] = {sg * xa[a].re, sg * xa[a].im};
      }
    }
    if constexpr (PH != 0) {
      // a reduce-first pass runs with a phase only when the phase opens it
      // (the branch is compiled out: a join here costs register copies)
      if (RF || (flags & F_PHASE_PRE)) {
        // adjoint of a pass whose forward ends with the phase: d/dgamma at the
        // pass output, then both states are taken back through the phase
        DiagParts<R> dp;
        diag_prepare<R>(d, gthr, dp);
        double xacc = 0.0;
#pragma unroll
        for (int a = 0; a < A; ++a) {
          const double D = diag_value<R>(d, dp, a);
          if constexpr (TWO) {
            if (RF || (flags & F_DRED_PRE)) {
              xacc += 2.0 * D * imcj(xb[a], xa[a]);
              // taken here, from the unphased states (otherwise the sum may
              // be sunk past the loop, keeping both copies of the states live)
              asm volatile("" : "+v"(xacc));
            }
          }
          const Cd ph = phase_of<PH>(d, tab, D);
          xa[a] = cmul(ph, xa[a]);
          if constexpr (TWO) xb[a] = cmul(ph, xb[a]);
          __builtin_amdgcn_sched_barrier(0);
        }
        dacc += xacc;
      }
    }
    // ---- reduce-first adjoint: all reductions in the natural layout -------
    if constexpr (TWO && !POST) {
      if (reduce_bits) {
        // register bits: amplitude pairs inside the thread
        static_for<R>([&](auto kc) __attribute__((always_inline)) {
          constexpr int kk = decltype(kc)::value;
          int lb = __builtin_amdgcn_readfirstlane(t + kk);
          asm volatile("" : "+s"(lb));
          if (ldc(d->own, lb)) {
            const double v = red_bit<R, kk>(xb, xa, ldc(d->gtype, lb));
            if constexpr (ACCL) {
              lds_add(&tacc[lb * nthr + tid], v);
            } else {
              const double w = wave_sum(v, width);
              if ((tid & 63) == 0) wacc[(tid >> 6) * (K + 1) + lb] += w;
            }
          }
        });
        // thread bits: Z needs only the thread's sum of Im(conj(B) A); X and
        // Y read the partner amplitude of A from LDS (natural layout)
        double m0 = 0.0, m1 = 0.0;
#pragma unroll
        for (int a = 0; a < A; ++a) {
          double& m = (a & 1) ? m1 : m0;
          m = fmd(xb[a].re, xa[a].im, m);
          m = fmd(-xb[a].im, xa[a].re, m);
        }
        const double msum = m0 + m1;
        constexpr int SH = sizeof(LT) == 4 ? 3 : 4;
        const int tu = __builtin_amdgcn_readfirstlane(t);
        // tid < 2^t: thread and register bits are disjoint
        uint32_t base = lds_addr(lds) + (swz<LT>(tid) << SH);
        asm volatile("" : "+v"(base));  // opaque (see lds_transpose)
#pragma unroll
        for (int a = 0; a < A; ++a) Stor<LT>::lds_st_a(base ^ swz_u<LT>((uint32_t)a << tu), xa[a]);
        __syncthreads();
        auto accum = [&](int lb, double v) __attribute__((always_inline)) {
          if constexpr (ACCL) {
            lds_add(&tacc[lb * nthr + tid], v);
          } else {
            const double w = wave_sum(v, width);
            if ((tid & 63) == 0) wacc[(tid >> 6) * (K + 1) + lb] += w;
          }
        };
        // Z on a thread bit: +-(the thread's sum), by the bit of the thread
        for (uint32_t zm = ldc(&d->zmask, 0); zm; zm &= zm - 1) {
          const int lb = __builtin_ctz(zm);
          accum(lb, ((tid >> lb) & 1u) ? -msum : msum);
        }
        // X / Y: the partner amplitudes of A (thread tid ^ 2^lb) from LDS,
        // raw; with FP32 tiles the next bit's reads are in flight while the
        // current bit is reduced.  The X bits and the Y bits are separate
        // masks, so each loop has one fixed formula; the bits come from the
        // mask with scalar bit scans (no table load, whose wait would also
        // wait for the partner reads in flight).
        typedef Cx<LT> Raw[A];
        auto rd = [&](int lb, Raw& P) __attribute__((always_inline)) {
          const uint32_t pb = base ^ swz_u<LT>(1u << lb);  // the partner thread tid ^ 2^lb
#pragma unroll
          for (int a = 0; a < A; ++a) P[a] = Stor<LT>::template lds_ld_a<LT>(pb ^ swz_u<LT>((uint32_t)a << tu));
        };
        auto use = [&](auto xc, int lb, const Raw& P) __attribute__((always_inline)) {
          constexpr bool isx = decltype(xc)::value;
          double v0 = 0.0, v1 = 0.0;
#pragma unroll
          for (int a = 0; a < A; ++a) {
            const double pr = (double)P[a].re, pi = (double)P[a].im;
            double& acc = (a & 1) ? v1 : v0;
            if constexpr (isx) {  // Im(conj(B_x) A_{x^q})
              acc = fmd(xb[a].re, pi, acc);
              acc = fmd(-xb[a].im, pr, acc);
            } else {              // Re(conj(B_x) A_{x^q}), signed below
              acc = fmd(xb[a].re, pr, acc);
              acc = fmd(xb[a].im, pi, acc);
            }
          }
          const double v = v0 + v1;
          // Y: -sg * Re(...), sg = +1 where the bit is 0
          accum(lb, isx ? v : (((tid >> lb) & 1u) ? v : -v));
        };
        // pops the lowest bit of a uniform mask
        auto pop = [](uint32_t& m) __attribute__((always_inline)) -> int {
          const int b = __builtin_ctz(m);
          m &= m - 1;
          return b;
        };
        auto run = [&](auto xc, uint32_t m) __attribute__((always_inline)) {
          if (!m) return;
          if constexpr (sizeof(LT) == 4 && !WIDE) {
            Raw P0, P1;
            int b0 = pop(m);
            rd(b0, P0);
#pragma unroll 1
            while (true) {
              if (!m) {
                use(xc, b0, P0);
                break;
              }
              const int b1 = pop(m);
              rd(b1, P1);
              use(xc, b0, P0);
              if (!m) {
                use(xc, b1, P1);
                break;
              }
              b0 = pop(m);
              rd(b0, P0);
              use(xc, b1, P1);
            }
          } else {
            Raw P0;
#pragma unroll 1
            while (m) {
              const int b = pop(m);
              rd(b, P0);
              use(xc, b, P0);
            }
          }
        };
        run(std::true_type{}, ldc(&d->xmask, 0));
        run(std::false_type{}, ldc(&d->ymask, 0));
        __syncthreads();  // the tile buffer is B's from here on
      }
    }

    // the next tile's loads go out after everything that reads this tile's
    // inputs, so no wait for this tile's data covers them.  They are issued at
    // raised wave priority, so the other wave of the SIMD, busy with its
    // stages, does not hold them back (complex64 adjoint, 24 qubits: 101.3-101.6
    // -> 100.3 us per pass; forward passes unchanged)
    if (PF && r + 1 < rounds) {
      __builtin_amdgcn_s_setprio(2);
      fetch(tile_of(r + 1));
      __builtin_amdgcn_s_setprio(0);
    }

    // ---- stages: reductions then gates on each newly visible register bit --
#pragma unroll 1
    for (int s = 0; s < kMaxStages; ++s) {
      if (s < ns) {
        if (s > 0) {
          if constexpr (!TWO || POST) lds_transpose<LT, R>(xa, lds, tid, los[s - 1], los[s]);
          if constexpr (TWO) lds_transpose<LT, R>(xb, lds, tid, los[s - 1], los[s]);
        }
        static_for<R>([&](auto kc) __attribute__((always_inline)) {
          constexpr int kk = decltype(kc)::value;
          // opaque to the optimiser: keeps the uniform slot loads next to
          // their use instead of hoisting them out of the loops
          int sl = __builtin_amdgcn_readfirstlane(s * R + kk);
          asm volatile("" : "+s"(sl));
          const int ty = ldc(d->stype, sl), fl = ldc(d->sflag, sl);
          if constexpr (TWO && POST) {
            if (fl & SL_RED) {
              const int lb = ldc(d->slb, sl);
              const double v = red_bit<R, kk>(xb, xa, ty);
              if constexpr (ACCL) {
                lds_add(&tacc[lb * nthr + tid], v);
              } else {
                const double w = wave_sum(v, width);
                if ((tid & 63) == 0) wacc[(tid >> 6) * (K + 1) + lb] += w;
              }
            }
          }
          const double ga = ldc(d->sa, sl), gb = ldc(d->sb, sl);
          if constexpr (TWO) {
            if constexpr (POST) {
              if (fl & SL_ROT_A) gate_bit_s<R, kk>(xa, ty, ga, gb);
            }
            if (fl & SL_ROT_B) gate_bit_s<R, kk>(xb, ty, ga, gb);
          } else {
            gate_bit_s<R, kk>(xa, ty, ga, gb);
          }
        });
      }
    }
    if (ns > 1) {
      if constexpr (!TWO || POST) lds_transpose<LT, R>(xa, lds, tid, los[ns - 1], t);
      if constexpr (TWO) lds_transpose<LT, R>(xb, lds, tid, los[ns - 1], t);
    }

    // index-derived values after the stages come from a laundered base, so
    // the compiler does not keep the per-amplitude indices of the pre-gate
    // ops live across the stages
    uint32_t gst = gthr;
    asm volatile("" : "+v"(gst));

    // ---- diagonal ops in the natural layout, after the gates --------------
    if (POST_DIAG && (flags & (F_DRED_POST | F_PHASE_POST | F_OBS_DIAG))) {
      DiagParts<R> dp;
      diag_prepare<R>(d, gst, dp);
      double xacc = 0.0;
#pragma unroll
      for (int a = 0; a < A; ++a) {
        const double D = diag_value<R>(d, dp, a);
        if constexpr (TWO) {
          if (flags & F_DRED_POST) xacc += 2.0 * D * imcj(xb[a], xa[a]);
        }
        if constexpr (PH != 0) {
          if (flags & F_PHASE_POST) {
            const Cd ph = phase_of<PH>(d, tab, D);
            xa[a] = cmul(ph, xa[a]);
            if constexpr (TWO) xb[a] = cmul(ph, xb[a]);
          }
        }
        if constexpr (TWO) {
          if (flags & F_OBS_DIAG) {
            const double ar = (double)xa[a].re, ai = (double)xa[a].im;
            xb[a] = {D * ar, D * ai};
            xacc += D * (ar * ar + ai * ai);
          }
        }
        // one amplitude at a time: keeps the scheduler from evaluating all
        // sixteen diagonal values and phases ahead, and the sum from being
        // sunk past the loop (register pressure)
        asm volatile("" : "+v"(xacc));
        __builtin_amdgcn_sched_barrier(0);
      }
      dacc += xacc;
    }
    if (flags & F_CZ_POST) {
#pragma unroll
      for (int a = 0; a < A; ++a) {
        uint32_t x = gst;
#pragma unroll
        for (int k = 0; k < R; ++k)
          if (a & (1 << k)) x |= rbit[k];
        const double sg = cz_sign(x, n);
        if constexpr (!RF) xa[a] = {(TAr)sg * xa[a].re, (TAr)sg * xa[a].im};
        if constexpr (TWO) xb[a] = {sg * xb[a].re, sg * xb[a].im};
      }
    }

    // ---- store: addresses recomputed from the laundered base ----------------
    uint32_t adr[A];
    tile_addresses<R, BIG ? 0 : GSH>(d, (flags & F_SCATTER) != 0, gst, rpos, rbit, n, adr);
    if (!RF && (flags & F_STORE_A)) {
#pragma unroll
      for (int a = 0; a < A; ++a) {
        if (BIG) Stor<G>::st_big(d->a_out, adr[a], xa[a]);
        else Stor<G>::st_o(d->a_out, adr[a], xa[a]);
      }
    }
    if constexpr (TWO) {
      if (flags & F_STORE_B) {
#pragma unroll
        for (int a = 0; a < A; ++a) {
          if (BIG) Stor<G>::st_big(d->b_out, adr[a], xb[a]);
          else Stor<G>::st_o(d->b_out, adr[a], xb[a]);
        }
      }
    }
  }

  // ---- fixed-order reduction: per wave (wave_sum), then the waves in order
  if (!reduce) return;
  if constexpr (ACCL) {
    for (int lb = 0; lb < K; ++lb) {
      if (reduce_bits && d->own[lb]) {
        const double v = wave_sum(tacc[lb * nthr + tid_in], width);
        if (lane0) wacc[wave * (K + 1) + lb] = v;
      }
    }
  }
  if (DSUM && (flags & (F_DRED_POST | F_DRED_PRE | F_OBS_DIAG))) {
    const double v = wave_sum(dacc, width);
    if (lane0) wacc[wave * (K + 1) + K] = v;
  }
  __syncthreads();
  for (uint32_t i = tid_in; i < (uint32_t)(K + 1); i += nthr) {
    int row = -1;
    if (i == (uint32_t)K) {
      if (flags & (F_DRED_POST | F_DRED_PRE | F_OBS_DIAG)) row = d->row_base + K;
    } else if ((flags & F_REDUCE) && d->own[i]) {
      row = d->row_base + (int)i;
    }
    if (row >= 0) {
      double v = 0.0;
      for (int w = 0; w < nwaves; ++w) v += wacc[w * (K + 1) + i];
      d->part[(size_t)row * d->grid + blk] = v;
    }
  }
}

// blockIdx.y selects one of gridDim.y consecutive descriptors: a batched
// launch runs the same pass of gridDim.y independent circuits.
template <class S, int R, bool TWO, int PH, bool POST, bool BIG>
__global__ __launch_bounds__(512) void k_pass(const PassDesc* __restrict__ d) {
  extern __shared__ __align__(16) char lds_dyn[];
  run_pass<S, R, TWO, PH, POST, BIG>(d + blockIdx.y, lds_dyn, threadIdx.x, blockIdx.x, gridDim.x);
}
// 1024-thread workgroups: 14 local bits with 16 amplitudes per thread.
template <class S, int R, bool TWO, int PH, bool POST, bool BIG>
__global__ __launch_bounds__(1024) void k_pass_w(const PassDesc* __restrict__ d) {
  extern __shared__ __align__(16) char lds_dyn[];
  run_pass<S, R, TWO, PH, POST, BIG, true>(d + blockIdx.y, lds_dyn, threadIdx.x, blockIdx.x, gridDim.x);
}
// Variants built with 1024-thread workgroups: complex64 (FP32 tiles),
// 16 amplitudes per thread, phase-free forward and reduce-first adjoint passes.
template <class S, int R, bool TWO, int PH, bool POST>
constexpr bool has_wide() {
  return std::is_same<S, float>::value && R == 4 && PH == 0 && !POST;
}

// LDS bytes of a pass: the tile in the storage type, the per-wave slots and,
// for complex64 adjoint passes, the per-thread accumulators.
// tab_bytes: the staged table of the pass (0 if none).
// dtype: 1 complex64, 0 complex128, 2 complex64 with FP64 LDS tiles.
// wide: a 1024-thread variant (no per-thread accumulators).
__host__ __device__ inline size_t pass_lds_bytes(int dtype, int K, int threads, bool two, int tab_bytes = 0,
                                                 bool wide = false) {
  return (size_t)(1u << K) * (dtype == 1 ? 8 : 16) + (size_t)((threads + 63) / 64) * (kMaxK + 1) * sizeof(double) +
         (two && dtype == 1 && !wide ? (size_t)K * threads * sizeof(double) : 0) + (size_t)tab_bytes;
}
template <class S>
constexpr int kernel_dtype() {
  return std::is_same<S, float>::value ? 1 : std::is_same<S, double>::value ? 0 : 2;
}

// BIG: the state is 4 GiB or larger, so byte offsets need 64-bit addresses.
// With occ != NULL nothing is launched: *occ receives the number of
// workgroups of this instantiation that fit one CU.
template <class S, int R, bool TWO, int PH, bool POST>
hipError_t launch_pass_t(const PassDesc* d, int K, int grid, int batch, int threads, bool big, hipStream_t st,
                         int* occ, int tab_bytes) {
  const size_t lds = pass_lds_bytes(kernel_dtype<S>(), K, threads, TWO, tab_bytes, threads > 512);
  if (threads > 512) {
    if constexpr (has_wide<S, R, TWO, PH, POST>()) {
      if (occ) {
        if (big)
          return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, k_pass_w<S, R, TWO, PH, POST, true>, threads, lds);
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, k_pass_w<S, R, TWO, PH, POST, false>, threads, lds);
      }
      if (big)
        hipLaunchKernelGGL((k_pass_w<S, R, TWO, PH, POST, true>), dim3(grid, batch), dim3(threads), lds, st, d);
      else
        hipLaunchKernelGGL((k_pass_w<S, R, TWO, PH, POST, false>), dim3(grid, batch), dim3(threads), lds, st, d);
      return hipGetLastError();
    }
    return hipErrorInvalidValue;
  }
  if (occ) {
    if (big)
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, k_pass<S, R, TWO, PH, POST, true>, threads, lds);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(occ, k_pass<S, R, TWO, PH, POST, false>, threads, lds);
  }
  if (big)
    hipLaunchKernelGGL((k_pass<S, R, TWO, PH, POST, true>), dim3(grid, batch), dim3(threads), lds, st, d);
  else
    hipLaunchKernelGGL((k_pass<S, R, TWO, PH, POST, false>), dim3(grid, batch), dim3(threads), lds, st, d);
  return hipGetLastError();
}

template <class S, int R, bool TWO, int PH, bool POST>
hipError_t set_attr_t() {
  // the largest tile a 512-thread variant runs (13 local bits), and the LDS
  // of a 1024-thread variant (14 local bits)
  const int maxlds = (int)pass_lds_bytes(kernel_dtype<S>(), 13, 512, TWO, 16 * kLutMax + 8 * (kMaxQubits + 1));
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pass<S, R, TWO, PH, POST, false>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, maxlds);
  if (e == hipSuccess)
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pass<S, R, TWO, PH, POST, true>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, maxlds);
  if constexpr (has_wide<S, R, TWO, PH, POST>()) {
    const int wlds = (int)pass_lds_bytes(kernel_dtype<S>(), kMaxKWide, 1024, TWO, 8 * (kMaxQubits + 1), true);
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pass_w<S, R, TWO, PH, POST, false>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, wlds);
    if (e == hipSuccess)
      e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pass_w<S, R, TWO, PH, POST, true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, wlds);
  }
  return e;
}

}  // namespace qg
