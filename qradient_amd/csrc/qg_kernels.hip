// HIP kernels for the statevector grad_run path on CDNA4 (gfx950).
//
// A statevector of 2^n amplitudes is processed in "passes".  Each pass views
// the index bits as K local bits and n-K tile bits; one workgroup handles one
// tile (2^K amplitudes) at a time:
//
//   natural layout : thread bits = local bits [0, t), register bits = [t, K)
//                    (t = K - R, each thread holds 2^R amplitudes in VGPRs)
//   stage s >= 1   : register bits = local [lo_s, lo_s + R) with
//                    lo_s = max(0, t - R*s); reached by an LDS transpose
//
// Every single-qubit gate of the pass is applied to the register bits of the
// stage in which its bit first appears, so all K gates of a tile are applied
// with ceil(t/R) + 1 LDS round trips and one HBM read + write per state.
// Amplitudes are held in FP64 in registers and LDS whatever the storage type,
// so a complex64 state is rounded once per pass, not once per gate.
//
// The CNOT ladder of the McClean circuit (reference
// qradient/physical_components/state.py:209-251) is a GF(2)-linear map of the
// index whose high output bits depend only on high input bits, so it is
// applied as an address map on the global load (forward) or store (reverse)
// of the low-bit pass: no pass of its own.
//
// Reductions Im<B|P_q|A> (reference grad entries, mc_clean.py:68-76) are
// accumulated in FP64 per thread, reduced per workgroup in a fixed order and
// written to partial rows; k_finalize sums the rows in a fixed order, so
// results are bitwise reproducible for a given plan.
#include <hip/hip_runtime.h>
#include <type_traits>
#include <utility>
#include "qg_internal.h"

namespace qg {

struct Cd {
  double re, im;
};

// Global (address space 1) views: loads and stores through them compile to
// global_load / global_store instead of flat accesses.
typedef float f2v __attribute__((ext_vector_type(2)));
typedef double d2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) f2v gf2v;
typedef __attribute__((address_space(1))) d2v gd2v;

template <class S>
struct Stor;
template <>
struct Stor<float> {
  static __device__ __forceinline__ Cd ld(const void* p, uint32_t i) {
    const f2v v = ((const gf2v*)p)[i];
    return {(double)v.x, (double)v.y};
  }
  static __device__ __forceinline__ void st(void* p, uint32_t i, Cd v) {
    f2v w;
    w.x = (float)v.re;
    w.y = (float)v.im;
    ((gf2v*)p)[i] = w;
  }
};
template <>
struct Stor<double> {
  static __device__ __forceinline__ Cd ld(const void* p, uint32_t i) {
    const d2v v = ((const gd2v*)p)[i];
    return {v.x, v.y};
  }
  static __device__ __forceinline__ void st(void* p, uint32_t i, Cd v) {
    d2v w;
    w.x = v.re;
    w.y = v.im;
    ((gd2v*)p)[i] = w;
  }
};

// LDS slot of local index j: XOR swizzle so that strided stage layouts stay
// free of bank conflicts for 16-byte accesses.
__device__ __forceinline__ uint32_t swz(uint32_t j) { return j ^ ((j >> 4) & 15u); }

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

__device__ __forceinline__ uint32_t apply_map(const uint32_t* __restrict__ m, uint32_t v, int n) {
  uint32_t r = 0;
  for (int k = 0; k < n; ++k) r ^= m[k] & (0u - ((v >> k) & 1u));
  return r;
}

__device__ __forceinline__ double cz_sign(uint32_t x, int n) {
  const uint32_t chain = (n >= 2) ? (uint32_t)((1ull << (n - 1)) - 1ull) : 0u;
  return (__popc(x & (x >> 1) & chain) & 1) ? -1.0 : 1.0;
}

// ---------------------------------------------------------------------------
// 2x2 gates on register bit KK: U = c*I - i*s*P for P = X, Y, Z (reference
// state.py:90-92, 142-144, 168-170).  The planner negates s for inverses.
// ---------------------------------------------------------------------------
template <int R, int KK>
__device__ __forceinline__ void gate_bit(Cd (&x)[1 << R], int type, double c, double s) {
  constexpr int A = 1 << R;
  if (type == G_X) {
#pragma unroll
    for (int a = 0; a < A; ++a) {
      if (a & (1 << KK)) continue;
      const Cd x0 = x[a], x1 = x[a | (1 << KK)];
      // x0' = c x0 - i s x1 ; x1' = c x1 - i s x0
      x[a] = {c * x0.re + s * x1.im, c * x0.im - s * x1.re};
      x[a | (1 << KK)] = {c * x1.re + s * x0.im, c * x1.im - s * x0.re};
    }
  } else if (type == G_Y) {
#pragma unroll
    for (int a = 0; a < A; ++a) {
      if (a & (1 << KK)) continue;
      const Cd x0 = x[a], x1 = x[a | (1 << KK)];
      // -iY = [[0,-1],[1,0]]: x0' = c x0 - s x1 ; x1' = c x1 + s x0
      x[a] = {c * x0.re - s * x1.re, c * x0.im - s * x1.im};
      x[a | (1 << KK)] = {c * x1.re + s * x0.re, c * x1.im + s * x0.im};
    }
  } else if (type == G_Z) {
#pragma unroll
    for (int a = 0; a < A; ++a) {
      const Cd v = x[a];
      // bit 0: (c - i s) v ; bit 1: (c + i s) v
      const double ss = (a & (1 << KK)) ? s : -s;
      x[a] = {c * v.re - ss * v.im, c * v.im + ss * v.re};
    }
  }
}

// sum over the register pairs of bit KK of Im<l|P|p>.
template <int R, int KK>
__device__ __forceinline__ double red_bit(const Cd (&l)[1 << R], const Cd (&p)[1 << R], int type) {
  constexpr int A = 1 << R;
  double acc = 0.0;
  if (type == G_Z) {
#pragma unroll
    for (int a = 0; a < A; ++a) {
      const double v = l[a].re * p[a].im - l[a].im * p[a].re;
      acc += (a & (1 << KK)) ? -v : v;
    }
  } else if (type == G_X) {
#pragma unroll
    for (int a = 0; a < A; ++a) {
      if (a & (1 << KK)) continue;
      const int b = a | (1 << KK);
      acc += (l[a].re * p[b].im - l[a].im * p[b].re) + (l[b].re * p[a].im - l[b].im * p[a].re);
    }
  } else if (type == G_Y) {
#pragma unroll
    for (int a = 0; a < A; ++a) {
      if (a & (1 << KK)) continue;
      const int b = a | (1 << KK);
      acc += -(l[a].re * p[b].re + l[a].im * p[b].im) + (l[b].re * p[a].re + l[b].im * p[a].im);
    }
  }
  return acc;
}

template <int R>
__device__ __forceinline__ void lds_transpose(Cd (&x)[1 << R], double2* lds, uint32_t tid, int lo_from, int lo_to) {
  constexpr int A = 1 << R;
#pragma unroll
  for (int a = 0; a < A; ++a) lds[swz(lidx<R>(tid, a, lo_from))] = make_double2(x[a].re, x[a].im);
  __syncthreads();
#pragma unroll
  for (int a = 0; a < A; ++a) {
    const double2 v = lds[swz(lidx<R>(tid, a, lo_to))];
    x[a] = {v.x, v.y};
  }
  __syncthreads();
}

// D(x) for the 2^R register values of one thread, with the thread's other
// bits fixed: D(a) = c0 + sum_k s_k(a) h_k + sum_{k<l} s_k(a) s_l(a) pw_kl.
template <int R>
struct DiagParts {
  double c0;
  double h[R];
  double pw[R][R];
};

__device__ __forceinline__ double zsgn(uint32_t x, int pos) { return ((x >> pos) & 1u) ? -1.0 : 1.0; }

template <int R>
__device__ __forceinline__ void diag_prepare(const PassDesc* __restrict__ d, uint32_t gthr, const int (&rpos)[R],
                                             DiagParts<R>& o) {
  o.c0 = 0.0;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    o.h[k] = 0.0;
#pragma unroll
    for (int l = 0; l < R; ++l) o.pw[k][l] = 0.0;
  }
  const int nt = d->ndterms;
  for (int i = 0; i < nt; ++i) {
    const int a = d->dterms[i].a, b = d->dterms[i].b;
    const double w = d->dterms[i].w;
    int ka = -1, kb = -1;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      if (rpos[k] == a) ka = k;
      if (b >= 0 && rpos[k] == b) kb = k;
    }
    const double za = (ka >= 0) ? 1.0 : zsgn(gthr, a);
    const double zb = (b < 0 || kb >= 0) ? 1.0 : zsgn(gthr, b);
    if (ka < 0 && kb < 0) {
      o.c0 += w * za * zb;
    } else if (ka >= 0 && kb < 0) {
#pragma unroll
      for (int k = 0; k < R; ++k)
        if (ka == k) o.h[k] += w * zb;
    } else if (ka < 0 && kb >= 0) {
#pragma unroll
      for (int k = 0; k < R; ++k)
        if (kb == k) o.h[k] += w * za;
    } else {
#pragma unroll
      for (int k = 0; k < R; ++k)
#pragma unroll
        for (int l = k + 1; l < R; ++l)
          if ((ka == k && kb == l) || (ka == l && kb == k)) o.pw[k][l] += w;
    }
  }
}

template <int R>
__device__ __forceinline__ double diag_value(const DiagParts<R>& o, int a) {
  double v = o.c0;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const double sk = ((a >> k) & 1) ? -1.0 : 1.0;
    v += sk * o.h[k];
#pragma unroll
    for (int l = k + 1; l < R; ++l) {
      const double sl = ((a >> l) & 1) ? -1.0 : 1.0;
      v += sk * sl * o.pw[k][l];
    }
  }
  return v;
}

// exp(-i * gamma * D): PH = 1 reads a table (integer-valued D), PH = 2 evaluates sincos.
template <int PH>
__device__ __forceinline__ Cd phase_of(const PassDesc* __restrict__ d, double D) {
  if constexpr (PH == 1) {
    const int idx = (int)rint(D) + d->lut_off;
    return {d->lut[2 * idx], d->lut[2 * idx + 1]};
  } else {
    double s, c;
    sincos(d->phase_gamma * D, &s, &c);
    return {c, -s};
  }
}

__device__ __forceinline__ Cd cmul(Cd a, Cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

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

__device__ __forceinline__ double wave_sum(double v, int width) {
  for (int off = width >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, width);
  return v;
}

// ---------------------------------------------------------------------------
// One pass over the tiles assigned to this workgroup.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int nnew_of(const int* los, int s, int R) {
  if (s == 0) return R;
  const int d = los[s - 1] - los[s];
  return d < R ? d : R;
}

template <class S, int R, bool TWO, int PH>
__device__ __forceinline__ void run_pass(const PassDesc* __restrict__ d, double2* lds, uint32_t tid_in,
                                         uint32_t blk, uint32_t nblk) {
  constexpr int A = 1 << R;
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

  int rpos[R];
  uint32_t rbit[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {
    rpos[k] = gpos_of_local(t + k, c, p);
    rbit[k] = 1u << rpos[k];
  }
  const bool use_map = (flags & (F_GATHER | F_SCATTER)) != 0;

  // Per-wave FP64 accumulators in LDS, after the tile buffer: [wave][K + 1].
  // Each (wave, slot) has exactly one writer (lane 0), so the summation order
  // is fixed.
  double* wacc = reinterpret_cast<double*>(lds + (1u << K));
  const bool reduce = TWO && (flags & (F_REDUCE | F_DRED_POST | F_OBS_DIAG)) && d->row_base >= 0;
  if (reduce) {
    for (uint32_t i = tid_in; i < (uint32_t)(nwaves * (K + 1)); i += nthr) wacc[i] = 0.0;
    __syncthreads();
  }

  // XCD-aware tile order: workgroup blk runs on XCD blk % 8; give each XCD a
  // contiguous range of tiles so tiles that share cache lines share an L2.
  const bool xcd = (nblk >= 8) && (nblk % 8 == 0) && (ntiles % nblk == 0);
  const uint32_t per_x = ntiles / 8, slots = nblk / 8;
  const uint32_t rounds = (ntiles + nblk - 1) / nblk;

  for (uint32_t r = 0; r < rounds; ++r) {
    // Recompute every thread-index-derived address inside the loop rather
    // than hoisting loop-invariant LDS/global offsets into registers.
    uint32_t tid = tid_in;
    asm volatile("" : "+v"(tid));
    uint32_t tt;
    if (xcd) {
      tt = (blk % 8) * per_x + r * slots + blk / 8;
    } else {
      tt = blk + r * nblk;
      if (tt >= ntiles) break;
    }
    const uint32_t gthr = gl_tile(tt, c, p, K) | gl_local(tid, c, p);

    // ---- addresses and loads ------------------------------------------------
    uint32_t adr[A];
    if (flags & F_GATHER) {
      const uint32_t mbase = apply_map(d->map, gthr, n);
#pragma unroll
      for (int a = 0; a < A; ++a) {
        uint32_t v = mbase;
#pragma unroll
        for (int k = 0; k < R; ++k)
          if (a & (1 << k)) v ^= d->map[rpos[k]];
        adr[a] = v;
      }
    } else {
#pragma unroll
      for (int a = 0; a < A; ++a) {
        uint32_t v = gthr;
#pragma unroll
        for (int k = 0; k < R; ++k)
          if (a & (1 << k)) v |= rbit[k];
        adr[a] = v;
      }
    }
    Cd xa[A], xb[A];
    if (flags & F_INIT_TAB) {
#pragma unroll
      for (int a = 0; a < A; ++a) xa[a] = {d->init_tab[__popc(adr[a])], 0.0};
    } else {
#pragma unroll
      for (int a = 0; a < A; ++a) xa[a] = Stor<S>::ld(d->a_in, adr[a]);
    }
    if (TWO) {
      if (flags & F_LOAD_B) {
#pragma unroll
        for (int a = 0; a < A; ++a) xb[a] = Stor<S>::ld(d->b_in, adr[a]);
      } else {
#pragma unroll
        for (int a = 0; a < A; ++a) xb[a] = {0.0, 0.0};
      }
    }

    // ---- diagonal ops in the natural layout, before the gates --------------
    if (flags & F_CZ_PRE) {
#pragma unroll
      for (int a = 0; a < A; ++a) {
        uint32_t x = gthr;
#pragma unroll
        for (int k = 0; k < R; ++k)
          if (a & (1 << k)) x |= rbit[k];
        const double sg = cz_sign(x, n);
        xa[a] = {sg * xa[a].re, sg * xa[a].im};
      }
    }
    if constexpr (PH != 0) {
      if (flags & F_PHASE_PRE) {
        DiagParts<R> dp;
        diag_prepare<R>(d, gthr, rpos, dp);
#pragma unroll
        for (int a = 0; a < A; ++a) xa[a] = cmul(phase_of<PH>(d, diag_value<R>(dp, a)), xa[a]);
      }
    }

    // ---- stages: reductions then gates on each newly visible register bit --
#pragma unroll 1
    for (int s = 0; s < ns; ++s) {
      if (s > 0) {
        lds_transpose<R>(xa, lds, tid, los[s - 1], los[s]);
        if (TWO) lds_transpose<R>(xb, lds, tid, los[s - 1], los[s]);
      }
      const int lo = los[s];
      const int nnew = nnew_of(los, s, R);
      static_for<R>([&](auto kc) __attribute__((always_inline)) {
        constexpr int kk = decltype(kc)::value;
        if (kk < nnew) {
          // opaque to the optimiser: keeps the uniform gate loads next to
          // their use instead of hoisting them out of the loops
          int lb = __builtin_amdgcn_readfirstlane(lo + kk);
          asm volatile("" : "+s"(lb));
          const int ty = d->gtype[lb];
          if (TWO && (flags & F_REDUCE) && d->own[lb] && d->row_base >= 0) {
            const double v = wave_sum(red_bit<R, kk>(xb, xa, ty), width);
            if ((tid & 63) == 0) wacc[(tid >> 6) * (K + 1) + lb] += v;
          }
          if (flags & F_ROT_A) gate_bit<R, kk>(xa, ty, d->gc[lb], d->gs[lb]);
          if (TWO && (flags & F_ROT_B)) gate_bit<R, kk>(xb, ty, d->gc[lb], d->gs[lb]);
        }
      });
    }
    if (ns > 1) {
      lds_transpose<R>(xa, lds, tid, los[ns - 1], t);
      if (TWO) lds_transpose<R>(xb, lds, tid, los[ns - 1], t);
    }

    // ---- diagonal ops in the natural layout, after the gates --------------
    if (flags & (F_DRED_POST | F_PHASE_POST | F_OBS_DIAG)) {
      DiagParts<R> dp;
      diag_prepare<R>(d, gthr, rpos, dp);
      double xacc = 0.0;
#pragma unroll
      for (int a = 0; a < A; ++a) {
        const double D = diag_value<R>(dp, a);
        if (TWO && (flags & F_DRED_POST)) xacc += 2.0 * D * (xb[a].re * xa[a].im - xb[a].im * xa[a].re);
        if constexpr (PH != 0) {
          if (flags & F_PHASE_POST) {
            const Cd ph = phase_of<PH>(d, D);
            xa[a] = cmul(ph, xa[a]);
            if (TWO) xb[a] = cmul(ph, xb[a]);
          }
        }
        if (TWO && (flags & F_OBS_DIAG)) {
          xb[a] = {D * xa[a].re, D * xa[a].im};
          xacc += D * (xa[a].re * xa[a].re + xa[a].im * xa[a].im);
        }
      }
      if (reduce && (flags & (F_DRED_POST | F_OBS_DIAG))) {
        const double v = wave_sum(xacc, width);
        if ((tid & 63) == 0) wacc[(tid >> 6) * (K + 1) + K] += v;
      }
    }
    if (flags & F_CZ_POST) {
#pragma unroll
      for (int a = 0; a < A; ++a) {
        uint32_t x = gthr;
#pragma unroll
        for (int k = 0; k < R; ++k)
          if (a & (1 << k)) x |= rbit[k];
        const double sg = cz_sign(x, n);
        xa[a] = {sg * xa[a].re, sg * xa[a].im};
        if (TWO) xb[a] = {sg * xb[a].re, sg * xb[a].im};
      }
    }

    // ---- store: addresses recomputed from a laundered base, not kept live ----
    uint32_t gst = gthr;
    asm volatile("" : "+v"(gst));
    if (flags & F_SCATTER) {
      const uint32_t mbase = apply_map(d->map, gst, n);
#pragma unroll
      for (int a = 0; a < A; ++a) {
        uint32_t v = mbase;
#pragma unroll
        for (int k = 0; k < R; ++k)
          if (a & (1 << k)) v ^= d->map[rpos[k]];
        adr[a] = v;
      }
    } else {
#pragma unroll
      for (int a = 0; a < A; ++a) {
        uint32_t v = gst;
#pragma unroll
        for (int k = 0; k < R; ++k)
          if (a & (1 << k)) v |= rbit[k];
        adr[a] = v;
      }
    }
    if (flags & F_STORE_A) {
#pragma unroll
      for (int a = 0; a < A; ++a) Stor<S>::st(d->a_out, adr[a], xa[a]);
    }
    if (TWO && (flags & F_STORE_B)) {
#pragma unroll
      for (int a = 0; a < A; ++a) Stor<S>::st(d->b_out, adr[a], xb[a]);
    }
  }

  // ---- fixed-order workgroup reduction: sum the per-wave slots in order -------
  if (!reduce) return;
  __syncthreads();
  const uint32_t tid = tid_in;
  for (uint32_t i = tid; i < (uint32_t)(K + 1); i += nthr) {
    int row = -1;
    if (i == (uint32_t)K) {
      if (flags & (F_DRED_POST | F_OBS_DIAG)) row = d->row_base + K;
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

template <class S, int R, bool TWO, int PH>
__global__ __launch_bounds__(512) void k_pass(const PassDesc* __restrict__ d) {
  extern __shared__ double2 lds_dyn[];
  run_pass<S, R, TWO, PH>(d, lds_dyn, threadIdx.x, blockIdx.x, gridDim.x);
}

// out[i] = scale[i] * sum_{r in rows[ptr[i]..ptr[i+1])} sum_{w < cnt[r]} part[r*stride + w]
__global__ void k_finalize(int nout, const int* __restrict__ ptr, const int* __restrict__ rows,
                           const int* __restrict__ cnt, const double* __restrict__ scale,
                           const double* __restrict__ part, int stride, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nout) return;
  double s = 0.0;
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    const int r = rows[k];
    double rs = 0.0;
    for (int w = 0; w < cnt[r]; ++w) rs += part[(size_t)r * stride + w];
    s += rs;
  }
  out[i] = scale[i] * s;
}

}  // namespace qg

// ---------------------------------------------------------------------------
// launch wrappers used by the host planner
// ---------------------------------------------------------------------------
namespace qg {

// per-wave accumulator slots appended to the tile buffer: [wave][kMaxK + 1]
static inline size_t lds_acc_bytes(int threads) { return (size_t)((threads + 63) / 64) * (kMaxK + 1) * sizeof(double); }

template <class S, int R, bool TWO, int PH>
static hipError_t launch_pass_t(const PassDesc* d, int grid, int threads, hipStream_t st) {
  const size_t lds = (size_t)threads * (1u << R) * sizeof(double2) + lds_acc_bytes(threads);
  hipLaunchKernelGGL((k_pass<S, R, TWO, PH>), dim3(grid), dim3(threads), lds, st, d);
  return hipGetLastError();
}

template <class S, int R, bool TWO, int PH>
static hipError_t set_attr_t() {
  const int maxlds = (1 << kMaxK) * (int)sizeof(double2) + (int)lds_acc_bytes(1 << (kMaxK - kRegBits));
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pass<S, R, TWO, PH>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, maxlds);
}

hipError_t set_pass_attributes() {
  hipError_t e = hipSuccess;
#define QG_ATTR(S, PH)                                         \
  if (e == hipSuccess) e = set_attr_t<S, 4, false, PH>();      \
  if (e == hipSuccess) e = set_attr_t<S, 4, true, PH>();
  QG_ATTR(float, 0) QG_ATTR(float, 1) QG_ATTR(float, 2) QG_ATTR(double, 0) QG_ATTR(double, 1) QG_ATTR(double, 2)
#undef QG_ATTR
  return e;
}

// dtype: 0 complex128, 1 complex64.  reg_bits = R (4 normally; K for K < 4).
// phase: 0 none, 1 table, 2 sincos.
template <class S, int R>
static hipError_t launch_r(bool two, int phase, const PassDesc* d, int grid, int threads, hipStream_t st) {
  if (phase == 0) return two ? launch_pass_t<S, R, true, 0>(d, grid, threads, st) : launch_pass_t<S, R, false, 0>(d, grid, threads, st);
  if (phase == 1) return two ? launch_pass_t<S, R, true, 1>(d, grid, threads, st) : launch_pass_t<S, R, false, 1>(d, grid, threads, st);
  return two ? launch_pass_t<S, R, true, 2>(d, grid, threads, st) : launch_pass_t<S, R, false, 2>(d, grid, threads, st);
}

hipError_t launch_pass(int dtype, int reg_bits, bool two, int phase, const PassDesc* d, int grid, int threads,
                       hipStream_t st) {
  if (dtype == 1) {
    switch (reg_bits) {
      case 4: return launch_r<float, 4>(two, phase, d, grid, threads, st);
      case 3: return launch_r<float, 3>(two, phase, d, grid, threads, st);
      case 2: return launch_r<float, 2>(two, phase, d, grid, threads, st);
      case 1: return launch_r<float, 1>(two, phase, d, grid, threads, st);
    }
  } else {
    switch (reg_bits) {
      case 4: return launch_r<double, 4>(two, phase, d, grid, threads, st);
      case 3: return launch_r<double, 3>(two, phase, d, grid, threads, st);
      case 2: return launch_r<double, 2>(two, phase, d, grid, threads, st);
      case 1: return launch_r<double, 1>(two, phase, d, grid, threads, st);
    }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_finalize(int nout, const int* ptr, const int* rows, const int* cnt, const double* scale,
                           const double* part, int stride, double* out, hipStream_t st) {
  const int threads = 64;
  hipLaunchKernelGGL(k_finalize, dim3((nout + threads - 1) / threads), dim3(threads), 0, st, nout, ptr, rows,
                     cnt, scale, part, stride, out);
  return hipGetLastError();
}

}  // namespace qg
