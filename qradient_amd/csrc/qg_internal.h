// Internal descriptors shared by the host planner (qg_api.cpp) and the HIP
// kernels (qg_pass.hpp and the files that instantiate it).  Plain-old-data
// only: the planner writes arrays of these structs into device memory and
// every kernel reads them through a pointer, so a captured hipGraph stays
// valid while the per-call contents (angles, observable weights) change.
#pragma once
#include <stdint.h>

namespace qg {

constexpr int kMaxK = 14;        // max local (tile) bits of a pass (14: complex64 only, 2^14 x 8 B = 128 KiB LDS)
constexpr int kMaxKWide = 14;    // local bits that need 1024-thread workgroups (complex64 history passes)
constexpr int kFusedMaxN = 13;   // largest circuit run whole in one workgroup (2^13 FP64 complex = 128 KiB LDS)
constexpr int kRegBits = 4;      // amplitudes held per thread = 2^kRegBits
constexpr int kMaxStages = 4;    // natural stage + ceil((kMaxK - kRegBits) / kRegBits)
constexpr int kMaxQubits = 32;     // 32-bit amplitude indices
constexpr int kMaxR = 5;           // log2 amplitudes per thread
constexpr int kMaxSlots = kMaxStages * kMaxR;
enum : int32_t { SL_ROT_A = 1, SL_ROT_B = 2, SL_RED = 4 };
constexpr int kLutMax = 1024;      // phase-table entries staged in LDS (2W + 1 <= kLutMax)
// Device scratch of the allocation-free whole-circuit entry point (one static
// array per GPU; equals QG_NOALLOC_SCRATCH_BYTES in the public header).
constexpr unsigned long kNoAllocScratchBytes = 256ul * 1024ul;
constexpr int kNoAllocMaxDevices = 64;

enum GateType : int32_t { G_NONE = 0, G_X = 1, G_Y = 2, G_Z = 3 };

// Pass flags.
enum : uint32_t {
  F_LOAD_A      = 1u << 0,   // A is read from a_in
  F_INIT_TAB    = 1u << 1,   // A is generated: amp(x) = init_tab[popcount(x)]
  F_LOAD_B      = 1u << 2,   // B (the adjoint state lambda) is read from b_in
  F_ROT_A       = 1u << 3,   // apply the pass gates to A
  F_ROT_B       = 1u << 4,   // apply the pass gates to B
  F_STORE_A     = 1u << 5,
  F_STORE_B     = 1u << 6,
  F_REDUCE      = 1u << 7,   // per owned bit: Im<B|P_q|A> before the gates
  F_GATHER      = 1u << 8,   // load address = map(x) (ladder gather, forward)
  F_SCATTER     = 1u << 9,   // store address = map(x) (ladder scatter, reverse)
  F_CZ_PRE      = 1u << 10,  // multiply by the CZ-ladder sign after load
  F_CZ_POST     = 1u << 11,  // multiply by the CZ-ladder sign before store
  F_PHASE_PRE   = 1u << 12,  // A (and B) *= exp(-i*phase_gamma*D(x)) after load (QAOA cost phase)
  F_PHASE_POST  = 1u << 13,  // A (and B) *= exp(-i*phase_gamma*D(x)) before store
  F_DRED_POST   = 1u << 14,  // extra += 2*sum D(x) Im(conj(B)A) before the post phase (dE/dgamma)
  F_OBS_DIAG    = 1u << 15,  // B = D(x)*A at the end; extra += D(x)|A|^2 (lambda = O psi, E)
  F_PHASE_LUT   = 1u << 16,  // D(x) is integer valued: phase from lut[D + lut_off]
  F_DRED_PRE    = 1u << 17,  // (adjoint, with F_PHASE_PRE) extra += 2*sum D(x) Im(conj(B)A) before the pre phase
  F_XCD_ORDER   = 1u << 18,  // XCD-aware tile order (else round-robin); set by the planner
  F_NT_A        = 1u << 19,  // A is read with non-temporal loads (read once in this sweep; set by the planner)
};

// One term of the diagonal quadratic form D(x) = sum w * z_a(x) [* z_b(x)],
// with z_p(x) = +1 if global bit p of x is 0 else -1.  b = -1 for a one-bit
// term.  (In PassDesc::cterms, a holds a bit mask instead; see there.)
struct DTerm {
  int32_t a, b;
  double w;
};

// One Pauli term whose expectation value <P> is measured (finite-shot
// sampling): kind 0 = Z_a (b < 0) or Z_a Z_b, 1 = X_a, 2 = Y_a, 3 = the
// identity (the squared norm); a, b are global bit positions.
struct PTerm {
  int32_t kind, a, b, pad;
};
constexpr int kTermsPerLaunch = 16;  // terms measured per pass over the state (general kernel)
constexpr int kDiagTermsPerLaunch = 24;  // diagonal (Z / ZZ / identity) terms per pass of the mask kernel
constexpr int kTermPartRows = kDiagTermsPerLaunch;  // partial rows a term launch may write
constexpr int kTermBlocks = 2048;    // workgroups of a term pass (partials per term)
constexpr int kSampleBlock = 1024;   // positions per block of the basis sampler

struct PassDesc {
  // geometry: local bits = [0, c) U [p, p + K - c); tile bits = the rest, in order
  int32_t n, K, t, c, p;
  uint32_t flags;
  int32_t grid;          // stride of the partial rows (>= number of workgroups)
  int32_t row_base;      // first partial row written by this pass (-1: none)
  const void* a_in;
  void* a_out;
  const void* b_in;
  void* b_out;
  // gates, indexed by local bit: U = cos(phi) I - i sin(phi) P with the
  // half-angle phi reduced to [-pi/2, pi/2] (a global sign of -1 per reduced
  // gate is dropped: it cancels in every expectation and gradient), stored as
  // ga = tan(phi / 2), gb = sin(phi) for the three-shear rotation.  Inverse
  // gates have phi negated.  G_NONE: no gate on that bit.
  double ga[kMaxK];
  double gb[kMaxK];
  int32_t gtype[kMaxK];
  int32_t own[kMaxK];    // 1: accumulate the reduction of this local bit into row_base + bit
  // The same gates per stage slot (stage s, register bit kk -> slot s * R + kk),
  // filled by the host from the per-bit arrays for the launch's R: a slot whose
  // bit was handled in an earlier stage is G_NONE.  sflag: SL_ROT_A / SL_ROT_B
  // (apply to A / B) and SL_RED (reduce this slot's bit).
  double sa[kMaxSlots];
  double sb[kMaxSlots];
  int32_t stype[kMaxSlots];
  int32_t sflag[kMaxSlots];
  int32_t slb[kMaxSlots];  // local bit of the slot
  int32_t nstage;          // stages that hold a gate or a reduction (later stages are skipped)
  // reduce-first adjoint passes: the owned thread bits (local bits below t)
  // by gate type, as masks of local bits
  uint32_t xmask, ymask, zmask;
  // the ladder map (F_GATHER / F_SCATTER) in shift form:
  // map(v) = v ^ ((v & lad1) >> 1) ^ ((v & lad2) >> 2)
  uint32_t lad1, lad2;
  const double* init_tab;// n + 1 entries when F_INIT_TAB
  const DTerm* dterms;   // diagonal quadratic form (global bit positions a, b)
  int32_t ndterms;
  // The same terms split by the launch's register bits (filled by the host):
  // cterms[0, nct) have no register bit, then nht[k] terms pair register bit
  // k with another bit; each as (a = mask of its other bits, w).  pwc[a]: the
  // terms on register bits only, at register pattern a.
  const DTerm* cterms;
  int32_t nct;
  int32_t nht[kMaxR];
  double pwc[1 << kMaxR];
  int32_t lut_off;       // LUT index offset (F_PHASE_LUT)
  double phase_gamma;    // angle multiplying D(x) in the phase
  const double* lut;     // (re, im) pairs: exp(-i*gamma*d) at d = idx - lut_off
  double* part;          // partial sums: part[row * grid + block]
};

}  // namespace qg
