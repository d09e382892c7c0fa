// Host side of the qradient_amd C ABI: contexts, plans (pass schedules,
// workspace layout, descriptor tables), the grad_run drivers and the
// device-resident State object.  See include/qradient_amd.h for the contract.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <limits>
#include <vector>

#include "../../include/qradient_amd.h"
#include "qg_internal.h"

namespace qg {
hipError_t set_pass_attributes();
hipError_t launch_pass(int dtype, int R, bool two, int phase, bool post, const PassDesc* d, int K, int grid,
                       int batch, int threads, int n, hipStream_t st, int* occ, int tab);
hipError_t launch_finalize(int nout, const int* ptr, const int* rows, const int* cnt, const double* scale,
                           const double* part, int stride, double* out, hipStream_t st);
hipError_t simple_rot1(int dtype, void* v, int n, int b, int type, double c, double s, hipStream_t st);
hipError_t simple_permute(int dtype, const void* in, void* out, int n, const uint32_t* dmap, hipStream_t st);
hipError_t simple_diag_phase(int dtype, void* v, int n, const DTerm* dt, int nt, double angle, int cz,
                             hipStream_t st);
hipError_t simple_obs_apply(int dtype, const void* in, void* out, int n, const DTerm* dt, int ndt, const DTerm* xy,
                            int nxy, hipStream_t st);
hipError_t simple_dot(int dtype, const void* a, const void* b, int n, int mode, double* part, double* out,
                      hipStream_t st);
hipError_t simple_init(int dtype, void* v, int n, const double* tab, hipStream_t st);
hipError_t simple_scale(int dtype, void* v, int n, double re, double im, hipStream_t st);
hipError_t simple_convert_in(int dtype, const void* in_c128, void* out, int n, hipStream_t st);
hipError_t simple_convert_out(int dtype, const void* in, void* out_c128, int n, hipStream_t st);
hipError_t simple_term_expec(int dtype, const void* v, int n, const PTerm* t, int nt, double* part, double* out,
                             hipStream_t st, const PTerm* h_t);
hipError_t simple_sample_basis(int dtype, const void* v, int n, const int32_t* order, const double* u, int shots,
                               int32_t* idx, double* mass, hipStream_t st);
constexpr int kDotBlocks = 512;
hipError_t launch_fused(const void* descs, int batch, int n, hipStream_t st);
hipError_t set_fused_attributes();
size_t fused_desc_size();
hipError_t noalloc_scratch_init(int device);
hipError_t launch_stream_copy(void* dst, const void* src, size_t bytes, bool nontemporal, int grid, hipStream_t st);
void* noalloc_scratch(int device);

// Mirror of the device-side FusedDesc in qg_fused.hip (same layout).
struct FusedDescH {
  int32_t n, L, kind, ent;
  const int32_t* gtype;
  const double* gcs;
  const double* params;
  const DTerm* dterms;
  int32_t ndterms;
  const DTerm* xy;
  int32_t nxy;
  uint32_t lad1, lad2;
  const double* init_tab;
  double* out;
  int32_t grad;
  int32_t psi_c64;
  void* psi_out;
};
}  // namespace qg

using namespace qg;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(call)                                                                                    \
  do {                                                                                                  \
    hipError_t e_ = (call);                                                                             \
    if (e_ != hipSuccess) {                                                                             \
      return fail(e_ == hipErrorOutOfMemory ? QG_ERR_NOMEM : QG_ERR_HIP,                                \
                  std::string(#call) + ": " + hipGetErrorString(e_));                                   \
    }                                                                                                   \
  } while (0)

#define CHK(call)                  \
  do {                             \
    int r_ = (call);               \
    if (r_ != QG_OK) return r_;    \
  } while (0)

namespace qg {
// The calling thread's qg_last_error, for the other host translation units.
int set_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace qg

extern "C" const char* qg_last_error(void) { return g_err.c_str(); }
extern "C" int qg_version(void) { return 1; }

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct qg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int num_cu = 256;
  size_t max_lds = 65536;
  size_t total_mem = 0;  // device memory, queried once (plan layouts are sized against it)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // pinned host mirror of the device scratch of qg_mcclean_grad_run_noalloc
  // (tables staged here, copied with one H2D; results copied back here)
  unsigned char* noalloc_host = nullptr;
};

extern "C" int qg_ctx_create(int device, void* stream, qg_ctx** out) {
  if (!out) return fail(QG_ERR_INVALID, "qg_ctx_create: out is NULL");
  int count = 0;
  HIPCHK(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(QG_ERR_INVALID, "qg_ctx_create: invalid device index");
  HIPCHK(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
    return fail(QG_ERR_UNSUPPORTED, std::string("qradient_amd is built for gfx950, device is ") + prop.gcnArchName);
  qg_ctx* c = new qg_ctx();
  c->device = device;
  c->num_cu = prop.multiProcessorCount;
  c->max_lds = prop.sharedMemPerBlock;
  {
    size_t free_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &c->total_mem));
  }
  if (stream) {
    c->stream = (hipStream_t)stream;
  } else {
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      return fail(QG_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    c->own_stream = true;
  }
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      set_pass_attributes() != hipSuccess || set_fused_attributes() != hipSuccess ||
      fused_desc_size() != sizeof(FusedDescH) ||
      hipHostMalloc((void**)&c->noalloc_host, kNoAllocScratchBytes, hipHostMallocDefault) != hipSuccess) {
    if (c->noalloc_host) (void)hipHostFree(c->noalloc_host);
    delete c;
    return fail(QG_ERR_HIP, "qg_ctx_create: event creation or kernel attribute setup failed");
  }
  // the allocation-free entry point's device scratch: its address on this
  // device, taken once per device with the device current
  {
    const hipError_t e = noalloc_scratch_init(device);
    if (e != hipSuccess) {
      (void)hipHostFree(c->noalloc_host);
      delete c;
      return fail(QG_ERR_HIP, std::string("qg_ctx_create: scratch address: ") + hipGetErrorString(e));
    }
  }
  *out = c;
  return QG_OK;
}

extern "C" int qg_ctx_destroy(qg_ctx* c) {
  if (!c) return QG_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  if (c->noalloc_host) (void)hipHostFree(c->noalloc_host);
  delete c;
  return QG_OK;
}

extern "C" int qg_stream_copy(qg_ctx* c, void* dst, const void* src, size_t bytes, int nontemporal,
                              int grid_per_cu, int reps, double* ms_out) {
  if (!c || !dst || !src || !ms_out) return fail(QG_ERR_INVALID, "NULL argument");
  if (bytes == 0 || bytes % 16 != 0 || ((uintptr_t)dst | (uintptr_t)src) % 16 != 0)
    return fail(QG_ERR_INVALID, "qg_stream_copy: bytes and both pointers must be multiples of 16");
  if (reps < 1 || grid_per_cu < 1 || grid_per_cu > 64)
    return fail(QG_ERR_INVALID, "qg_stream_copy: reps >= 1 and grid_per_cu in [1, 64]");
  HIPCHK(hipSetDevice(c->device));
  const int grid = c->num_cu * grid_per_cu;
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  for (int r = 0; r < reps; ++r) HIPCHK(launch_stream_copy(dst, src, bytes, nontemporal != 0, grid, c->stream));
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *ms_out = (double)ms / reps;
  return QG_OK;
}

extern "C" int qg_ctx_synchronize(qg_ctx* c) {
  if (!c) return fail(QG_ERR_INVALID, "NULL context");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  return QG_OK;
}

// ---------------------------------------------------------------------------
// observables
// ---------------------------------------------------------------------------
struct ObsSplit {
  std::vector<DTerm> diag;  // z / zz with global bit positions
  std::vector<DTerm> xy;    // a = global bit, b = 0 (X) or 1 (Y)
};

static int split_observable(int n, const qg_term* terms, int nterms, ObsSplit& o) {
  if (nterms < 0 || (nterms > 0 && !terms)) return fail(QG_ERR_INVALID, "observable: bad term array");
  for (int k = 0; k < nterms; ++k) {
    const qg_term& t = terms[k];
    if (t.i < 0 || t.i >= n) return fail(QG_ERR_INVALID, "observable: qubit index out of range");
    if (!std::isfinite(t.w)) return fail(QG_ERR_INVALID, "observable: non-finite weight");
    const int bi = n - 1 - t.i;
    switch (t.kind) {
      case QG_TERM_Z: o.diag.push_back({bi, -1, t.w}); break;
      case QG_TERM_ZZ:
        if (t.j <= t.i || t.j >= n)
          return fail(QG_ERR_INVALID, "observable: zz terms need i < j < n (upper triangle, observable.py:56-64)");
        o.diag.push_back({bi, n - 1 - t.j, t.w});
        break;
      case QG_TERM_X: o.xy.push_back({bi, 0, t.w}); break;
      case QG_TERM_Y: o.xy.push_back({bi, 1, t.w}); break;
      default: return fail(QG_ERR_INVALID, "observable: unknown term kind");
    }
  }
  return QG_OK;
}

// ---------------------------------------------------------------------------
// pass geometry
// ---------------------------------------------------------------------------
struct Geo {
  int K, t, c, p, R, threads;
  int own_lo, own_hi;  // owned local bits [own_lo, own_hi)
};

static Geo make_geo(int K, int c, int p, int own_lo, int own_hi) {
  Geo g;
  g.K = K;
  g.c = c;
  g.p = p;
  g.R = std::min(kRegBits, K);
  g.t = K - g.R;
  g.threads = 1 << g.t;
  g.own_lo = own_lo;
  g.own_hi = own_hi;
  return g;
}

// Passes covering all n bits once: a low pass with local bits [0, Kl) and high
// passes with local bits [0, c) U [p, p + K - c) that own [p, p + K - c).
static std::vector<Geo> layer_geometry(int n, int kmax, int cmin) {
  std::vector<Geo> g;
  if (n <= kmax) {
    g.push_back(make_geo(n, n, n, 0, n));
    return g;
  }
  g.push_back(make_geo(kmax, kmax, n, 0, kmax));
  const int rest = n - kmax, wmax = kmax - cmin;
  const int h = (rest + wmax - 1) / wmax;
  int start = kmax;
  for (int i = 0; i < h; ++i) {
    const int w = (rest - (start - kmax) + (h - i) - 1) / (h - i);
    const int c = kmax - w;
    g.push_back(make_geo(kmax, c, start, c, kmax));
    start += w;
  }
  return g;
}

// Default local bits per pass: the fewest passes per layer that 13 local bits
// allow, with the smallest tile that keeps that count (at least 11 local bits
// for complex64, 12 for complex128).  Smaller tiles mean smaller workgroups
// (2^(K-4) threads): more of them are resident per CU, so one workgroup's
// barrier waits overlap another's work.
// complex64 passes may take 14 local bits (1024-thread workgroups, phase-free
// forward and reduce-first adjoint passes only) when that saves a pass per
// layer (wide = true: the caller can run such passes).  Not the default: one
// 1024-thread workgroup per CU (128 KiB of LDS) runs the two-state pass at
// about 40% of the copy bandwidth, so two 14-bit passes take as long as three
// 11-bit ones; qg_plan_tune and the autotuner can still select it.
static int passes_per_layer(int n, int K, int cmin) {
  return n <= K ? 1 : 1 + (n - K + (K - cmin) - 1) / (K - cmin);
}
// complex64 states of 4 GiB and more (n >= 29) take 13 local bits: the high
// passes then keep 4 and 5 contiguous low bits (128-256 B runs) where 12 local
// bits leave 3 (64 B runs, 8-16 MiB apart), which streamed at 3.1 TB/s against
// 4.4 TB/s.  Measured at 30 qubits, 100 layers, checkpoints every 5 layers:
// 4.89 s against 5.48 s, the same gradient error.
static int default_kmax(int n, int dtype, int cmin, bool wide = false) {
  const int kmin = dtype == QG_C64 ? 11 : 12;
  const int best = passes_per_layer(n, 13, cmin);
  if (wide && dtype == QG_C64 && n > kFusedMaxN && passes_per_layer(n, kMaxKWide, cmin) < best) return kMaxKWide;
  if (dtype == QG_C64 && n >= 29) return 13;
  for (int K = kmin; K < 13; ++K)
    if (passes_per_layer(n, K, cmin) == best) return K;
  return 13;
}

// Register bits of a pass by role: complex64 forward and history-adjoint
// passes hold 32 amplitudes per thread (the adjoint's read-only companion in
// FP32 registers); passes that propagate two FP64 states hold 16.
enum Role { ROLE_FWD = 0, ROLE_REV_HIST = 1, ROLE_TWO_FP64 = 2 };

static Geo with_role_r(const Geo& g0, int dtype, int role, int r_fwd, int r_rev, bool* post) {
  Geo g = g0;
  int R = 4;
  if (dtype == QG_C64 && role == ROLE_FWD) R = r_fwd;
  if (dtype == QG_C64 && role == ROLE_REV_HIST) R = r_rev;
  g.R = std::min(R, g.K);
  g.t = g.K - g.R;
  g.threads = 1 << g.t;
  // the kernel variant with post-gate diagonal ops is chosen per launch
  *post = false;
  return g;
}
#define with_role(g0, dtype, role, tf) with_role_r(g0, dtype, role, p->r_fwd, p->r_rev, tf)

// Columns of the basis-index map of CNOT-ladder stacking `s` (state.py:229-241):
// stacking 1 acts the even bonds first, stacking 0 the odd bonds first.
static void ladder_map(int n, int stacking, uint32_t* cols) {
  for (int k = 0; k < n; ++k) {
    uint32_t v = 1u << k;
    for (int grp = 0; grp < 2; ++grp) {
      const int first = (stacking == 1) ? (grp == 0 ? 0 : 1) : (grp == 0 ? 1 : 0);
      for (int i = first; i + 1 < n; i += 2)
        if ((v >> (n - 1 - i)) & 1u) v ^= 1u << (n - 2 - i);
    }
    cols[k] = v;
  }
}

static void ladder_map(int n, int stacking, std::vector<uint32_t>& cols) {
  cols.assign(n, 0);
  ladder_map(n, stacking, cols.data());
}

// The same map in the shift form the kernels evaluate (ladder_of):
// map(v) = v ^ ((v & m1) >> 1) ^ ((v & m2) >> 2).  Each bond's target is the
// bit below its control and the two groups of bonds act one after the other,
// so column k holds bit k and at most bits k-1 and k-2; the masks are read off
// the columns, and the form is checked column by column.
static bool ladder_shift_form(int n, int stacking, uint32_t* m1, uint32_t* m2) {
  std::vector<uint32_t> cols;
  ladder_map(n, stacking, cols);
  uint32_t a = 0, b = 0;
  for (int k = 0; k < n; ++k) {
    const uint32_t c = cols[k];
    if (k >= 1 && ((c >> (k - 1)) & 1u)) a |= 1u << k;
    if (k >= 2 && ((c >> (k - 2)) & 1u)) b |= 1u << k;
    const uint32_t expect = (1u << k) | (((a >> k) & 1u) << (k >= 1 ? k - 1 : 0)) |
                            (((b >> k) & 1u) << (k >= 2 ? k - 2 : 0));
    if (c != expect) return false;
  }
  *m1 = a;
  *m2 = b;
  return true;
}

// Amplitude of |x> in the RY(pi/4) product state by popcount k = |x|
// (mc_clean.py:55-56): cos(pi/8)^(n-k) sin(pi/8)^k, for k = 0..n.
static void ry_init_table(int n, double* tab) {
  const double c8 = std::cos(M_PI / 8), s8 = std::sin(M_PI / 8);
  for (int k = 0; k <= n; ++k) tab[k] = std::pow(c8, n - k) * std::pow(s8, k);
}

// ---------------------------------------------------------------------------
// plans
// ---------------------------------------------------------------------------
enum PlanKind { PK_MCCLEAN = 0, PK_QAOA = 1 };

struct Launch {
  int desc;  // index into the descriptor table (first of `batch` consecutive ones)
  int batch = 1;
  int K = 0;
  bool post = false;  // the pass has diagonal ops after its gates (set at launch from the flags)
  int tab = 0;        // bytes of the table the pass stages in LDS (set at launch)
  int R, threads, grid, phase;
  bool two;
  int kind;  // 0 pass kernel, 1 generic observable, 2 convert-in (ini state)
};

struct qg_plan {
  qg_ctx* ctx = nullptr;
  int kind = PK_MCCLEAN;
  int n = 0, L = 0, dtype = 0, ent = 0, batch = 1;
  bool history = true;
  int ckpt = 0;   // checkpoint period in layers (uncompute mode)
  int ckpt_req = 0;  // requested period (0: history when it fits)
  int nckpt = 0;
  size_t off_work = 0, off_ckpt = 0;
  int kmax = 13, cmin = 3;  // cmin 3: contiguous runs of >= 8 amplitudes in high-bit passes
  bool kmax_auto = true;    // kmax from default_kmax (not set by qg_plan_tune)
  int r_fwd = 4, r_rev = 4;  // amplitudes per thread (log2) of forward / history-adjoint passes (complex64)
  int lds64_req = -1;  // complex64 McClean: FP64 LDS tiles (-1 auto, 0 no, 1 all passes, 2 forward passes)
  bool lds64 = false;      // forward passes use FP64 LDS tiles
  bool lds64_adj = false;  // adjoint (two-state) passes too
  int chunk = 1;  // circuits resident at once (QAOA and McClean batches)
  int chunk_req = 0;  // QAOA: requested chunk (0: as many circuits as fit)
  int wg_per_cu = 16; // QAOA: workgroups per CU of a batched pass launch, in all circuits
  size_t qaoa_terms_cap = 0;  // QAOA: DTerm capacity of d_terms (chunk x terms per circuit)
  int mc_batch = 0;  // McClean batch plans (qg_mcclean_batch_plan): circuits per call
  bool fused = true;       // n <= 13: whole circuit per workgroup (qg_fused.hip)
  char* h_stage = nullptr;  // pinned staging of QAOA term and phase tables
  size_t stage_cap = 0;
  DTerm* h_cstage = nullptr;  // pinned staging of split terms (streamed launches)
  size_t cstage_cap = 0;
  size_t stream_bound = 0;  // streamed run: split-term capacity
  size_t stream_next = 0;   // streamed run: first launch not yet flushed
  char* h_arena = nullptr;  // per-call tables, pinned host
  char* d_arena = nullptr;  // and their device copy
  size_t arena_cap = 0;
  void* d_fstate = nullptr;  // final state of the last fused McClean run
  std::vector<Geo> geo;
  size_t S = 0;  // bytes per state
  // workspace
  char* ws = nullptr;
  size_t ws_bytes = 0;  // what the current layout needs
  size_t ws_cap = 0;    // size of the attached workspace (allocated or supplied)
  bool ws_owned = false;
  size_t off_hist = 0, off_lam = 0, off_ini = 0, off_prep = 0;
  // small device tables
  uint32_t lad1 = 0, lad2 = 0;  // the ladder map in shift form (gather/scatter)
  double* d_init = nullptr;     // n + 1
  DTerm* d_terms = nullptr;     // diagonal terms (max_terms)
  std::vector<DTerm> h_terms;   // host copy of what d_terms holds for the current run
  std::vector<DTerm> h_cterms;  // the terms split per pass launch (PassDesc::cterms)
  DTerm* d_cterms = nullptr;
  void* d_ini_states = nullptr;  // QAOA: the chunk's initial states (when given)
  size_t ini_cap = 0;            // its size in bytes
  size_t cterm_cap = 0;
  DTerm* d_xy = nullptr;
  int max_terms = 0;
  double* d_lut = nullptr;
  int lut_len = 0;
  double* d_part = nullptr;
  int gmax = 1, nrows = 0;
  int* d_fin = nullptr;         // ptr | rows | cnt
  size_t fin_cap = 0;
  // capacities of the small tables (they only grow, so a plan re-laid out for
  // a shape no larger than one it held before allocates nothing)
  int cap_n = 0;            // d_init (n + 1), h_tabs
  int cap_terms = 0;        // d_terms, d_xy
  size_t cap_part = 0;      // d_part elements (rows x gmax)
  int cap_nout = 0;         // d_scale, d_out, h_out
  size_t cap_fstate = 0;    // d_fstate bytes
  char* h_tabs = nullptr;   // pinned staging of the initial-state table
  double* d_scale = nullptr;
  double* d_out = nullptr;
  double* d_dot = nullptr;      // kDotBlocks + 1
  PassDesc* d_desc = nullptr;
  int ndesc_cap = 0;
  // pinned staging
  PassDesc* h_desc = nullptr;
  double* h_out = nullptr;
  int nout = 0;
  std::vector<Launch> launches;
  float last_ms = 0.f;
  // per-launch profiling: one event before every launch and one after the last
  bool profile = false;
  // one event set per profiled run since profiling was switched on (the
  // launches' kinds beside it); times are read when they are asked for
  std::vector<std::vector<hipEvent_t>> prof_ev;
  std::vector<std::vector<int>> prof_kind;
  size_t prof_runs = 0;
  mutable std::vector<float> pass_ms;  // durations of those runs' launches
  mutable std::vector<int> pass_kind;  // 0 forward, 1 reverse (two-state), 2 forward+observable, 3 other
  int last_state = -1;  // 0: final state in hist, used by get_state
  // the pass launches of a run as a HIP graph, replayed while the launch list
  // (variants, grids, descriptor slots) is unchanged; the kernels read their
  // per-call parameters from the descriptor table, uploaded before each replay.
  // Off by default: a capture is invalidated by device-synchronising calls
  // (allocation, synchronous copies) that other threads make meanwhile
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  std::vector<int64_t> gsig;
  // McClean with x / y observable terms on the pass kernels: lambda = O psi
  // and E = Re<psi|lambda> by the elementwise kernels (a kind-1 launch)
  const void* gen_in = nullptr;
  void* gen_out = nullptr;
  int gen_ndiag = 0, gen_nxy = 0;
  int obs_desc = -1;     // McClean: descriptor whose row K holds E
  int adj_desc = -1;     // McClean: first adjoint descriptor
  bool state_sign = false;  // the pass kernels' final state is -1 times the circuit's
  void* last_state_ptr = nullptr;
};

static void* wsp(qg_plan* p, size_t off) { return p->ws + off; }
static void free_small(qg_plan* p);
static void drop_graph(qg_plan* p);

static int ensure_workspace(qg_plan* p) {
  if (p->ws) return QG_OK;
  void* m = nullptr;
  hipError_t e = hipMalloc(&m, p->ws_bytes);
  if (e != hipSuccess)
    return fail(QG_ERR_NOMEM, "workspace allocation of " + std::to_string(p->ws_bytes) + " bytes failed: " +
                                  hipGetErrorString(e));
  p->ws = (char*)m;
  p->ws_cap = p->ws_bytes;
  p->ws_owned = true;
  return QG_OK;
}

// Kernel storage code of a plan's passes: complex128 (0), complex64 (1), or
// complex64 with FP64 LDS tiles (2).
static int kdtype(const qg_plan* p, bool two) {
  return (p->dtype == QG_C64 && (two ? p->lds64_adj : p->lds64)) ? 2 : p->dtype;
}

static size_t lds_bytes_of(int dtype, int K, int threads) {  // the larger (adjoint) variant
  return (size_t)(1u << K) * (dtype == QG_C64 ? 8 : 16) + (size_t)((threads + 63) / 64) * (kMaxK + 1) * 8 +
         (dtype == QG_C64 ? (size_t)K * threads * 8 : 0);
}

// Workgroups per pass launch: as many as are resident at once on the device
// (occupancy of the kernel variant, by its registers and LDS), rounded down to
// a power of two and at most one per tile.
constexpr size_t kMaxProfRuns = 256;  // profiled runs whose launch events a plan keeps
constexpr int kBatchWgPerCu = 16;  // workgroups per CU of a batched QAOA pass launch
constexpr int kWgOversub = 4;      // forward pass grids: this many times the resident workgroups
// McClean history-mode passes read the state they read once (a forward
// pass's input, an adjoint pass's history state) with non-temporal loads
// (F_NT_A) while one launch covers at most 2^25 amplitudes: the lines the
// pass writes then stay in the caches for the next pass.  Measured per pass,
// non-temporal against ordinary loads (us): forward, complex64 24 qubits
// 48-55 / 55-59, 25 qubits 98.5 / 120.5, complex128 24 qubits 93 / 105,
// 25 qubits 199 / 207; adjoint, complex64 24 qubits 105.8 / 109.8,
// 25 qubits 212.7 / 222, complex128 24 qubits 164 / 176.6, 25 qubits 329 /
// 345.  From 26 qubits the hint costs: complex64 26 qubits forward 285 /
// 254.5 and adjoint 494 / 469.5, 27 qubits 711 / 565 and 1117 / 1017.
constexpr uint64_t kNtMaxLaunchAmps = (uint64_t)1 << 25;

// Workgroups of this pass variant resident per CU (occupancy API, or an LDS
// estimate where the query fails).
static int pass_occupancy(const qg_plan* p, const Geo& g, bool two, bool post, int phase) {
  int occ = 0;
  if (launch_pass(kdtype(p, two), g.R, two, phase, post, nullptr, g.K, 0, 1, g.threads, p->n, nullptr, &occ, 0) !=
          hipSuccess ||
      occ < 1) {
    const size_t lds = lds_bytes_of(kdtype(p, two), g.K, g.threads);
    occ = (int)std::max<size_t>(1, (160 * 1024) / lds);
    occ = std::min(occ, std::max(1, 512 / g.threads));
  }
  return occ;
}

static int grid_for_pass(const qg_plan* p, const Geo& g, bool two, bool post, int phase) {
  const int occ = pass_occupancy(p, g, two, post, phase);
  const uint64_t ntiles = 1ull << (p->n - g.K);
  // forward passes: more workgroups than are resident spread the tiles more
  // evenly; adjoint passes keep one resident wave of long-lived workgroups
  // (their reduction accumulators last across the workgroup's tiles)
  uint64_t target = (uint64_t)p->ctx->num_cu * occ * (two ? 1 : kWgOversub);
  uint64_t grid = 1;
  while (grid * 2 <= target && grid * 2 <= ntiles) grid *= 2;
  return (int)grid;
}

// Geometry of the gate-free pass that forms lambda = D psi and E when the
// last forward pass cannot carry it (a 1024-thread pass).
static Geo obs_geo(int n) {
  const int K = std::min(n, 11);
  return make_geo(K, K, n, 0, 0);
}

// The largest grid any pass variant of these geometries may use (the stride
// of the partial rows).
static int grid_max(const qg_plan* p) {
  int gm = 1;
  std::vector<Geo> geos = p->geo;
  if (p->kind == PK_MCCLEAN) geos.push_back(obs_geo(p->n));  // the observable's own pass
  for (const Geo& g : geos)
    for (int role = 0; role < 3; ++role)
      for (int phase = 0; phase < 3; ++phase) {
        bool tf;
        const Geo G = with_role(g, p->dtype, role, &tf);
        gm = std::max(gm, grid_for_pass(p, G, role != ROLE_FWD, false, phase));
        gm = std::max(gm, grid_for_pass(p, G, true, true, phase));
      }
  return gm;
}

// Grows the small device tables of a plan to hold ndesc descriptors, nrows
// partial rows of gmax workgroups and nout results (and the terms and static
// tables of an n-qubit circuit); tables already large enough are kept.
static int ensure_small(qg_plan* p, int ndesc, int nrows, int nout) {
  const int n = p->n;
  const bool grow_n = n > p->cap_n || !p->d_init;
  const bool grow_terms = std::max(64, n * n) > p->cap_terms || !p->d_terms;
  const size_t part = (size_t)std::max(1, nrows) * p->gmax;
  const size_t fin = (size_t)(3 * (nout + 1) + 3 * nrows + 8);
  const bool grow_part = part > p->cap_part || !p->d_part;
  const bool grow_out = nout > p->cap_nout || !p->d_out;
  const bool grow_fin = fin > p->fin_cap || !p->d_fin;
  const bool grow_desc = ndesc > p->ndesc_cap || !p->d_desc;
  if (grow_n || grow_terms || grow_part || grow_out || grow_fin || grow_desc) {
    // the tables being replaced may still be read by queued work
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
    drop_graph(p);  // a captured graph holds their addresses
  }
  if (grow_n) {
    (void)hipFree(p->d_init);
    if (p->h_tabs) (void)hipHostFree(p->h_tabs);
    p->d_init = nullptr;
    p->h_tabs = nullptr;
    p->cap_n = 0;
    HIPCHK(hipMalloc(&p->d_init, sizeof(double) * (n + 1)));
    HIPCHK(hipHostMalloc(&p->h_tabs, sizeof(double) * (n + 1), hipHostMallocDefault));
    p->cap_n = n;
  }
  if (grow_terms) {
    (void)hipFree(p->d_terms);
    (void)hipFree(p->d_xy);
    p->d_terms = p->d_xy = nullptr;
    p->cap_terms = 0;
    const int mt = std::max(64, n * n);
    HIPCHK(hipMalloc(&p->d_terms, sizeof(DTerm) * mt));
    HIPCHK(hipMalloc(&p->d_xy, sizeof(DTerm) * mt));
    p->cap_terms = mt;
  }
  p->max_terms = p->cap_terms;
  if (grow_part) {
    (void)hipFree(p->d_part);
    p->d_part = nullptr;
    p->cap_part = 0;
    HIPCHK(hipMalloc(&p->d_part, sizeof(double) * part));
    HIPCHK(hipMemset(p->d_part, 0, sizeof(double) * part));
    p->cap_part = part;
  }
  if (grow_fin) {
    (void)hipFree(p->d_fin);
    p->d_fin = nullptr;
    p->fin_cap = 0;
    HIPCHK(hipMalloc(&p->d_fin, sizeof(int) * fin));
    p->fin_cap = fin;
  }
  if (grow_out) {
    (void)hipFree(p->d_scale);
    (void)hipFree(p->d_out);
    if (p->h_out) (void)hipHostFree(p->h_out);
    p->d_scale = p->d_out = nullptr;
    p->h_out = nullptr;
    p->cap_nout = 0;
    HIPCHK(hipMalloc(&p->d_scale, sizeof(double) * std::max(1, nout)));
    HIPCHK(hipMalloc(&p->d_out, sizeof(double) * std::max(1, nout)));
    HIPCHK(hipHostMalloc(&p->h_out, sizeof(double) * std::max(1, nout), hipHostMallocDefault));
    p->cap_nout = nout;
  }
  if (!p->d_dot) HIPCHK(hipMalloc(&p->d_dot, sizeof(double) * (kDotBlocks + 1)));
  if (grow_desc) {
    (void)hipFree(p->d_desc);
    if (p->h_desc) (void)hipHostFree(p->h_desc);
    p->d_desc = nullptr;
    p->h_desc = nullptr;
    p->ndesc_cap = 0;
    HIPCHK(hipMalloc(&p->d_desc, sizeof(PassDesc) * std::max(1, ndesc)));
    HIPCHK(hipHostMalloc(&p->h_desc, sizeof(PassDesc) * std::max(1, ndesc), hipHostMallocDefault));
    p->ndesc_cap = ndesc;
  }
  p->nout = nout;
  return QG_OK;
}

extern "C" int qg_plan_destroy(qg_plan* p) {
  if (!p) return QG_OK;
  (void)hipSetDevice(p->ctx->device);
  (void)hipStreamSynchronize(p->ctx->stream);
  if (p->ws_owned && p->ws) (void)hipFree(p->ws);
  free_small(p);
  drop_graph(p);
  for (auto& set : p->prof_ev)
    for (hipEvent_t e : set) (void)hipEventDestroy(e);
  delete p;
  return QG_OK;
}

extern "C" int qg_plan_workspace_size(const qg_plan* p, size_t* bytes) {
  if (!p || !bytes) return fail(QG_ERR_INVALID, "NULL argument");
  *bytes = p->ws_bytes;
  return QG_OK;
}

extern "C" int qg_plan_set_workspace(qg_plan* p, void* dev, size_t bytes) {
  if (!p || !dev) return fail(QG_ERR_INVALID, "NULL argument");
  if (bytes < p->ws_bytes) return fail(QG_ERR_INVALID, "workspace smaller than qg_plan_workspace_size");
  if (((uintptr_t)dev) % 256 != 0) return fail(QG_ERR_INVALID, "workspace must be 256-byte aligned");
  if (p->ws_owned && p->ws) {
    (void)hipStreamSynchronize(p->ctx->stream);
    (void)hipFree(p->ws);
  }
  p->ws = (char*)dev;
  p->ws_cap = bytes;
  p->ws_owned = false;
  return QG_OK;
}

extern "C" int qg_plan_last_ms(const qg_plan* p, double* ms) {
  if (!p || !ms) return fail(QG_ERR_INVALID, "NULL argument");
  *ms = p->last_ms;
  return QG_OK;
}

// prob[k] = <v|(I + P_k)/2|v> = |P_k^+ v|^2 for each term, in the given order
// (weights ignored): the probability of the +1 outcome that the projectors of
// observable.py:126-177 give in base.py:22-46, also for a state whose norm is
// not exactly 1.  One pass over v per 32 diagonal or 16 other terms.
// The term table of the term-expectation kernel: terms[0, nterms) in order,
// then the identity (|v|^2).
static int pterm_table(int n, const qg_term* terms, int nterms, std::vector<PTerm>& h) {
  if (nterms < 0 || (nterms > 0 && !terms)) return fail(QG_ERR_INVALID, "NULL argument");
  if (n > 31) return fail(QG_ERR_UNSUPPORTED, "term expectations need n <= 31");
  h.assign(nterms + 1, PTerm{});
  for (int k = 0; k < nterms; ++k) {
    const qg_term& t = terms[k];
    if (t.i < 0 || t.i >= n) return fail(QG_ERR_INVALID, "observable: qubit index out of range");
    PTerm& q = h[k];
    q.a = n - 1 - t.i;
    q.b = -1;
    q.pad = 0;
    switch (t.kind) {
      case QG_TERM_Z: q.kind = 0; break;
      case QG_TERM_ZZ:
        if (t.j <= t.i || t.j >= n)
          return fail(QG_ERR_INVALID, "observable: zz terms need i < j < n (upper triangle, observable.py:56-64)");
        q.kind = 0;
        q.b = n - 1 - t.j;
        break;
      case QG_TERM_X: q.kind = 1; break;
      case QG_TERM_Y: q.kind = 2; break;
      default: return fail(QG_ERR_INVALID, "observable: unknown term kind");
    }
  }
  h[nterms] = PTerm{3, 0, -1, 0};  // |v|^2
  return QG_OK;
}

// e[k] = <v|P_k|v> for the terms in order, e[nterms] = |v|^2 (weights
// ignored).  One pass over v per 32 diagonal or 16 other terms.
static int term_expec(qg_ctx* c, int dtype, const void* v, int n, const qg_term* terms, int nterms, double* e) {
  std::vector<PTerm> h;
  CHK(pterm_table(n, terms, nterms, h));
  const int nt = nterms + 1;
  HIPCHK(hipSetDevice(c->device));
  const size_t tb = ((sizeof(PTerm) * nt + 255) / 256) * 256;
  const size_t pb = sizeof(double) * kTermBlocks * kTermPartRows;
  char* buf = nullptr;
  HIPCHK(hipMalloc(&buf, tb + pb + sizeof(double) * nt));
  PTerm* d_t = (PTerm*)buf;
  double* d_part = (double*)(buf + tb);
  double* d_out = (double*)(buf + tb + pb);
  hipError_t err = hipMemcpyAsync(d_t, h.data(), sizeof(PTerm) * nt, hipMemcpyHostToDevice, c->stream);
  if (err == hipSuccess) err = simple_term_expec(dtype, v, n, d_t, nt, d_part, d_out, c->stream, h.data());
  if (err == hipSuccess) err = hipMemcpyAsync(e, d_out, sizeof(double) * nt, hipMemcpyDeviceToHost, c->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
  (void)hipFree(buf);
  if (err != hipSuccess) return fail(QG_ERR_HIP, std::string("term expectations: ") + hipGetErrorString(err));
  return QG_OK;
}

// prob[k] = <v|(I + P_k)/2|v> = |P_k^+ v|^2 for each term, in the given order
// (weights ignored): the probability of the +1 outcome that the projectors of
// observable.py:126-177 give in base.py:22-46, also for a state whose norm is
// not exactly 1.
static int term_prob(qg_ctx* c, int dtype, const void* v, int n, const qg_term* terms, int nterms, double* prob) {
  if (nterms < 0 || (nterms > 0 && (!terms || !prob))) return fail(QG_ERR_INVALID, "NULL argument");
  if (nterms == 0) return QG_OK;
  std::vector<double> e(nterms + 1);
  CHK(term_expec(c, dtype, v, n, terms, nterms, e.data()));
  for (int k = 0; k < nterms; ++k) prob[k] = 0.5 * (e[nterms] + e[k]);
  return QG_OK;
}

extern "C" int qg_plan_expec_val(qg_plan* p, const qg_term* terms, int nterms, double* E_out) {
  if (!p || !E_out) return fail(QG_ERR_INVALID, "NULL argument");
  if (!p->last_state_ptr) return fail(QG_ERR_INVALID, "no state: run the circuit first");
  for (int k = 0; k < nterms; ++k)
    if (terms && !std::isfinite(terms[k].w)) return fail(QG_ERR_INVALID, "observable: non-finite weight");
  std::vector<double> e(nterms + 1);
  CHK(term_expec(p->ctx, p->dtype, p->last_state_ptr, p->n, terms, nterms, e.data()));
  double E = 0.0;
  for (int k = 0; k < nterms; ++k) E += terms[k].w * e[k];
  *E_out = E;
  return QG_OK;
}

extern "C" int qg_plan_term_prob(qg_plan* p, const qg_term* terms, int nterms, double* prob) {
  if (!p) return fail(QG_ERR_INVALID, "NULL plan");
  if (!p->last_state_ptr) return fail(QG_ERR_INVALID, "no state: run the circuit first");
  return term_prob(p->ctx, p->dtype, p->last_state_ptr, p->n, terms, nterms, prob);
}

extern "C" int qg_plan_get_state(qg_plan* p, double* out) {
  if (!p || !out) return fail(QG_ERR_INVALID, "NULL argument");
  if (!p->last_state_ptr) return fail(QG_ERR_INVALID, "no state: run the circuit first");
  HIPCHK(hipSetDevice(p->ctx->device));
  const size_t N = 1ull << p->n;
  double* tmp = nullptr;
  HIPCHK(hipMalloc(&tmp, N * 16));
  hipError_t e = simple_convert_out(p->dtype, p->last_state_ptr, tmp, p->n, p->ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, tmp, N * 16, hipMemcpyDeviceToHost, p->ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(p->ctx->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(QG_ERR_HIP, std::string("qg_plan_get_state: ") + hipGetErrorString(e));
  if (p->state_sign)
    for (size_t k = 0; k < 2 * N; ++k) out[k] = -out[k];
  return QG_OK;
}

// ---------------------------------------------------------------------------
// descriptor helpers
// ---------------------------------------------------------------------------
static void desc_geo(PassDesc& d, int n, const Geo& g) {
  std::memset(&d, 0, sizeof(PassDesc));
  d.n = n;
  d.K = g.K;
  d.t = g.t;
  d.c = g.c;
  d.p = g.p;
  d.row_base = -1;
  for (int k = 0; k < kMaxK; ++k) {
    d.gtype[k] = G_NONE;
    d.ga[k] = 0.0;
    d.gb[k] = 0.0;
  }
}

// U = cos(phi) I - i sin(phi) P with phi = sign * angle / 2.  phi is reduced
// modulo pi to [-pi/2, pi/2]; each reduction multiplies U by -1, a global sign
// the kernels drop (returns true when it did, so the caller can track the
// parity for the returned state vector).
static bool gate_flips(double angle, double sign) {
  const double phi0 = sign * 0.5 * angle;
  const double phi = std::remainder(phi0, M_PI);
  return std::fabs(std::remainder((phi0 - phi) / M_PI, 2.0)) > 0.5;
}

static bool set_gate(PassDesc& d, int lb, int axis, double angle, double sign) {
  const double phi0 = sign * 0.5 * angle;
  const double phi = std::remainder(phi0, M_PI);
  const bool flipped = gate_flips(angle, sign);
  const double c = std::cos(phi), s = std::sin(phi);
  d.gtype[lb] = axis == 0 ? G_X : axis == 1 ? G_Y : G_Z;
  d.ga[lb] = s / (1.0 + c);
  d.gb[lb] = s;
  return flipped;
}

// The (ga, gb) pair set_gate stores for a gate of this angle and sign.
struct GateAB {
  double a, b;
};
static GateAB gate_ab(double angle, double sign) {
  const double phi = std::remainder(sign * 0.5 * angle, M_PI);
  const double c = std::cos(phi), s = std::sin(phi);
  return {s / (1.0 + c), s};
}

// Pinned staging for per-call tables, grown as needed; the previous call's
// copies out of it have completed (every run ends with a stream sync).
static int ensure_stage(qg_plan* p, size_t bytes) {
  if (bytes <= p->stage_cap) return QG_OK;
  if (p->h_stage) (void)hipHostFree(p->h_stage);
  p->h_stage = nullptr;
  p->stage_cap = 0;
  HIPCHK(hipHostMalloc(&p->h_stage, bytes, hipHostMallocDefault));
  p->stage_cap = bytes;
  return QG_OK;
}

// Fills the per-stage-slot gate tables of a descriptor from its per-bit
// arrays, for a launch holding 2^R amplitudes per thread (the kernel's stage
// layout: stage 0 holds local bits [t, t + R), stage s >= 1 holds
// [max(0, t - R s), +R); a bit is gated in the first stage that holds it).
// red_first: the adjoint pass takes its reductions before the gates (A is not
// rotated; see the kernel's POST parameter).
// Tile order of a pass (F_XCD_ORDER in the kernel): XCD-aware for QAOA plans
// and for passes whose contiguous runs (2^c amplitudes) are shorter than a
// 128-byte cache line, so that neighbouring tiles, which share lines, share an
// L2; round-robin otherwise.  Measured: McClean 0.6-2.5% faster round-robin at
// 24-30 qubits, the 512-circuit QAOA batch 1.6% slower.  A function of the plan
// kind and the pass geometry only, so a circuit gets the same bits alone and
// in a batch.
static void desc_order(const qg_plan* p, PassDesc& d) {
  const int amp_log2 = p->dtype == QG_C64 ? 3 : 4;
  if (p->kind == PK_QAOA || d.c + amp_log2 < 7) d.flags |= F_XCD_ORDER;
  else d.flags &= ~F_XCD_ORDER;
}

static void desc_slots(PassDesc& d, int R, bool two, bool red_first) {
  int los[kMaxStages];
  int ns = 1;
  los[0] = d.t;
  for (int s = 1; s < kMaxStages; ++s) {
    los[s] = 0;
    if (ns == s && los[s - 1] > 0) {
      los[s] = std::max(0, d.t - R * s);
      ns = s + 1;
    }
  }
  for (int k = 0; k < kMaxSlots; ++k) {
    d.stype[k] = G_NONE;
    d.sa[k] = d.sb[k] = 0.0;
    d.sflag[k] = 0;
    d.slb[k] = 0;
  }
  const bool a_used_after = !two || (d.flags & (F_STORE_A | F_DRED_POST | F_OBS_DIAG));
  int last_a = -1;
  for (int s = 0; s < ns; ++s) {
    const int nnew = s == 0 ? R : std::min(los[s - 1] - los[s], R);
    for (int kk = 0; kk < nnew; ++kk) {
      const int slot = s * R + kk, lb = los[s] + kk;
      d.slb[slot] = lb;
      d.stype[slot] = (two || (d.flags & F_ROT_A)) ? d.gtype[lb] : G_NONE;
      d.sa[slot] = d.ga[lb];
      d.sb[slot] = d.gb[lb];
      int f = 0;
      if (d.gtype[lb] != G_NONE) {
        if ((d.flags & F_ROT_A) && !red_first) {
          f |= SL_ROT_A;
          last_a = slot;
        }
        if (two && (d.flags & F_ROT_B)) f |= SL_ROT_B;
      }
      if (two && !red_first && (d.flags & F_REDUCE) && d.own[lb] && d.row_base >= 0) f |= SL_RED;
      d.sflag[slot] = f;
    }
  }
  // A's last gate matters only if A is used after the gates
  if (!a_used_after && last_a >= 0) d.sflag[last_a] &= ~SL_ROT_A;
  // owned thread bits of a reduce-first pass
  d.xmask = d.ymask = d.zmask = 0;
  if (two && red_first && (d.flags & F_REDUCE) && d.row_base >= 0) {
    for (int lb = 0; lb < d.t; ++lb) {
      if (!d.own[lb]) continue;
      if (d.gtype[lb] == G_X) d.xmask |= 1u << lb;
      if (d.gtype[lb] == G_Y) d.ymask |= 1u << lb;
      if (d.gtype[lb] == G_Z) d.zmask |= 1u << lb;
    }
  }
  // stages after the last one with work (only spectator bits) are skipped
  d.nstage = 1;
  for (int s = 0; s < ns; ++s)
    for (int kk = 0; kk < R; ++kk)
      if (d.stype[s * R + kk] != G_NONE || (d.sflag[s * R + kk] & SL_RED)) d.nstage = s + 1;
}

// Splits the diagonal terms of a descriptor by its launch's register bits
// (PassDesc::cterms, nct, nht, pwc); appends the lists to out and returns
// their offset there.
static size_t split_terms(const qg_plan* p, PassDesc& d, int R, std::vector<DTerm>& out) {
  const size_t off0 = out.size();
  const size_t base = (size_t)(d.dterms - p->d_terms);
  int rpos[kMaxR];
  for (int k = 0; k < R; ++k) {
    const int l = d.t + k;
    rpos[k] = l < d.c ? l : d.p + (l - d.c);
  }
  auto regk = [&](int b) {
    for (int k = 0; k < R; ++k)
      if (rpos[k] == b) return k;
    return -1;
  };
  auto sgn = [](int a, int k) { return ((a >> k) & 1) ? -1.0 : 1.0; };
  std::vector<DTerm> hl[kMaxR];
  for (int a = 0; a < (1 << kMaxR); ++a) d.pwc[a] = 0.0;
  for (int i = 0; i < d.ndterms; ++i) {
    const DTerm& t = p->h_terms[base + i];
    const int ka = regk(t.a), kb = t.b >= 0 ? regk(t.b) : -1;
    if (ka < 0 && kb < 0) {
      const uint32_t m = (1u << t.a) | (t.b >= 0 ? (1u << t.b) : 0u);
      out.push_back({(int32_t)m, 0, t.w});
    } else if (t.b < 0 || (ka >= 0 && kb >= 0)) {
      for (int a = 0; a < (1 << R); ++a) d.pwc[a] += t.w * sgn(a, ka) * (t.b < 0 ? 1.0 : sgn(a, kb));
    } else {
      const int k = ka >= 0 ? ka : kb, other = ka >= 0 ? t.b : t.a;
      hl[k].push_back({(int32_t)(1u << other), 0, t.w});
    }
  }
  d.nct = (int32_t)(out.size() - off0);
  for (int k = 0; k < kMaxR; ++k) {
    d.nht[k] = k < R ? (int32_t)hl[k].size() : 0;
    if (k < R) out.insert(out.end(), hl[k].begin(), hl[k].end());
  }
  return off0;
}

static void drop_graph(qg_plan* p) {
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  p->gexec = nullptr;
  p->gsig.clear();
}

// lambda = O psi with x / y terms (k_obs_apply) and E = Re<psi|lambda> into
// d_dot[kDotBlocks]: the observable step of a McClean run whose observable
// the pass kernels' diagonal form cannot carry.
static hipError_t launch_generic_obs(qg_plan* p) {
  hipError_t e = simple_obs_apply(p->dtype, p->gen_in, p->gen_out, p->n, p->d_terms, p->gen_ndiag, p->d_xy,
                                  p->gen_nxy, p->ctx->stream);
  if (e == hipSuccess)
    e = simple_dot(p->dtype, p->gen_in, p->gen_out, p->n, 0, p->d_dot, p->d_dot + kDotBlocks, p->ctx->stream);
  return e;
}

static hipError_t launch_all(qg_plan* p) {
  qg_ctx* c = p->ctx;
  for (const Launch& l : p->launches) {
    if (l.kind == 1) {
      const hipError_t e = launch_generic_obs(p);
      if (e != hipSuccess) return e;
    }
    if (l.kind != 0) continue;
    const hipError_t e = launch_pass(kdtype(p, l.two), l.R, l.two, l.phase, l.post, p->d_desc + l.desc, l.K, l.grid,
                                     l.batch, l.threads, p->n, c->stream, nullptr, l.tab);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Prepares and uploads descriptors [first_desc, ndesc) and launches the
// plan's launch list (whose descriptors lie in that range).
static int run_launches(qg_plan* p, int ndesc, int first_desc = 0) {
  qg_ctx* c = p->ctx;
  p->h_cterms.clear();
  std::vector<std::pair<int, size_t>> cterm_at;  // (descriptor, offset of its split terms)
  for (Launch& l : p->launches)
    if (l.kind == 0) {
      // partial rows are strided by gmax: a larger grid would overlap them
      if (l.grid > p->gmax) return fail(QG_ERR_HIP, "internal: pass grid exceeds the partial-row stride");
      const PassDesc& d = p->h_desc[l.desc];
      // passes without a phase use the kernel variant without phase code
      if (!(d.flags & (F_PHASE_PRE | F_PHASE_POST))) l.phase = 0;
      // adjoint passes whose A is a read-only pass output take the
      // reduce-first form (POST = false), also with a pre phase that carries
      // the d/dgamma reduction; the rest the general form
      l.post = l.two && ((d.flags & (F_DRED_POST | F_PHASE_POST | F_OBS_DIAG | F_STORE_A)) != 0 ||
                         !(d.flags & F_LOAD_A) || ((d.flags & F_PHASE_PRE) && !(d.flags & F_DRED_PRE)));
      for (int b = 0; b < l.batch; ++b) {
        desc_slots(p->h_desc[l.desc + b], l.R, l.two, l.two && !l.post);
        desc_order(p, p->h_desc[l.desc + b]);
      }
      if (d.flags & (F_PHASE_PRE | F_PHASE_POST | F_DRED_PRE | F_DRED_POST | F_OBS_DIAG))
        for (int b = 0; b < l.batch; ++b)
          cterm_at.push_back({l.desc + b, split_terms(p, p->h_desc[l.desc + b], l.R, p->h_cterms)});
      l.tab = ((d.flags & F_INIT_TAB) ? (int)sizeof(double) * (p->n + 1) : 0) +
              ((l.phase == 1 && (d.flags & (F_PHASE_PRE | F_PHASE_POST))) ? (int)sizeof(double) * 2 * (2 * d.lut_off + 1)
                                                                           : 0);
    }
  if (!p->h_cterms.empty()) {
    if (p->h_cterms.size() > p->cterm_cap) {
      HIPCHK(hipStreamSynchronize(c->stream));  // the old table may still be read
      (void)hipFree(p->d_cterms);
      p->d_cterms = nullptr;
      p->cterm_cap = 0;
      HIPCHK(hipMalloc(&p->d_cterms, sizeof(DTerm) * p->h_cterms.size()));
      p->cterm_cap = p->h_cterms.size();
    }
    HIPCHK(hipMemcpyAsync(p->d_cterms, p->h_cterms.data(), sizeof(DTerm) * p->h_cterms.size(),
                          hipMemcpyHostToDevice, c->stream));
  }
  for (const auto& e : cterm_at) p->h_desc[e.first].cterms = p->d_cterms + e.second;
  HIPCHK(hipMemcpyAsync(p->d_desc + first_desc, p->h_desc + first_desc, sizeof(PassDesc) * (ndesc - first_desc),
                        hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  if (p->use_graph && !p->profile) {
    std::vector<int64_t> sig;
    sig.reserve(p->launches.size() * 11 + 2);
    sig.push_back(kdtype(p, false) * 4 + kdtype(p, true));
    sig.push_back(p->n);
    sig.push_back((int64_t)(intptr_t)p->d_desc);
    // the elementwise observable launch bakes its buffers and term counts in
    sig.push_back((int64_t)(intptr_t)p->gen_in);
    sig.push_back((int64_t)(intptr_t)p->gen_out);
    sig.push_back(p->gen_ndiag);
    sig.push_back(p->gen_nxy);
    for (const Launch& l : p->launches) {
      const int64_t v[11] = {l.kind, l.R, l.two, l.phase, l.post, l.desc, l.K, l.grid, l.batch, l.threads, l.tab};
      sig.insert(sig.end(), v, v + 11);
    }
    if (!p->gexec || sig != p->gsig) {
      drop_graph(p);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        e = launch_all(p);
        const hipError_t e2 = hipStreamEndCapture(c->stream, &g);
        if (e == hipSuccess) e = e2;
      }
      if (e == hipSuccess) e = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
      if (g) (void)hipGraphDestroy(g);
      if (e == hipSuccess) {
        p->gsig.swap(sig);
      } else {
        // capture unavailable: launch directly from now on
        (void)hipGetLastError();
        p->gexec = nullptr;
        p->use_graph = false;
      }
    }
    if (p->gexec) {
      HIPCHK(hipGraphLaunch(p->gexec, c->stream));
      return QG_OK;
    }
  }
  const size_t nl = p->launches.size();
  std::vector<hipEvent_t>* pev = nullptr;
  if (p->profile) {
    // a ring of the latest kMaxProfRuns runs: run r of the series uses slot r mod kMaxProfRuns
    const size_t slot = p->prof_runs % kMaxProfRuns;
    if (p->prof_ev.size() <= slot) p->prof_ev.resize(slot + 1);
    pev = &p->prof_ev[slot];
    while (pev->size() < nl + 1) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      pev->push_back(e);
    }
  }
  for (size_t k = 0; k < nl; ++k) {
    const Launch& l = p->launches[k];
    if (pev) HIPCHK(hipEventRecord((*pev)[k], c->stream));
    if (l.kind == 0) {
      HIPCHK(launch_pass(kdtype(p, l.two), l.R, l.two, l.phase, l.post, p->d_desc + l.desc, l.K, l.grid, l.batch,
                         l.threads, p->n, c->stream, nullptr, l.tab));
    } else if (l.kind == 1) {
      HIPCHK(launch_generic_obs(p));
    }
  }
  if (pev) HIPCHK(hipEventRecord((*pev)[nl], c->stream));
  return QG_OK;
}

// Batched QAOA runs carry thousands of descriptors.  A streamed run
// prepares, uploads and launches the plan's launches in groups while the
// host is still building later ones, so the device starts on the first
// passes early.  stream_begin sizes the split-term table for at most `bound`
// terms; stream_flush launches every launch built so far (from
// p->stream_next on); descriptors of flushed launches must not change.
static int stream_begin(qg_plan* p, size_t bound) {
  qg_ctx* c = p->ctx;
  if (bound > p->cterm_cap) {
    HIPCHK(hipStreamSynchronize(c->stream));  // the old table may still be read
    (void)hipFree(p->d_cterms);
    p->d_cterms = nullptr;
    p->cterm_cap = 0;
    HIPCHK(hipMalloc(&p->d_cterms, sizeof(DTerm) * bound));
    p->cterm_cap = bound;
  }
  if (bound > p->cstage_cap) {
    if (p->h_cstage) (void)hipHostFree(p->h_cstage);
    p->h_cstage = nullptr;
    p->cstage_cap = 0;
    HIPCHK(hipHostMalloc(&p->h_cstage, sizeof(DTerm) * bound, hipHostMallocDefault));
    p->cstage_cap = bound;
  }
  p->stream_bound = bound;
  p->stream_next = 0;
  p->h_cterms.clear();
  p->h_cterms.reserve(bound);
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  return QG_OK;
}

static int stream_flush(qg_plan* p) {
  qg_ctx* c = p->ctx;
  const uint32_t split_flags = F_PHASE_PRE | F_PHASE_POST | F_DRED_PRE | F_DRED_POST | F_OBS_DIAG;
  const size_t g0 = p->stream_next, g1 = p->launches.size();
  if (g1 <= g0) return QG_OK;
  const size_t ct0 = p->h_cterms.size();
  int d0 = 0x7fffffff, d1 = 0;
  for (size_t k = g0; k < g1; ++k) {
    Launch& l = p->launches[k];
    if (l.kind != 0) continue;
    if (l.grid > p->gmax) return fail(QG_ERR_HIP, "internal: pass grid exceeds the partial-row stride");
    const PassDesc& d = p->h_desc[l.desc];
    // the preparation of run_launches
    if (!(d.flags & (F_PHASE_PRE | F_PHASE_POST))) l.phase = 0;
    l.post = l.two && ((d.flags & (F_DRED_POST | F_PHASE_POST | F_OBS_DIAG | F_STORE_A)) != 0 ||
                       !(d.flags & F_LOAD_A) || ((d.flags & F_PHASE_PRE) && !(d.flags & F_DRED_PRE)));
    for (int b = 0; b < l.batch; ++b) {
      desc_slots(p->h_desc[l.desc + b], l.R, l.two, l.two && !l.post);
      desc_order(p, p->h_desc[l.desc + b]);
    }
    if (d.flags & split_flags)
      for (int b = 0; b < l.batch; ++b) {
        if (p->h_cterms.size() + (size_t)std::max(0, p->h_desc[l.desc + b].ndterms) > p->stream_bound)
          return fail(QG_ERR_HIP, "internal: split terms exceed their bound");
        const size_t off = split_terms(p, p->h_desc[l.desc + b], l.R, p->h_cterms);
        p->h_desc[l.desc + b].cterms = p->d_cterms + off;
      }
    l.tab = ((d.flags & F_INIT_TAB) ? (int)sizeof(double) * (p->n + 1) : 0) +
            ((l.phase == 1 && (d.flags & (F_PHASE_PRE | F_PHASE_POST))) ? (int)sizeof(double) * 2 * (2 * d.lut_off + 1)
                                                                         : 0);
    d0 = std::min(d0, l.desc);
    d1 = std::max(d1, l.desc + l.batch);
  }
  const size_t ct1 = p->h_cterms.size();
  if (ct1 > ct0) {
    std::memcpy(p->h_cstage + ct0, p->h_cterms.data() + ct0, sizeof(DTerm) * (ct1 - ct0));
    HIPCHK(hipMemcpyAsync(p->d_cterms + ct0, p->h_cstage + ct0, sizeof(DTerm) * (ct1 - ct0), hipMemcpyHostToDevice,
                          c->stream));
  }
  if (d1 > d0)
    HIPCHK(hipMemcpyAsync(p->d_desc + d0, p->h_desc + d0, sizeof(PassDesc) * (d1 - d0), hipMemcpyHostToDevice,
                          c->stream));
  for (size_t k = g0; k < g1; ++k) {
    const Launch& l = p->launches[k];
    if (l.kind == 0) {
      HIPCHK(launch_pass(kdtype(p, l.two), l.R, l.two, l.phase, l.post, p->d_desc + l.desc, l.K, l.grid, l.batch,
                         l.threads, p->n, c->stream, nullptr, l.tab));
    } else if (l.kind == 1) {
      HIPCHK(launch_generic_obs(p));
    }
  }
  p->stream_next = g1;
  return QG_OK;
}

// Notes the kinds of a profiled run's launches (its events were recorded by
// run_launches); the times are read in qg_plan_pass_times, so a profiled run
// costs its event records only.
static int collect_profile(qg_plan* p) {
  if (!p->profile) return QG_OK;
  const size_t nl = p->launches.size();
  const size_t slot = p->prof_runs % kMaxProfRuns;
  if (p->prof_kind.size() <= slot) p->prof_kind.resize(slot + 1);
  std::vector<int>& kinds = p->prof_kind[slot];
  kinds.assign(nl, 0);
  for (size_t k = 0; k < nl; ++k) {
    if (p->launches[k].kind != 0) {
      kinds[k] = 3;  // not a pass kernel
      continue;
    }
    const PassDesc& d = p->h_desc[p->launches[k].desc];
    kinds[k] = (d.flags & F_REDUCE) ? 1 : (d.flags & F_OBS_DIAG) ? 2 : 0;
  }
  ++p->prof_runs;
  return QG_OK;
}

extern "C" int qg_plan_set_profiling(qg_plan* p, int on) {
  if (!p) return fail(QG_ERR_INVALID, "NULL plan");
  p->profile = on != 0;
  p->prof_runs = 0;
  p->pass_ms.clear();
  p->pass_kind.clear();
  return QG_OK;
}

extern "C" int qg_plan_pass_times(const qg_plan* p, int cap, float* ms, int* kinds, int* count) {
  if (!p || !count) return fail(QG_ERR_INVALID, "NULL argument");
  if (p->prof_runs > 0) {
    HIPCHK(hipSetDevice(p->ctx->device));
    p->pass_ms.clear();
    p->pass_kind.clear();
    const size_t first = p->prof_runs > kMaxProfRuns ? p->prof_runs - kMaxProfRuns : 0;
    for (size_t r = first; r < p->prof_runs; ++r) {  // oldest kept run first
      const std::vector<hipEvent_t>& ev = p->prof_ev[r % kMaxProfRuns];
      const std::vector<int>& kd = p->prof_kind[r % kMaxProfRuns];
      HIPCHK(hipEventSynchronize(ev[kd.size()]));
      for (size_t k = 0; k < kd.size(); ++k) {
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
        p->pass_ms.push_back(t);
        p->pass_kind.push_back(kd[k]);
      }
    }
  }
  const int n = (int)p->pass_ms.size();
  *count = n;
  for (int k = 0; k < n && k < cap; ++k) {
    if (ms) ms[k] = p->pass_ms[k];
    if (kinds) kinds[k] = p->pass_kind[k];
  }
  return QG_OK;
}

// ---------------------------------------------------------------------------
// McClean
// ---------------------------------------------------------------------------
static void free_small(qg_plan* p) {
  drop_graph(p);  // its launches point into the descriptor table freed here
  (void)hipFree(p->d_cterms);
  p->d_cterms = nullptr;
  (void)hipFree(p->d_ini_states);
  p->d_ini_states = nullptr;
  p->ini_cap = 0;
  p->cterm_cap = 0;
  (void)hipFree(p->d_init);
  (void)hipFree(p->d_terms);
  (void)hipFree(p->d_xy);
  (void)hipFree(p->d_lut);
  (void)hipFree(p->d_part);
  (void)hipFree(p->d_fin);
  (void)hipFree(p->d_scale);
  (void)hipFree(p->d_out);
  (void)hipFree(p->d_dot);
  (void)hipFree(p->d_desc);
  if (p->h_desc) (void)hipHostFree(p->h_desc);
  if (p->h_out) (void)hipHostFree(p->h_out);
  if (p->h_arena) (void)hipHostFree(p->h_arena);
  if (p->h_stage) (void)hipHostFree(p->h_stage);
  if (p->h_cstage) (void)hipHostFree(p->h_cstage);
  p->h_cstage = nullptr;
  p->cstage_cap = 0;
  (void)hipFree(p->d_arena);
  (void)hipFree(p->d_fstate);
  p->h_stage = nullptr;
  p->stage_cap = 0;
  p->h_arena = nullptr;
  p->d_arena = nullptr;
  p->d_fstate = nullptr;
  p->arena_cap = 0;
  p->d_init = nullptr;
  p->d_terms = p->d_xy = nullptr;
  p->d_lut = p->d_part = p->d_scale = p->d_out = p->d_dot = nullptr;
  p->d_fin = nullptr;
  p->d_desc = nullptr;
  p->h_desc = nullptr;
  p->h_out = nullptr;
  p->lut_len = 0;
  p->fin_cap = 0;
  if (p->h_tabs) (void)hipHostFree(p->h_tabs);
  p->h_tabs = nullptr;
  p->cap_n = p->cap_terms = p->cap_nout = p->ndesc_cap = 0;
  p->cap_part = p->cap_fstate = 0;
}

// Memory a plan's layout may assume: this fraction of the device's total
// memory.  The layout (history or checkpoints, the period, the tile size) is a
// function of the circuit shape and the device alone, never of what happens
// to be free, so one shape always gives the same numerics; a workspace that
// cannot then be allocated is an error (QG_ERR_NOMEM), not a different mode.
constexpr double kPlanMemFraction = 0.75;

// Whether the automatic FP64-tile rule (circuits of more than 200 passes)
// covers the adjoint passes as well as the forward passes.  Measured at 30
// qubits, 100 layers, checkpoints every 5 layers: FP64 tiles everywhere 5.78 s
// and 6.2e-7 gradient error against complex128, forward only 5.44 s and 7.9e-7,
// none 5.42 s and 1.1e-6.  The forward roundings dominate the error.
constexpr bool kLds64Adjoint = false;

static int plan_budget(qg_plan* p, size_t* budget) {
  *budget = (size_t)(kPlanMemFraction * (double)p->ctx->total_mem);
  return QG_OK;
}

// Geometry, memory mode (history or checkpointed uncompute) and workspace
// layout of a McClean plan: a function of the shape, the tile options and the
// device's total memory only (no device call), so the footprint of a shape
// can be computed without a plan.
static int mcclean_geometry(qg_plan* p) {
  const int n = p->n, L = p->L;
  if (p->kmax_auto) p->kmax = default_kmax(n, p->dtype, p->cmin);
  p->geo = layer_geometry(n, p->kmax, p->cmin);
  int P = (int)p->geo.size();
  size_t budget = 0;
  CHK(plan_budget(p, &budget));
  if (p->kmax > 13) {
    // 14 local bits run only as complex64 history passes with FP32 tiles
    // (1024-thread variants exist for those); otherwise fall back to 13
    const size_t S = (size_t)(1ull << n) * 8;
    const bool hist_fits = p->ckpt_req <= 0 && (size_t)(P * L + 4) * S <= budget;
    const bool fp64_tiles = p->lds64_req >= 1 || (p->lds64_req < 0 && P * L > 200);
    if (p->dtype != QG_C64 || !hist_fits || fp64_tiles) {
      p->kmax = p->kmax_auto ? default_kmax(n, p->dtype, p->cmin, false) : 13;
      p->geo = layer_geometry(n, p->kmax, p->cmin);
      P = (int)p->geo.size();
    }
  }
  p->S = (size_t)(1ull << n) * (p->dtype == QG_C64 ? 8 : 16);
  // complex64 states are rounded at every LDS transpose of a pass with FP32
  // tiles; beyond about 200 passes these roundings would take the gradient
  // past a 1e-6 relative error, so long circuits use FP64 tiles (rounding
  // once per pass, at the store)
  p->lds64 = p->dtype == QG_C64 && n >= 4 && (p->lds64_req >= 1 || (p->lds64_req < 0 && P * L > 200));
  p->lds64_adj = p->lds64 && (p->lds64_req == 1 || (p->lds64_req < 0 && kLds64Adjoint));
  const size_t hist_bytes = (size_t)(P * L + 4) * p->S;
  if (p->ckpt_req == QG_HISTORY && hist_bytes > budget)
    return fail(QG_ERR_NOMEM, "the per-pass history does not fit in device memory");
  if (p->ckpt_req == QG_HISTORY || (p->ckpt_req == 0 && hist_bytes <= budget)) {
    // every pass output is kept: the adjoint sweep reads psi, never recomputes it
    p->history = true;
    p->off_hist = 0;
    p->off_lam = (size_t)P * L * p->S;
  } else {
    // psi is uncomputed in the adjoint sweep and restarted from a stored
    // checkpoint at the end of every k-th layer (and of the last layer)
    int k = p->ckpt_req;
    if (k == 0) {
      for (k = 1; k <= L; ++k)
        if ((size_t)(6 + (L + k - 1) / k) * p->S <= budget) break;
      if (k > L) k = L;
    }
    p->history = false;
    p->ckpt = std::min(k, L);
    p->nckpt = 0;
    for (int i = 0; i < L; ++i)
      if ((i + 1) % p->ckpt == 0 || i == L - 1) p->nckpt++;
    if ((size_t)(6 + p->nckpt) * p->S > budget)
      return fail(QG_ERR_NOMEM, "circuit does not fit in device memory even with one checkpoint");
    p->off_work = 0;
    p->off_ckpt = 2 * p->S;
    p->off_lam = p->off_ckpt + (size_t)p->nckpt * p->S;
  }
  p->off_ini = p->off_lam + 2 * p->S;
  p->off_prep = p->off_ini + p->S;
  p->ws_bytes = p->off_prep + p->S;
  return QG_OK;
}

// mcclean_geometry plus the partial-row stride and the small device tables.
static int mcclean_layout(qg_plan* p) {
  CHK(mcclean_geometry(p));
  const int n = p->n, L = p->L, P = (int)p->geo.size();
  p->gmax = grid_max(p);
  const int npass = 2 * P * L + 2 * P + 1;  // forward + reverse + ini-state preparation + observable
  p->nrows = (P * L + 1) * (kMaxK + 1);
  CHK(ensure_small(p, npass, p->nrows, 1 + L * n));
  // static tables, staged in pinned memory and copied in stream order (the
  // previous run on this stream ended with a synchronisation, so the staging
  // is free and the tables are no longer read)
  if (!ladder_shift_form(n, 1, &p->lad1, &p->lad2)) return fail(QG_ERR_INVALID, "ladder map not in shift form");
  double* hi = (double*)p->h_tabs;
  ry_init_table(n, hi);
  HIPCHK(hipMemcpyAsync(p->d_init, hi, sizeof(double) * (n + 1), hipMemcpyHostToDevice, p->ctx->stream));
  HIPCHK(hipStreamSynchronize(p->ctx->stream));
  return QG_OK;
}

extern "C" int qg_mcclean_plan(qg_ctx* ctx, int n, int L, int dtype, int entangler, int ckpt_every,
                               qg_plan** out) {
  if (!ctx || !out) return fail(QG_ERR_INVALID, "NULL argument");
  if (n < 1 || n > kMaxQubits) return fail(QG_ERR_INVALID, "qubit number must be in [1, 32]");
  if (L < 1) return fail(QG_ERR_INVALID, "layer number must be >= 1");
  if (dtype != QG_C128 && dtype != QG_C64) return fail(QG_ERR_INVALID, "dtype must be QG_C128 or QG_C64");
  if (entangler != QG_ENT_CNOT && entangler != QG_ENT_CZ) return fail(QG_ERR_INVALID, "unknown entangler");
  if (ckpt_every < QG_HISTORY) return fail(QG_ERR_INVALID, "ckpt_every must be >= 0 or QG_HISTORY");
  HIPCHK(hipSetDevice(ctx->device));
  qg_plan* p = new qg_plan();
  p->ctx = ctx;
  p->kind = PK_MCCLEAN;
  p->n = n;
  p->L = L;
  p->dtype = dtype;
  p->ent = entangler;
  p->ckpt_req = ckpt_every;
  int r = mcclean_layout(p);
  if (r != QG_OK) {
    qg_plan_destroy(p);
    return r;
  }
  *out = p;
  return QG_OK;
}

extern "C" int qg_mcclean_workspace_bytes(qg_ctx* ctx, int n, int L, int dtype, int ckpt_every, size_t* bytes,
                                          int* ckpt_chosen) {
  if (!bytes) return fail(QG_ERR_INVALID, "NULL argument");
  if (n < 1 || n > kMaxQubits) return fail(QG_ERR_INVALID, "qubit number must be in [1, 32]");
  if (L < 1) return fail(QG_ERR_INVALID, "layer number must be >= 1");
  if (dtype != QG_C128 && dtype != QG_C64) return fail(QG_ERR_INVALID, "dtype must be QG_C128 or QG_C64");
  if (ckpt_every < QG_HISTORY) return fail(QG_ERR_INVALID, "ckpt_every must be >= 0 or QG_HISTORY");
  if (!ctx && ckpt_every == 0)
    return fail(QG_ERR_INVALID, "without a context the memory mode must be given (QG_HISTORY or k > 0)");
  qg_ctx none;
  none.total_mem = std::numeric_limits<size_t>::max() / 2;  // no device: no memory bound
  qg_plan p;
  p.ctx = ctx ? ctx : &none;
  p.kind = PK_MCCLEAN;
  p.n = n;
  p.L = L;
  p.dtype = dtype;
  p.ckpt_req = ckpt_every;
  CHK(mcclean_geometry(&p));
  *bytes = n <= kFusedMaxN ? 0 : p.ws_bytes;  // whole-circuit runs need no workspace
  if (ckpt_chosen) *ckpt_chosen = p.history ? 0 : p.ckpt;
  return QG_OK;
}

// Builds the pass list of one McClean run into p->h_desc / p->launches and
// the finalize tables; returns the descriptor count.
// streamed: passes go to the device as they are built (stream_begin has been
// called); otherwise the caller launches the list with run_launches.
static int mcclean_schedule(qg_plan* p, const int32_t* axes, const double* angles, const ObsSplit& obs,
                            bool with_ini, bool grad, int& ndesc_out, bool streamed) {
  const int n = p->n, L = p->L, P = (int)p->geo.size();
  int flush_rc = QG_OK;
  const bool diag_obs = obs.xy.empty();
  p->launches.clear();
  p->obs_desc = p->adj_desc = -1;
  int nd = 0;
  int row = 0;
  auto hist = [&](int i, int g) { return wsp(p, p->off_hist + ((size_t)i * P + g) * p->S); };
  void* lam[2] = {wsp(p, p->off_lam), wsp(p, p->off_lam + p->S)};
  void* work[2] = {wsp(p, p->off_work), wsp(p, p->off_work + p->S)};
  std::vector<int> slot(L, -1);
  if (!p->history) {
    int m = 0;
    for (int i = 0; i < L; ++i)
      if ((i + 1) % p->ckpt == 0 || i == L - 1) slot[i] = m++;
  }
  auto ckpt = [&](int i) { return wsp(p, p->off_ckpt + (size_t)slot[i] * p->S); };
  auto other = [&](const void* b) { return b == work[0] ? work[1] : work[0]; };
  const uint32_t ent_flag_fwd = p->ent == QG_ENT_CNOT ? F_GATHER : F_CZ_PRE;
  const uint32_t ent_flag_rev = p->ent == QG_ENT_CNOT ? F_SCATTER : F_CZ_POST;

  auto push = [&](const Geo& g0, bool two, int role) -> PassDesc& {
    // every launch before this one is complete.  They go out in doubling
    // groups (2, 2, 4, 8, ...): the device starts after the first two, and
    // each later group holds more device work than the host takes to build
    // the next, while a grad_run sends about log2(passes) descriptor copies
    // (a copy between every two 50-100 us passes cost more than it hid)
    if (streamed && flush_rc == QG_OK && p->launches.size() - p->stream_next >= std::max<size_t>(2, p->stream_next))
      flush_rc = stream_flush(p);
    bool tf = false;
    const Geo g = with_role(g0, p->dtype, role, &tf);
    Launch l;
    l.desc = nd;
    l.R = g.R;
    l.K = g.K;
    l.post = tf;
    l.threads = g.threads;
    l.grid = grid_for_pass(p, g, two, tf, 0);
    l.phase = 0;
    l.two = two;
    l.kind = 0;
    p->launches.push_back(l);
    PassDesc& d = p->h_desc[nd++];
    desc_geo(d, n, g);
    d.grid = p->gmax;  // stride of the partial rows; the launch grid is l.grid
    d.lad1 = p->lad1;
    d.lad2 = p->lad2;
    d.init_tab = p->d_init;
    d.dterms = p->d_terms;
    d.ndterms = (int)obs.diag.size();
    d.part = p->d_part;
    return d;
  };
  auto gates = [&](PassDesc& d, const Geo& G, int i, double sign, bool own) {
    for (int lb = G.own_lo; lb < G.own_hi; ++lb) {
      const int b = (lb < G.c) ? lb : G.p + (lb - G.c);
      const int q = n - 1 - b;
      set_gate(d, lb, axes[i * n + q], angles[i * n + q], sign);
      if (own) d.own[lb] = 1;
    }
  };

  // optional ini_state preparation: RY(pi/4) on every qubit of the given vector
  const void* cur = nullptr;
  if (with_ini) {
    const void* src = wsp(p, p->off_ini);
    for (int g = 0; g < P; ++g) {
      const Geo& G = p->geo[g];
      PassDesc& d = push(G, false, ROLE_FWD);
      d.flags = F_LOAD_A | F_ROT_A | F_STORE_A;
      d.a_in = src;
      d.a_out = wsp(p, p->off_prep);
      for (int lb = G.own_lo; lb < G.own_hi; ++lb) set_gate(d, lb, QG_AXIS_Y, M_PI / 4, 1.0);
      src = wsp(p, p->off_prep);
    }
    cur = wsp(p, p->off_prep);
  }
  // forward sweep
  for (int i = 0; i < L; ++i) {
    for (int g = 0; g < P; ++g) {
      const Geo& G = p->geo[g];
      const bool last = (i == L - 1) && (g == P - 1);
      // the observable rides on the last forward pass, except after a
      // 1024-thread pass (no such variant of the general two-state form)
      const bool obs_fused = last && diag_obs && G.K <= 13;
      if (obs_fused) p->obs_desc = nd;
      PassDesc& d = push(G, obs_fused, obs_fused ? ROLE_TWO_FP64 : ROLE_FWD);
      uint32_t f = F_ROT_A | F_STORE_A;
      if (g == 0) {
        f |= ent_flag_fwd;
        f |= cur ? F_LOAD_A : F_INIT_TAB;
      } else {
        f |= F_LOAD_A;
      }
      void* out;
      if (p->history) {
        out = hist(i, g);
        // a fresh history buffer per pass: its input is read once here
        if ((f & F_LOAD_A) && ((uint64_t)1 << n) <= kNtMaxLaunchAmps) f |= F_NT_A;
      } else if (g == P - 1 && slot[i] >= 0) {
        out = ckpt(i);                       // checkpointed layer end
      } else if (g == 0) {
        out = other(cur);                    // the low pass gathers: out of place
      } else {
        out = (void*)cur;                    // high passes work in place
      }
      if (obs_fused) {
        f |= F_OBS_DIAG | F_STORE_B;
        d.b_out = lam[0];
        d.row_base = row;
        row += kMaxK + 1;
      }
      d.flags = f;
      d.a_in = cur;
      d.a_out = out;
      gates(d, G, i, 1.0, false);
      cur = out;
    }
  }
  if (!diag_obs) {
    // x / y terms: lambda = O psi and E by the elementwise kernels
    Launch l;
    l.desc = -1;
    l.R = l.K = l.threads = l.grid = l.phase = l.tab = 0;
    l.post = l.two = false;
    l.batch = 1;
    l.kind = 1;
    p->launches.push_back(l);
    p->gen_in = cur;
    p->gen_out = lam[0];
    p->gen_ndiag = (int)obs.diag.size();
    p->gen_nxy = (int)obs.xy.size();
  }
  if (p->obs_desc < 0 && diag_obs) {
    // lambda = D psi and E in a gate-free pass of its own
    const Geo og = obs_geo(n);
    p->obs_desc = nd;
    PassDesc& d = push(og, true, ROLE_TWO_FP64);
    d.flags = F_LOAD_A | F_OBS_DIAG | F_STORE_B;
    d.a_in = cur;
    d.b_out = lam[0];
    d.row_base = row;
    row += kMaxK + 1;
  }
  p->last_state_ptr = (void*)cur;
  p->state_sign = false;
  for (int k = 0; k < L * n; ++k) p->state_sign ^= gate_flips(angles[k], 1.0);
  p->adj_desc = nd;
  if (!grad) {
    ndesc_out = nd;
    return flush_rc;
  }
  // adjoint sweep
  int lc = 0;
  const void* acur = cur;
  for (int i = L - 1; i >= 0; --i) {
    for (int g = P - 1; g >= 0; --g) {
      const Geo& G = p->geo[g];
      PassDesc& d = push(G, true, p->history ? ROLE_REV_HIST : ROLE_TWO_FP64);
      uint32_t f = F_LOAD_A | F_ROT_A | F_LOAD_B | F_ROT_B | F_STORE_B | F_REDUCE;
      d.b_in = lam[lc];
      if (g == 0 && p->ent == QG_ENT_CNOT) {
        d.b_out = lam[lc ^ 1];
        lc ^= 1;
      } else {
        d.b_out = lam[lc];
      }
      if (p->history) {
        d.a_in = hist(i, g);
        if (((uint64_t)1 << n) <= kNtMaxLaunchAmps) f |= F_NT_A;  // read once here
      } else {
        const void* ain = (g == P - 1 && slot[i] >= 0) ? ckpt(i) : acur;
        void* aout = (g == 0) ? other(ain) : (void*)ain;
        d.a_in = ain;
        d.a_out = aout;
        f |= F_STORE_A;
        acur = aout;
      }
      if (g == 0) f |= ent_flag_rev;
      d.flags = f;
      d.row_base = row;
      gates(d, G, i, -1.0, true);
      row += kMaxK + 1;
    }
  }
  ndesc_out = nd;
  return flush_rc;
}

static int upload_finalize(qg_plan* p, const std::vector<std::vector<int>>& out_rows, const std::vector<int>& cnt,
                           const std::vector<double>& scale) {
  const int nout = (int)out_rows.size();
  std::vector<int> buf;
  buf.push_back(0);
  for (int i = 0; i < nout; ++i) buf.push_back(buf.back() + (int)out_rows[i].size());
  for (int i = 0; i < nout; ++i)
    for (int r : out_rows[i]) buf.push_back(r);
  const int rows_off = nout + 1;
  const int cnt_off = (int)buf.size();
  for (int r = 0; r < p->nrows; ++r) buf.push_back(r < (int)cnt.size() ? cnt[r] : 0);
  if (buf.size() > p->fin_cap) {
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
    (void)hipFree(p->d_fin);
    p->d_fin = nullptr;
    HIPCHK(hipMalloc(&p->d_fin, sizeof(int) * buf.size()));
    p->fin_cap = buf.size();
  }
  HIPCHK(hipMemcpyAsync(p->d_fin, buf.data(), sizeof(int) * buf.size(), hipMemcpyHostToDevice, p->ctx->stream));
  HIPCHK(hipMemcpyAsync(p->d_scale, scale.data(), sizeof(double) * nout, hipMemcpyHostToDevice, p->ctx->stream));
  HIPCHK(launch_finalize(nout, p->d_fin, p->d_fin + rows_off, p->d_fin + cnt_off, p->d_scale, p->d_part, p->gmax,
                         p->d_out, p->ctx->stream));
  // synchronous because the host vectors above die at return
  HIPCHK(hipStreamSynchronize(p->ctx->stream));
  return QG_OK;
}


// ---------------------------------------------------------------------------
// per-call arena: tables built in pinned host memory, copied with one H2D
// ---------------------------------------------------------------------------
struct Arena {
  qg_plan* p;
  size_t used = 0;
  // returns the host pointer; *dev receives the device address of the same bytes
  template <class T>
  T* take(size_t count, T** dev) {
    used = (used + 255) & ~(size_t)255;
    T* h = (T*)(p->h_arena + used);
    *dev = (T*)(p->d_arena + used);
    used += sizeof(T) * count;
    return h;
  }
};

static int ensure_arena(qg_plan* p, size_t bytes) {
  if (bytes <= p->arena_cap) return QG_OK;
  HIPCHK(hipStreamSynchronize(p->ctx->stream));
  if (p->h_arena) (void)hipHostFree(p->h_arena);
  (void)hipFree(p->d_arena);
  p->h_arena = nullptr;
  p->d_arena = nullptr;
  p->arena_cap = 0;
  HIPCHK(hipHostMalloc(&p->h_arena, bytes, hipHostMallocDefault));
  HIPCHK(hipMalloc(&p->d_arena, bytes));
  p->arena_cap = bytes;
  return QG_OK;
}

// Runs B circuits through the whole-circuit kernel.  The tables of circuit b
// are written by fill(b, desc, arena); results land in out[b * per_out ...].
template <class Fill>
static int run_fused(qg_plan* p, int B, size_t bytes_per_circuit, int per_out, bool grad, void* psi_out, Fill fill,
                     double* out) {
  qg_ctx* c = p->ctx;
  const size_t need = (size_t)B * (bytes_per_circuit + 16 * 256 + sizeof(FusedDescH) + sizeof(double) * per_out);
  CHK(ensure_arena(p, need));
  Arena ar{p};
  FusedDescH* d_desc = nullptr;
  FusedDescH* h_desc = ar.take<FusedDescH>(B, &d_desc);
  double* d_out = nullptr;
  ar.take<double>((size_t)B * per_out, &d_out);
  for (int b = 0; b < B; ++b) {
    std::memset(&h_desc[b], 0, sizeof(FusedDescH));
    h_desc[b].n = p->n;
    h_desc[b].L = p->L;
    h_desc[b].ent = p->ent;
    h_desc[b].lad1 = p->lad1;
    h_desc[b].lad2 = p->lad2;
    h_desc[b].init_tab = p->d_init;
    h_desc[b].out = d_out + (size_t)b * per_out;
    h_desc[b].grad = grad ? 1 : 0;
    h_desc[b].psi_c64 = p->dtype == QG_C64 ? 1 : 0;
    h_desc[b].psi_out = psi_out;
    CHK(fill(b, h_desc[b], ar));
  }
  HIPCHK(hipMemcpyAsync(p->d_arena, p->h_arena, ar.used, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  HIPCHK(launch_fused(d_desc, B, p->n, c->stream));
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  const int nout = grad ? per_out : 1;
  for (int b = 0; b < B && !grad; ++b)
    HIPCHK(hipMemcpyAsync(out + (size_t)b * per_out, d_out + (size_t)b * per_out, sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  if (grad)
    HIPCHK(hipMemcpyAsync(out, d_out, sizeof(double) * (size_t)B * nout, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipEventElapsedTime(&p->last_ms, c->ev0, c->ev1));
  p->launches.clear();
  return QG_OK;
}

static int obs_tables(const ObsSplit& obs, FusedDescH& d, Arena& ar) {
  DTerm* dd = nullptr;
  DTerm* hd = ar.take<DTerm>(std::max<size_t>(1, obs.diag.size()), &dd);
  std::copy(obs.diag.begin(), obs.diag.end(), hd);
  DTerm* dx = nullptr;
  DTerm* hx = ar.take<DTerm>(std::max<size_t>(1, obs.xy.size()), &dx);
  std::copy(obs.xy.begin(), obs.xy.end(), hx);
  d.dterms = dd;
  d.ndterms = (int)obs.diag.size();
  d.xy = dx;
  d.nxy = (int)obs.xy.size();
  return QG_OK;
}

static int mcclean_run(qg_plan* p, const int32_t* axes, const double* angles, const qg_term* terms, int nterms,
                       const double* ini_state, double* E_out, double* grad_out, bool grad) {
  if (!p || p->kind != PK_MCCLEAN) return fail(QG_ERR_INVALID, "not a McClean plan");
  if (p->mc_batch) return fail(QG_ERR_INVALID, "a batch plan runs through qg_mcclean_batch_grad_run");
  if (!axes || !angles || !E_out || (grad && !grad_out)) return fail(QG_ERR_INVALID, "NULL argument");
  const int n = p->n, L = p->L;
  for (int k = 0; k < L * n; ++k) {
    if (axes[k] < 0 || axes[k] > 2) return fail(QG_ERR_INVALID, "Invalid axis " + std::to_string(axes[k]));
    if (!std::isfinite(angles[k])) return fail(QG_ERR_INVALID, "non-finite angle");
  }
  ObsSplit obs;
  CHK(split_observable(n, terms, nterms, obs));
  if ((int)obs.diag.size() > p->max_terms || (int)obs.xy.size() > p->max_terms)
    return fail(QG_ERR_INVALID, "too many observable terms");
  qg_ctx* c = p->ctx;
  HIPCHK(hipSetDevice(c->device));
  if (p->fused && n <= kFusedMaxN && !ini_state) {
    // the whole circuit in one workgroup (qg_fused.hip): gates as (type, cos, sin)
    std::vector<double> res(1 + (size_t)L * n);
    auto fill = [&](int, FusedDescH& d, Arena& ar) -> int {
      d.kind = 0;
      int32_t* dt = nullptr;
      int32_t* ht = ar.take<int32_t>((size_t)L * n, &dt);
      double* dg = nullptr;
      double* hg = ar.take<double>((size_t)2 * L * n, &dg);
      for (int k = 0; k < L * n; ++k) {
        ht[k] = axes[k] == 0 ? G_X : axes[k] == 1 ? G_Y : G_Z;
        hg[2 * k] = std::cos(0.5 * angles[k]);
        hg[2 * k + 1] = std::sin(0.5 * angles[k]);
      }
      d.gtype = dt;
      d.gcs = dg;
      return obs_tables(obs, d, ar);
    };
    const size_t bytes = (size_t)L * n * 20 + (obs.diag.size() + obs.xy.size() + 2) * sizeof(DTerm);
    if (p->cap_fstate < ((size_t)16 << n)) {
      HIPCHK(hipStreamSynchronize(c->stream));
      (void)hipFree(p->d_fstate);
      p->d_fstate = nullptr;
      p->cap_fstate = 0;
      HIPCHK(hipMalloc(&p->d_fstate, ((size_t)16) << n));
      p->cap_fstate = (size_t)16 << n;
    }
    CHK(run_fused(p, 1, bytes, 1 + L * n, grad, p->d_fstate, fill, res.data()));
    *E_out = res[0];
    if (grad) std::memcpy(grad_out, res.data() + 1, sizeof(double) * L * n);
    p->last_state_ptr = p->d_fstate;
    p->state_sign = false;
    return QG_OK;
  }
  if (!obs.xy.empty() && n > 31) return fail(QG_ERR_UNSUPPORTED, "observables with x/y terms need n <= 31");
  CHK(ensure_workspace(p));
  if (!obs.diag.empty())
    HIPCHK(hipMemcpyAsync(p->d_terms, obs.diag.data(), sizeof(DTerm) * obs.diag.size(), hipMemcpyHostToDevice,
                          c->stream));
  if (!obs.xy.empty())
    HIPCHK(hipMemcpyAsync(p->d_xy, obs.xy.data(), sizeof(DTerm) * obs.xy.size(), hipMemcpyHostToDevice,
                          c->stream));
  p->h_terms = obs.diag;
  if (ini_state) {
    const size_t N = 1ull << n;
    double* tmp = nullptr;
    HIPCHK(hipMalloc(&tmp, N * 16));
    hipError_t e = hipMemcpyAsync(tmp, ini_state, N * 16, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = simple_convert_in(p->dtype, tmp, wsp(p, p->off_ini), n, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(QG_ERR_HIP, std::string("ini_state upload: ") + hipGetErrorString(e));
  }
  int nd = 0;
  // passes go to the device as they are built, so the device starts while
  // the host builds the rest (graph replay and profiled runs launch the list
  // at the end); only the observable's pass splits terms
  const bool streamed = !p->use_graph && !p->profile;
  if (streamed) CHK(stream_begin(p, std::max<size_t>(1, obs.diag.size())));
  CHK(mcclean_schedule(p, axes, angles, obs, ini_state != nullptr, grad, nd, streamed));
  // uncompute passes of a checkpointed adjoint overwrite the final state
  if (grad && !p->history) p->last_state_ptr = nullptr;
  // rows written by each pass and the finalize table
  const int P = (int)p->geo.size();
  std::vector<std::vector<int>> out_rows(grad ? 1 + L * n : 1);
  std::vector<int> cnt(p->nrows, 0);
  for (const Launch& l : p->launches) {
    if (l.kind != 0) continue;  // no descriptor
    const PassDesc& d = p->h_desc[l.desc];
    if (d.row_base < 0) continue;
    for (int k = 0; k <= kMaxK; ++k) cnt[d.row_base + k] = l.grid;
  }
  {
    // E: row K of the observable pass (none with x / y terms: from d_dot)
    if (p->obs_desc >= 0) {
      const PassDesc& d = p->h_desc[p->obs_desc];
      out_rows[0].push_back(d.row_base + d.K);
    }
    if (grad) {
      int idx = p->adj_desc;
      for (int i = L - 1; i >= 0; --i) {
        for (int g = P - 1; g >= 0; --g, ++idx) {
          const PassDesc& dr = p->h_desc[idx];
          const Geo& G = p->geo[g];
          for (int lb = G.own_lo; lb < G.own_hi; ++lb) {
            const int b = (lb < G.c) ? lb : G.p + (lb - G.c);
            const int q = n - 1 - b;
            out_rows[1 + i * n + q].push_back(dr.row_base + lb);
          }
        }
      }
    }
  }
  std::vector<double> scale(out_rows.size(), 1.0);
  if (streamed) CHK(stream_flush(p));
  else CHK(run_launches(p, nd));
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  CHK(collect_profile(p));
  CHK(upload_finalize(p, out_rows, cnt, scale));
  HIPCHK(hipEventElapsedTime(&p->last_ms, c->ev0, c->ev1));
  const int nout = (int)out_rows.size();
  HIPCHK(hipMemcpyAsync(p->h_out, p->d_out, sizeof(double) * nout, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *E_out = p->h_out[0];
  if (p->obs_desc < 0) HIPCHK(hipMemcpy(E_out, p->d_dot + kDotBlocks, sizeof(double), hipMemcpyDeviceToHost));
  if (grad) std::memcpy(grad_out, p->h_out + 1, sizeof(double) * L * n);
  return QG_OK;
}

extern "C" int qg_mcclean_grad_run(qg_plan* p, const int32_t* axes, const double* angles, const qg_term* terms,
                                   int nterms, const double* ini_state, double* E_out, double* grad_out) {
  return mcclean_run(p, axes, angles, terms, nterms, ini_state, E_out, grad_out, true);
}

extern "C" int qg_mcclean_expec_val(qg_plan* p, const int32_t* axes, const double* angles, const qg_term* terms,
                                    int nterms, const double* ini_state, double* E_out) {
  return mcclean_run(p, axes, angles, terms, nterms, ini_state, E_out, nullptr, false);
}

// ---------------------------------------------------------------------------
// Batched McClean: B circuits of one shape (their own axes and angles, one
// observable), every pass one launch over the circuits of a chunk
// ---------------------------------------------------------------------------
// The tile shape, memory mode (the per-pass history) and per-circuit grid are
// those of a single plan of the shape, so each circuit's result is bitwise
// the one a single plan gives.
static int mcclean_batch_layout(qg_plan* p) {
  const int n = p->n, L = p->L;
  p->kmax = default_kmax(n, p->dtype, p->cmin);
  p->geo = layer_geometry(n, p->kmax, p->cmin);
  const int P = (int)p->geo.size();
  p->S = (size_t)(1ull << n) * (p->dtype == QG_C64 ? 8 : 16);
  p->lds64 = p->dtype == QG_C64 && n >= 4 && P * L > 200;
  p->lds64_adj = p->lds64 && kLds64Adjoint;
  p->history = true;
  size_t budget = 0;
  CHK(plan_budget(p, &budget));
  const size_t per = (size_t)(P * L + 2) * p->S;  // history of every pass + two lambda buffers
  if (n > kFusedMaxN) {
    if (per > budget) return fail(QG_ERR_NOMEM, "one circuit's per-pass history does not fit in device memory");
    p->chunk = (int)std::min<size_t>((size_t)p->mc_batch, budget / per);
  } else {
    p->chunk = p->mc_batch;  // one workgroup per circuit, no workspace
  }
  p->off_hist = 0;
  p->off_lam = (size_t)p->chunk * P * L * p->S;
  p->ws_bytes = n > kFusedMaxN ? (size_t)p->chunk * per : 0;
  p->gmax = grid_max(p);
  const int ch = n > kFusedMaxN ? p->chunk : 1;
  p->nrows = ch * (P * L + 1) * (kMaxK + 1);
  CHK(ensure_small(p, ch * 2 * P * L, p->nrows, ch * (1 + L * n)));
  if (!ladder_shift_form(n, 1, &p->lad1, &p->lad2)) return fail(QG_ERR_INVALID, "ladder map not in shift form");
  double* hi = (double*)p->h_tabs;
  ry_init_table(n, hi);
  HIPCHK(hipMemcpyAsync(p->d_init, hi, sizeof(double) * (n + 1), hipMemcpyHostToDevice, p->ctx->stream));
  HIPCHK(hipStreamSynchronize(p->ctx->stream));
  return QG_OK;
}

extern "C" int qg_mcclean_batch_plan(qg_ctx* ctx, int n, int L, int dtype, int entangler, int batch,
                                     qg_plan** out) {
  if (!ctx || !out) return fail(QG_ERR_INVALID, "NULL argument");
  if (n < 1 || n > kMaxQubits) return fail(QG_ERR_INVALID, "qubit number must be in [1, 32]");
  if (L < 1) return fail(QG_ERR_INVALID, "layer number must be >= 1");
  if (dtype != QG_C128 && dtype != QG_C64) return fail(QG_ERR_INVALID, "dtype must be QG_C128 or QG_C64");
  if (entangler != QG_ENT_CNOT && entangler != QG_ENT_CZ) return fail(QG_ERR_INVALID, "unknown entangler");
  if (batch < 1) return fail(QG_ERR_INVALID, "batch must be >= 1");
  HIPCHK(hipSetDevice(ctx->device));
  qg_plan* p = new qg_plan();
  p->ctx = ctx;
  p->kind = PK_MCCLEAN;
  p->n = n;
  p->L = L;
  p->dtype = dtype;
  p->ent = entangler;
  p->ckpt_req = QG_HISTORY;
  p->mc_batch = batch;
  const int r = mcclean_batch_layout(p);
  if (r != QG_OK) {
    qg_plan_destroy(p);
    return r;
  }
  *out = p;
  return QG_OK;
}

extern "C" int qg_mcclean_batch_grad_run(qg_plan* p, int B, const int32_t* axes, const double* angles,
                                         const qg_term* terms, int nterms, double* E_out, double* grad_out) {
  if (!p || p->kind != PK_MCCLEAN || !p->mc_batch) return fail(QG_ERR_INVALID, "not a McClean batch plan");
  if (!axes || !angles || !E_out || !grad_out) return fail(QG_ERR_INVALID, "NULL argument");
  if (B < 1 || B > p->mc_batch) return fail(QG_ERR_INVALID, "B must be in [1, the plan's batch]");
  const int n = p->n, L = p->L, P = (int)p->geo.size(), per_out = 1 + L * n;
  const size_t np = (size_t)L * n;
  for (size_t k = 0; k < (size_t)B * np; ++k) {
    if (axes[k] < 0 || axes[k] > 2) return fail(QG_ERR_INVALID, "Invalid axis " + std::to_string(axes[k]));
    if (!std::isfinite(angles[k])) return fail(QG_ERR_INVALID, "non-finite angle");
  }
  ObsSplit obs;
  CHK(split_observable(n, terms, nterms, obs));
  if (!obs.xy.empty()) return fail(QG_ERR_UNSUPPORTED, "batched McClean runs take z / zz observables");
  if ((int)obs.diag.size() > p->max_terms) return fail(QG_ERR_INVALID, "too many observable terms");
  qg_ctx* c = p->ctx;
  HIPCHK(hipSetDevice(c->device));
  if (n <= kFusedMaxN) {
    // the whole-circuit kernel, one workgroup per circuit, one launch
    std::vector<double> res((size_t)B * per_out);
    auto fill = [&](int b, FusedDescH& d, Arena& ar) -> int {
      d.kind = 0;
      int32_t* dt = nullptr;
      int32_t* ht = ar.take<int32_t>(np, &dt);
      double* dg = nullptr;
      double* hg = ar.take<double>(2 * np, &dg);
      const int32_t* ax = axes + (size_t)b * np;
      const double* an = angles + (size_t)b * np;
      for (size_t k = 0; k < np; ++k) {
        ht[k] = ax[k] == 0 ? G_X : ax[k] == 1 ? G_Y : G_Z;
        hg[2 * k] = std::cos(0.5 * an[k]);
        hg[2 * k + 1] = std::sin(0.5 * an[k]);
      }
      d.gtype = dt;
      d.gcs = dg;
      d.psi_out = nullptr;
      return obs_tables(obs, d, ar);
    };
    const size_t bytes = np * 20 + (obs.diag.size() + 2) * sizeof(DTerm);
    CHK(run_fused(p, B, bytes, per_out, true, nullptr, fill, res.data()));
    for (int b = 0; b < B; ++b) {
      E_out[b] = res[(size_t)b * per_out];
      std::memcpy(grad_out + (size_t)b * np, res.data() + (size_t)b * per_out + 1, sizeof(double) * np);
    }
    p->last_state_ptr = nullptr;
    return QG_OK;
  }
  CHK(ensure_workspace(p));
  if (!obs.diag.empty())
    HIPCHK(hipMemcpyAsync(p->d_terms, obs.diag.data(), sizeof(DTerm) * obs.diag.size(), hipMemcpyHostToDevice,
                          c->stream));
  p->h_terms = obs.diag;
  const uint32_t ent_fwd = p->ent == QG_ENT_CNOT ? F_GATHER : F_CZ_PRE;
  const uint32_t ent_rev = p->ent == QG_ENT_CNOT ? F_SCATTER : F_CZ_POST;
  float total_ms = 0.f;
  for (int c0 = 0; c0 < B; c0 += p->chunk) {
    const int nb = std::min(p->chunk, B - c0);
    auto hist = [&](int j, int i, int g) { return wsp(p, p->off_hist + (((size_t)j * L + i) * P + g) * p->S); };
    auto lam = [&](int j, int k) { return wsp(p, p->off_lam + ((size_t)j * 2 + k) * p->S); };
    p->launches.clear();
    int nd = 0, row = 0;
    std::vector<std::vector<int>> out_rows((size_t)nb * per_out);
    std::vector<int> cnt(p->nrows, 0);
    // one launch per pass over the chunk's circuits, at a single plan's grid
    auto begin = [&](const Geo& G0, bool two, int role, int* grid) -> Geo {
      bool tf = false;
      const Geo G = with_role(G0, p->dtype, role, &tf);
      Launch l;
      l.desc = nd;
      l.batch = nb;
      l.R = G.R;
      l.K = G.K;
      l.post = tf;
      l.threads = G.threads;
      l.grid = grid_for_pass(p, G, two, tf, 0);
      l.phase = 0;
      l.two = two;
      l.kind = 0;
      p->launches.push_back(l);
      *grid = l.grid;
      return G;
    };
    auto base = [&](const Geo& G) -> PassDesc& {
      PassDesc& d = p->h_desc[nd++];
      desc_geo(d, n, G);
      d.grid = p->gmax;
      d.lad1 = p->lad1;
      d.lad2 = p->lad2;
      d.init_tab = p->d_init;
      d.dterms = p->d_terms;
      d.ndterms = (int)obs.diag.size();
      d.part = p->d_part;
      return d;
    };
    auto gates = [&](PassDesc& d, const Geo& G, int j, int i, double sign, int rb) {
      const int32_t* ax = axes + (size_t)(c0 + j) * np;
      const double* an = angles + (size_t)(c0 + j) * np;
      for (int lb = G.own_lo; lb < G.own_hi; ++lb) {
        const int b = (lb < G.c) ? lb : G.p + (lb - G.c);
        const int q = n - 1 - b;
        set_gate(d, lb, ax[i * n + q], an[i * n + q], sign);
        if (rb >= 0) {
          d.own[lb] = 1;
          out_rows[(size_t)j * per_out + 1 + i * n + q].push_back(rb + lb);
        }
      }
    };
    // forward sweep; lambda = D psi and E ride on the last pass
    for (int i = 0; i < L; ++i) {
      for (int g = 0; g < P; ++g) {
        const bool last = (i == L - 1) && (g == P - 1);
        int gr = 0;
        const Geo G = begin(p->geo[g], last, last ? ROLE_TWO_FP64 : ROLE_FWD, &gr);
        for (int j = 0; j < nb; ++j) {
          PassDesc& d = base(G);
          uint32_t f = F_ROT_A | F_STORE_A;
          if (g == 0) f |= ent_fwd | (i == 0 ? F_INIT_TAB : F_LOAD_A);
          else f |= F_LOAD_A;
          // history passes: the input is read once here
          if ((f & F_LOAD_A) && ((uint64_t)nb << n) <= kNtMaxLaunchAmps) f |= F_NT_A;
          d.a_in = (i == 0 && g == 0) ? nullptr : (g == 0 ? hist(j, i - 1, P - 1) : hist(j, i, g - 1));
          d.a_out = hist(j, i, g);
          gates(d, p->geo[g], j, i, 1.0, -1);
          if (last) {
            f |= F_OBS_DIAG | F_STORE_B;
            d.b_out = lam(j, 0);
            d.row_base = row;
            out_rows[(size_t)j * per_out].push_back(row + G.K);
            for (int k = 0; k <= kMaxK; ++k) cnt[row + k] = gr;
            row += kMaxK + 1;
          }
          d.flags = f;
        }
      }
    }
    // adjoint sweep (reduce-first history passes)
    int lc = 0;
    for (int i = L - 1; i >= 0; --i) {
      for (int g = P - 1; g >= 0; --g) {
        int gr = 0;
        const Geo G = begin(p->geo[g], true, ROLE_REV_HIST, &gr);
        const bool swap = g == 0 && p->ent == QG_ENT_CNOT;
        for (int j = 0; j < nb; ++j) {
          PassDesc& d = base(G);
          uint32_t f = F_LOAD_A | F_ROT_A | F_LOAD_B | F_ROT_B | F_STORE_B | F_REDUCE;
          if (g == 0) f |= ent_rev;
          if (((uint64_t)nb << n) <= kNtMaxLaunchAmps) f |= F_NT_A;  // the history state is read once here
          d.a_in = hist(j, i, g);
          d.b_in = lam(j, lc);
          d.b_out = lam(j, swap ? lc ^ 1 : lc);
          d.flags = f;
          d.row_base = row;
          gates(d, p->geo[g], j, i, -1.0, row);
          for (int k = 0; k <= kMaxK; ++k) cnt[row + k] = gr;
          row += kMaxK + 1;
        }
        if (swap) lc ^= 1;
      }
    }
    std::vector<double> scale(out_rows.size(), 1.0);
    CHK(run_launches(p, nd));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    CHK(collect_profile(p));
    CHK(upload_finalize(p, out_rows, cnt, scale));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    total_ms += ms;
    HIPCHK(hipMemcpyAsync(p->h_out, p->d_out, sizeof(double) * (size_t)nb * per_out, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int j = 0; j < nb; ++j) {
      const double* o = p->h_out + (size_t)j * per_out;
      E_out[c0 + j] = o[0];
      std::memcpy(grad_out + (size_t)(c0 + j) * np, o + 1, sizeof(double) * np);
    }
  }
  p->last_ms = total_ms;
  p->last_state_ptr = nullptr;
  return QG_OK;
}

// The whole-circuit McClean grad run of qg_mcclean_grad_run (n <= 13, no
// ini_state), with every table and result in the device's static scratch
// array instead of plan allocations.  The tables are built exactly as
// mcclean_run builds them for the whole-circuit kernel, so the results are
// bitwise the same.  No device or host allocation: the tables are staged in
// the context's pinned mirror of the scratch, made at qg_ctx_create.
extern "C" int qg_mcclean_grad_run_noalloc(qg_ctx* c, int n, int L, int entangler, const int32_t* axes,
                                           const double* angles, const qg_term* terms, int nterms, double* E_out,
                                           double* grad_out) {
  if (!c || !axes || !angles || !E_out || !grad_out) return fail(QG_ERR_INVALID, "NULL argument");
  if (n < 1 || n > kFusedMaxN)
    return fail(QG_ERR_INVALID, "qg_mcclean_grad_run_noalloc: qubit number must be in [1, 13] (one workgroup per "
                                "circuit); larger circuits need a plan and qg_mcclean_grad_run");
  if (L < 1) return fail(QG_ERR_INVALID, "layer number must be >= 1");
  if (entangler != QG_ENT_CNOT && entangler != QG_ENT_CZ) return fail(QG_ERR_INVALID, "unknown entangler");
  if (nterms < 0 || (nterms > 0 && !terms)) return fail(QG_ERR_INVALID, "observable: bad term array");
  for (int k = 0; k < L * n; ++k) {
    if (axes[k] < 0 || axes[k] > 2) return fail(QG_ERR_INVALID, "Invalid axis " + std::to_string(axes[k]));
    if (!std::isfinite(angles[k])) return fail(QG_ERR_INVALID, "non-finite angle");
  }
  // observable terms, validated as split_observable validates them
  int ndiag = 0, nxy = 0;
  for (int k = 0; k < nterms; ++k) {
    const qg_term& t = terms[k];
    if (t.i < 0 || t.i >= n) return fail(QG_ERR_INVALID, "observable: qubit index out of range");
    if (!std::isfinite(t.w)) return fail(QG_ERR_INVALID, "observable: non-finite weight");
    switch (t.kind) {
      case QG_TERM_Z: ++ndiag; break;
      case QG_TERM_ZZ:
        if (t.j <= t.i || t.j >= n)
          return fail(QG_ERR_INVALID, "observable: zz terms need i < j < n (upper triangle, observable.py:56-64)");
        ++ndiag;
        break;
      case QG_TERM_X:
      case QG_TERM_Y: ++nxy; break;
      default: return fail(QG_ERR_INVALID, "observable: unknown term kind");
    }
  }
  const int max_terms = std::max(64, n * n);  // as a McClean plan allows
  if (ndiag > max_terms || nxy > max_terms) return fail(QG_ERR_INVALID, "too many observable terms");
  // scratch layout (256-byte aligned pieces)
  const size_t np = (size_t)L * n;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = (off + bytes + 255) & ~(size_t)255;
    return at;
  };
  const size_t o_desc = take(sizeof(FusedDescH));
  const size_t o_type = take(sizeof(int32_t) * np);
  const size_t o_gcs = take(sizeof(double) * 2 * np);
  const size_t o_diag = take(sizeof(DTerm) * std::max(1, ndiag));
  const size_t o_xy = take(sizeof(DTerm) * std::max(1, nxy));
  const size_t o_init = take(sizeof(double) * (n + 1));
  const size_t o_out = take(sizeof(double) * (1 + np));
  if (off > kNoAllocScratchBytes)
    return fail(QG_ERR_INVALID, "qg_mcclean_grad_run_noalloc: the circuit's tables need " + std::to_string(off) +
                                    " bytes, more than the " + std::to_string(kNoAllocScratchBytes) +
                                    "-byte scratch (fewer layers, or a plan and qg_mcclean_grad_run)");
  HIPCHK(hipSetDevice(c->device));
  unsigned char* dev = (unsigned char*)noalloc_scratch(c->device);
  if (!dev) return fail(QG_ERR_HIP, "qg_mcclean_grad_run_noalloc: no scratch address for this device");
  unsigned char* h = c->noalloc_host;
  // gates as (type, cos, sin), as mcclean_run's whole-circuit tables
  int32_t* ht = (int32_t*)(h + o_type);
  double* hg = (double*)(h + o_gcs);
  for (size_t k = 0; k < np; ++k) {
    ht[k] = axes[k] == 0 ? G_X : axes[k] == 1 ? G_Y : G_Z;
    hg[2 * k] = std::cos(0.5 * angles[k]);
    hg[2 * k + 1] = std::sin(0.5 * angles[k]);
  }
  // observable, in split_observable's order
  DTerm* hd = (DTerm*)(h + o_diag);
  DTerm* hx = (DTerm*)(h + o_xy);
  int id = 0, ix = 0;
  for (int k = 0; k < nterms; ++k) {
    const qg_term& t = terms[k];
    const int bi = n - 1 - t.i;
    if (t.kind == QG_TERM_Z) hd[id++] = {bi, -1, t.w};
    else if (t.kind == QG_TERM_ZZ) hd[id++] = {bi, n - 1 - t.j, t.w};
    else if (t.kind == QG_TERM_X) hx[ix++] = {bi, 0, t.w};
    else hx[ix++] = {bi, 1, t.w};
  }
  // the ladder's gather map and the RY(pi/4) table, as a McClean plan's
  uint32_t lad1 = 0, lad2 = 0;
  if (!ladder_shift_form(n, 1, &lad1, &lad2)) return fail(QG_ERR_INVALID, "ladder map not in shift form");
  ry_init_table(n, (double*)(h + o_init));
  FusedDescH* hdsc = (FusedDescH*)(h + o_desc);
  std::memset(hdsc, 0, sizeof(FusedDescH));
  hdsc->n = n;
  hdsc->L = L;
  hdsc->kind = 0;
  hdsc->ent = entangler;
  hdsc->gtype = (const int32_t*)(dev + o_type);
  hdsc->gcs = (const double*)(dev + o_gcs);
  hdsc->dterms = (const DTerm*)(dev + o_diag);
  hdsc->ndterms = ndiag;
  hdsc->xy = (const DTerm*)(dev + o_xy);
  hdsc->nxy = nxy;
  hdsc->lad1 = lad1;
  hdsc->lad2 = lad2;
  hdsc->init_tab = (const double*)(dev + o_init);
  hdsc->out = (double*)(dev + o_out);
  hdsc->grad = 1;
  hdsc->psi_c64 = 0;
  hdsc->psi_out = nullptr;
  // tables in, one launch, results out (the kernel writes every result, so
  // nothing in the scratch needs clearing)
  HIPCHK(hipMemcpyAsync(dev, h, o_out, hipMemcpyHostToDevice, c->stream));
  HIPCHK(launch_fused(dev + o_desc, 1, n, c->stream));
  HIPCHK(hipMemcpyAsync(h + o_out, dev + o_out, sizeof(double) * (1 + np), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const double* res = (const double*)(h + o_out);
  *E_out = res[0];
  std::memcpy(grad_out, res + 1, sizeof(double) * np);
  return QG_OK;
}

// ---------------------------------------------------------------------------
// Parameter-shifted McClean circuits (McClean.sample_grad, mc_clean.py:117-198)
// ---------------------------------------------------------------------------

// Appends the forward passes of layers [first, L) to the plan's launch list,
// from the state at src (nullptr: the RY(pi/4) product state, generated by the
// first pass).  The states go to the two buffers buf: the low pass gathers
// out of place, high passes work in place, so src is only read.  Returns the
// buffer that holds the final state.
static void* mcclean_forward_suffix(qg_plan* p, const int32_t* axes, const double* angles, int first,
                                    const void* src, void* const buf[2], int& nd) {
  const int n = p->n, L = p->L, P = (int)p->geo.size();
  const uint32_t ent = p->ent == QG_ENT_CNOT ? F_GATHER : F_CZ_PRE;
  const void* cur = src;
  for (int i = first; i < L; ++i) {
    for (int g = 0; g < P; ++g) {
      const Geo& G0 = p->geo[g];
      bool tf = false;
      const Geo G = with_role(G0, p->dtype, ROLE_FWD, &tf);
      Launch l;
      l.desc = nd;
      l.R = G.R;
      l.K = G.K;
      l.post = tf;
      l.threads = G.threads;
      l.grid = grid_for_pass(p, G, false, tf, 0);
      l.phase = 0;
      l.two = false;
      l.kind = 0;
      p->launches.push_back(l);
      PassDesc& d = p->h_desc[nd++];
      desc_geo(d, n, G);
      d.grid = p->gmax;
      d.lad1 = p->lad1;
      d.lad2 = p->lad2;
      d.init_tab = p->d_init;
      d.dterms = p->d_terms;
      d.ndterms = 0;
      d.part = p->d_part;
      uint32_t f = F_ROT_A | F_STORE_A;
      f |= (g == 0) ? (ent | (cur ? F_LOAD_A : F_INIT_TAB)) : F_LOAD_A;
      void* out = (g == 0) ? (cur == buf[0] ? buf[1] : buf[0]) : (void*)cur;
      d.flags = f;
      d.a_in = cur;
      d.a_out = out;
      for (int lb = G0.own_lo; lb < G0.own_hi; ++lb) {
        const int b = (lb < G0.c) ? lb : G0.p + (lb - G0.c);
        const int q = n - 1 - b;
        set_gate(d, lb, axes[i * n + q], angles[i * n + q], 1.0);
      }
      cur = out;
    }
  }
  return (void*)cur;
}

// Runs the unshifted circuit through the pass kernels, keeping its layer-end
// states (history) or checkpoints, then each shifted circuit from the nearest
// stored state before its shifted layer (the reference restarts from
// state_history[i], mc_clean.py:140-151), and calls on_shift(s, final state)
// to enqueue that shift's measurement on the plan's stream.  Afterwards the
// plan's state is the unshifted circuit's final state.
template <class OnShift>
static int mcclean_shift_run(qg_plan* p, const int32_t* axes, const double* angles, const double* ini_state,
                             int nshift, const int32_t* shift_param, const double* shift_delta, OnShift on_shift) {
  const int n = p->n, L = p->L;
  for (int s = 0; s < nshift; ++s) {
    if (shift_param[s] < 0 || shift_param[s] >= L * n)
      return fail(QG_ERR_INVALID, "shift parameter index out of range");
    if (!std::isfinite(shift_delta[s])) return fail(QG_ERR_INVALID, "non-finite shift");
  }
  const bool fused = p->fused;
  p->fused = false;
  double E = 0.0;
  const qg_term none{QG_TERM_Z, 0, -1, 0, 0.0};  // the forward run's own observable: unused
  const int r = mcclean_run(p, axes, angles, &none, 1, ini_state, &E, nullptr, false);
  p->fused = fused;
  CHK(r);
  qg_ctx* c = p->ctx;
  const int P = (int)p->geo.size();
  std::vector<const void*> layer_end(L, nullptr);
  if (p->history) {
    for (int i = 0; i < L; ++i) layer_end[i] = wsp(p, p->off_hist + ((size_t)i * P + P - 1) * p->S);
  } else {
    int m = 0;
    for (int i = 0; i < L; ++i)
      if ((i + 1) % p->ckpt == 0 || i == L - 1) layer_end[i] = wsp(p, p->off_ckpt + (size_t)(m++) * p->S);
  }
  float total_ms = p->last_ms;  // the unshifted run's
  const void* start0 = ini_state ? wsp(p, p->off_prep) : nullptr;
  // the adjoint's two lambda buffers are free in a forward-only run
  void* buf[2] = {wsp(p, p->off_lam), wsp(p, p->off_lam + p->S)};
  std::vector<double> ang(angles, angles + (size_t)L * n);
  // Two halves of the descriptor table (a suffix has at most P * L passes):
  // the host builds shift s + 1 while the device runs shift s, and waits for
  // shift s - 1, the last user of a half, before rewriting it.
  const int half = P * L;
  struct Ev {
    hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
    ~Ev() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } ev;
  for (int k = 0; k < 4; ++k) HIPCHK(hipEventCreate(&ev.e[k]));  // done[0], done[1], start, end
  const bool graph = p->use_graph;
  p->use_graph = false;  // suffixes differ in length: no graph replay here
  int rc = QG_OK;
  hipError_t he = hipSuccess;
  for (int s = 0; s < nshift && rc == QG_OK && he == hipSuccess; ++s) {
    const int k = shift_param[s], i = k / n, h = s & 1;
    if (s >= 2) he = hipEventSynchronize(ev.e[h]);
    if (he != hipSuccess) break;
    int j = i - 1;
    while (j >= 0 && !layer_end[j]) --j;
    ang[k] = angles[k] + shift_delta[s];
    p->launches.clear();
    int nd = h * half;
    void* fin = mcclean_forward_suffix(p, axes, ang.data(), j + 1, j >= 0 ? layer_end[j] : start0, buf, nd);
    ang[k] = angles[k];
    if (s == 0) he = hipEventRecord(ev.e[2], c->stream);
    if (he != hipSuccess) break;
    rc = run_launches(p, nd, h * half);
    if (rc != QG_OK) break;
    he = on_shift(s, (const void*)fin);
    if (he == hipSuccess) he = hipEventRecord(ev.e[h], c->stream);
  }
  p->use_graph = graph;
  if (rc == QG_OK && he == hipSuccess && nshift > 0) he = hipEventRecord(ev.e[3], c->stream);
  const hipError_t hs = hipStreamSynchronize(c->stream);
  if (rc != QG_OK) return rc;
  if (he == hipSuccess) he = hs;
  if (he != hipSuccess) return fail(QG_ERR_HIP, std::string("shifted circuits: ") + hipGetErrorString(he));
  if (nshift > 0) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, ev.e[2], ev.e[3]));
    total_ms += ms;
  }
  p->last_ms = total_ms;  // device time of the unshifted and every shifted circuit
  return QG_OK;
}

// Device scratch freed on every return path.
struct DevBuf {
  char* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

extern "C" int qg_mcclean_shift_term_prob(qg_plan* p, const int32_t* axes, const double* angles,
                                          const double* ini_state, int nshift, const int32_t* shift_param,
                                          const double* shift_delta, const qg_term* terms, int nterms,
                                          double* prob) {
  if (!p || p->kind != PK_MCCLEAN) return fail(QG_ERR_INVALID, "not a McClean plan");
  if (nshift < 0 || nterms < 0) return fail(QG_ERR_INVALID, "negative count");
  if (nshift > 0 && (!shift_param || !shift_delta || (nterms > 0 && !prob)))
    return fail(QG_ERR_INVALID, "NULL argument");
  std::vector<PTerm> h;
  CHK(pterm_table(p->n, terms, nterms, h));
  const int nt = nterms + 1;
  const size_t tb = ((sizeof(PTerm) * nt + 255) / 256) * 256;
  const size_t pb = sizeof(double) * kTermBlocks * kTermPartRows;
  HIPCHK(hipSetDevice(p->ctx->device));
  DevBuf db;
  HIPCHK(hipMalloc(&db.p, tb + pb + sizeof(double) * (size_t)nt * std::max(nshift, 1)));
  PTerm* d_t = (PTerm*)db.p;
  double* d_part = (double*)(db.p + tb);
  double* d_res = (double*)(db.p + tb + pb);
  HIPCHK(hipMemcpy(d_t, h.data(), sizeof(PTerm) * nt, hipMemcpyHostToDevice));
  CHK(mcclean_shift_run(p, axes, angles, ini_state, nterms ? nshift : 0, shift_param, shift_delta,
                        [&](int s, const void* fin) {
                          return simple_term_expec(p->dtype, fin, p->n, d_t, nt, d_part, d_res + (size_t)s * nt,
                                                   p->ctx->stream, h.data());
                        }));
  if (nshift == 0 || nterms == 0) return QG_OK;
  std::vector<double> res((size_t)nt * nshift);
  HIPCHK(hipMemcpy(res.data(), d_res, sizeof(double) * res.size(), hipMemcpyDeviceToHost));
  for (int s = 0; s < nshift; ++s)
    for (int t = 0; t < nterms; ++t)
      prob[(size_t)s * nterms + t] = 0.5 * (res[(size_t)s * nt + nterms] + res[(size_t)s * nt + t]);
  return QG_OK;
}

// Device copies of a sampling order (validated: every entry in [0, 2^n)) and
// of the uniforms, and room for the drawn positions and the block masses.
struct SampleBufs {
  DevBuf db;
  int32_t* order = nullptr;
  double* u = nullptr;
  int32_t* idx = nullptr;
  double* mass = nullptr;
};

static int sample_setup(int n, const int32_t* order, int64_t nu, const double* u, SampleBufs& b) {
  if (n > 31) return fail(QG_ERR_UNSUPPORTED, "basis sampling needs n <= 31");
  const uint64_t N = 1ull << n;
  if (order)
    for (uint64_t k = 0; k < N; ++k)
      if (order[k] < 0 || (uint64_t)order[k] >= N) return fail(QG_ERR_INVALID, "sampling order entry out of range");
  for (int64_t k = 0; k < nu; ++k)
    if (!(u[k] >= 0.0 && u[k] <= 1.0)) return fail(QG_ERR_INVALID, "uniform draws must lie in [0, 1]");
  const size_t nb = (N + kSampleBlock - 1) / kSampleBlock;
  const size_t ob = order ? ((N * 4 + 255) / 256) * 256 : 0;
  const size_t ub = (((size_t)nu * 8 + 255) / 256) * 256;
  const size_t ib = (((size_t)nu * 4 + 255) / 256) * 256;
  HIPCHK(hipMalloc(&b.db.p, ob + ub + ib + nb * 8));
  b.order = order ? (int32_t*)b.db.p : nullptr;
  b.u = (double*)(b.db.p + ob);
  b.idx = (int32_t*)(b.db.p + ob + ub);
  b.mass = (double*)(b.db.p + ob + ub + ib);
  if (order) HIPCHK(hipMemcpy(b.order, order, N * 4, hipMemcpyHostToDevice));
  if (nu) HIPCHK(hipMemcpy(b.u, u, (size_t)nu * 8, hipMemcpyHostToDevice));
  return QG_OK;
}

extern "C" int qg_mcclean_shift_sample(qg_plan* p, const int32_t* axes, const double* angles, const double* ini_state,
                                       int nshift, const int32_t* shift_param, const double* shift_delta,
                                       const int32_t* order, int shots, const double* u, int32_t* idx) {
  if (!p || p->kind != PK_MCCLEAN) return fail(QG_ERR_INVALID, "not a McClean plan");
  if (nshift < 0 || shots < 1) return fail(QG_ERR_INVALID, "need nshift >= 0 and shots >= 1");
  if (nshift > 0 && (!shift_param || !shift_delta || !u || !idx)) return fail(QG_ERR_INVALID, "NULL argument");
  HIPCHK(hipSetDevice(p->ctx->device));
  SampleBufs b;
  const int64_t nu = (int64_t)nshift * shots;
  CHK(sample_setup(p->n, order, nu, u, b));
  CHK(mcclean_shift_run(p, axes, angles, ini_state, nshift, shift_param, shift_delta,
                        [&](int s, const void* fin) {
                          return simple_sample_basis(p->dtype, fin, p->n, b.order, b.u + (size_t)s * shots, shots,
                                                     b.idx + (size_t)s * shots, b.mass, p->ctx->stream);
                        }));
  if (nu) HIPCHK(hipMemcpy(idx, b.idx, (size_t)nu * 4, hipMemcpyDeviceToHost));
  return QG_OK;
}

// ---------------------------------------------------------------------------
// QAOA (qradient/circuit_logic/qaoa.py:40-70), batched over circuits
// ---------------------------------------------------------------------------
// Geometry, chunk and tables of a QAOA plan (tile shape from kmax / cmin,
// chunk from chunk_req and device memory).
static int qaoa_layout(qg_plan* p) {
  const int n = p->n, L = p->L;
  if (p->kmax_auto) p->kmax = default_kmax(n, p->dtype, p->cmin);
  p->geo = layer_geometry(n, p->kmax, p->cmin);
  const int P = (int)p->geo.size();
  p->S = (size_t)(1ull << n) * (p->dtype == QG_C64 ? 8 : 16);
  const size_t per_circuit = (size_t)(P * L + 2) * p->S;
  size_t budget = 0;
  CHK(plan_budget(p, &budget));
  if (per_circuit > budget)
    return fail(QG_ERR_UNSUPPORTED, "one QAOA circuit's state history does not fit in device memory");
  const int fit = (int)std::min<size_t>((size_t)p->batch, budget / per_circuit);
  p->chunk = p->chunk_req > 0 ? std::min(p->chunk_req, fit) : fit;
  p->off_hist = 0;
  p->off_lam = (size_t)p->chunk * P * L * p->S;
  p->ws_bytes = p->off_lam + (size_t)p->chunk * 2 * p->S;
  p->gmax = grid_max(p);
  const int ndesc = p->chunk * 2 * P * L;
  p->nrows = p->chunk * (P * L + 1) * (kMaxK + 1);
  CHK(ensure_small(p, ndesc, p->nrows, p->chunk * (1 + 2 * L)));
  // each circuit of a chunk has its own classical Hamiltonian
  const int mt = n + n * (n - 1) / 2;
  const size_t need = (size_t)mt * p->chunk;
  if (need > p->qaoa_terms_cap) {
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
    (void)hipFree(p->d_terms);
    p->d_terms = nullptr;
    p->qaoa_terms_cap = 0;
    HIPCHK(hipMalloc(&p->d_terms, sizeof(DTerm) * need));
    p->qaoa_terms_cap = need;
  }
  p->max_terms = mt;
  double* hi = (double*)p->h_tabs;
  for (int k = 0; k <= n; ++k) hi[k] = std::pow(2.0, -0.5 * n);
  HIPCHK(hipMemcpyAsync(p->d_init, hi, sizeof(double) * (n + 1), hipMemcpyHostToDevice, p->ctx->stream));
  HIPCHK(hipStreamSynchronize(p->ctx->stream));
  return QG_OK;
}

extern "C" int qg_qaoa_plan(qg_ctx* ctx, int n, int p_layers, int dtype, int batch, qg_plan** out) {
  if (!ctx || !out) return fail(QG_ERR_INVALID, "NULL argument");
  if (n < 1 || n > kMaxQubits) return fail(QG_ERR_INVALID, "qubit number must be in [1, 32]");
  if (p_layers < 1) return fail(QG_ERR_INVALID, "layer number must be >= 1");
  if (dtype != QG_C128 && dtype != QG_C64) return fail(QG_ERR_INVALID, "dtype must be QG_C128 or QG_C64");
  if (batch < 1) return fail(QG_ERR_INVALID, "batch must be >= 1");
  HIPCHK(hipSetDevice(ctx->device));
  qg_plan* p = new qg_plan();
  p->ctx = ctx;
  p->kind = PK_QAOA;
  p->n = n;
  p->L = p_layers;
  p->dtype = dtype;
  p->batch = batch;
  p->wg_per_cu = kBatchWgPerCu;
  const int r = qaoa_layout(p);
  if (r != QG_OK) {
    qg_plan_destroy(p);
    return r;
  }
  *out = p;
  return QG_OK;
}

extern "C" int qg_qaoa_tune(qg_plan* p, int kmax, int cmin, int chunk, int wg_per_cu) {
  if (!p || p->kind != PK_QAOA) return fail(QG_ERR_INVALID, "not a QAOA plan");
  if (kmax != 0 && (kmax < 4 || kmax > 13)) return fail(QG_ERR_INVALID, "kmax must be 0 (default) or in [4, 13]");
  if (cmin < 0 || (kmax != 0 && cmin >= kmax)) return fail(QG_ERR_INVALID, "cmin must be in [0, kmax)");
  if (chunk < 0) return fail(QG_ERR_INVALID, "chunk must be >= 0 (0: as many as fit)");
  if (wg_per_cu < 1 || wg_per_cu > 64) return fail(QG_ERR_INVALID, "wg_per_cu must be in [1, 64]");
  HIPCHK(hipSetDevice(p->ctx->device));
  HIPCHK(hipStreamSynchronize(p->ctx->stream));
  drop_graph(p);
  p->kmax_auto = kmax == 0;
  if (kmax) p->kmax = kmax;
  p->cmin = cmin;
  p->chunk_req = chunk;
  p->wg_per_cu = wg_per_cu;
  if (p->ws && p->ws_owned) {
    (void)hipFree(p->ws);
    p->ws = nullptr;
    p->ws_cap = 0;
    p->ws_owned = false;
  }
  const size_t old_ws = p->ws_cap;
  CHK(qaoa_layout(p));
  if (p->ws && p->ws_bytes > old_ws) {
    p->ws = nullptr;  // a supplied workspace that is too small is detached
    p->ws_cap = 0;
  }
  p->last_state_ptr = nullptr;
  return QG_OK;
}

// ini: B initial states (complex128, 2^n each) or null for |+>^n (qaoa.py:17)
static int qaoa_run(qg_plan* p, int B, const qg_term* terms, const int32_t* term_off, const double* betas,
                    const double* gammas, const double* ini, double* E_out, double* grad_out, bool grad) {
  if (!p || p->kind != PK_QAOA) return fail(QG_ERR_INVALID, "not a QAOA plan");
  if (B < 1 || B > p->batch) return fail(QG_ERR_INVALID, "B must be in [1, batch]");
  if (!term_off || !betas || !gammas || !E_out || (grad && !grad_out)) return fail(QG_ERR_INVALID, "NULL argument");
  const int n = p->n, L = p->L, P = (int)p->geo.size();
  if (term_off[0] != 0) return fail(QG_ERR_INVALID, "term_off[0] must be 0");
  for (int b = 0; b < B; ++b)
    if (term_off[b + 1] < term_off[b] || term_off[b + 1] - term_off[b] > p->max_terms)
      return fail(QG_ERR_INVALID, "bad term offsets");
  for (int k = 0; k < B * L; ++k)
    if (!std::isfinite(betas[k]) || !std::isfinite(gammas[k])) return fail(QG_ERR_INVALID, "non-finite parameter");
  // split and classify the classical Hamiltonians
  std::vector<ObsSplit> obs(B);
  bool integral = true;
  double W = 0.0;
  for (int b = 0; b < B; ++b) {
    CHK(split_observable(n, terms + term_off[b], term_off[b + 1] - term_off[b], obs[b]));
    if (!obs[b].xy.empty())
      return fail(QG_ERR_INVALID, "the QAOA classical Hamiltonian takes z / zz terms only (state.py:268-271)");
    double wsum = 0.0;
    for (const DTerm& t : obs[b].diag) {
      wsum += std::fabs(t.w);
      if (t.w != std::rint(t.w)) integral = false;
    }
    W = std::max(W, wsum);
  }
  // integer weights: exp(-i gamma C) from a table of 2W + 1 entries staged in LDS
  const int phase = (integral && 2.0 * W + 1.0 <= (double)kLutMax) ? 1 : 2;
  const int Wi = (int)W;
  qg_ctx* c = p->ctx;
  HIPCHK(hipSetDevice(c->device));
  if (p->fused && n <= kFusedMaxN && !ini) {
    const int per = 1 + 2 * L;
    std::vector<double> res((size_t)B * per);
    auto fill = [&](int b, FusedDescH& d, Arena& ar) -> int {
      d.kind = 1;
      double* dp = nullptr;
      double* hp = ar.take<double>((size_t)2 * L, &dp);
      for (int i = 0; i < L; ++i) {
        hp[2 * i] = betas[(size_t)b * L + i];
        hp[2 * i + 1] = gammas[(size_t)b * L + i];
      }
      d.params = dp;
      return obs_tables(obs[b], d, ar);
    };
    size_t mx = 0;
    for (int b = 0; b < B; ++b) mx = std::max(mx, obs[b].diag.size());
    const size_t bytes = (size_t)L * 16 + (mx + 2) * sizeof(DTerm);
    // a single circuit keeps its final state (qg_plan_get_state, qg_plan_term_prob)
    if (B == 1 && p->cap_fstate < ((size_t)16 << n)) {
      HIPCHK(hipStreamSynchronize(c->stream));
      (void)hipFree(p->d_fstate);
      p->d_fstate = nullptr;
      p->cap_fstate = 0;
      HIPCHK(hipMalloc(&p->d_fstate, ((size_t)16) << n));
      p->cap_fstate = (size_t)16 << n;
    }
    CHK(run_fused(p, B, bytes, per, grad, B == 1 ? p->d_fstate : nullptr, fill, res.data()));
    for (int b = 0; b < B; ++b) {
      E_out[b] = res[(size_t)b * per];
      if (grad)
        for (int k = 0; k < 2 * L; ++k) grad_out[(size_t)b * 2 * L + k] = res[(size_t)b * per + 1 + k];
    }
    p->last_state_ptr = B == 1 ? p->d_fstate : nullptr;
    p->state_sign = false;
    return QG_OK;
  }
  CHK(ensure_workspace(p));
  const int chunk = p->chunk;
  const int lut_len = 2 * Wi + 1;
  if (ini && p->ini_cap < (size_t)chunk * p->S) {
    // one initial state per circuit of a chunk, in the storage type
    HIPCHK(hipStreamSynchronize(c->stream));
    (void)hipFree(p->d_ini_states);
    p->d_ini_states = nullptr;
    p->ini_cap = 0;
    HIPCHK(hipMalloc(&p->d_ini_states, (size_t)chunk * p->S));
    p->ini_cap = (size_t)chunk * p->S;
  }
  if (phase == 1) {
    const int need = chunk * 2 * L * lut_len * 2;
    if (need > p->lut_len) {
      (void)hipFree(p->d_lut);
      p->d_lut = nullptr;
      HIPCHK(hipMalloc(&p->d_lut, sizeof(double) * need));
      p->lut_len = need;
    }
  }
  const int per_out = 1 + 2 * L;
  float total_ms = 0.f;
  for (int c0 = 0; c0 < B; c0 += chunk) {
    const int nb = std::min(chunk, B - c0);
    // terms and phase tables of this chunk, staged in pinned memory and
    // copied on the stream ahead of the passes
    const size_t term_bytes = sizeof(DTerm) * (size_t)p->max_terms * nb;
    const size_t lut_doubles = phase == 1 ? (size_t)nb * 2 * L * lut_len * 2 : 0;
    CHK(ensure_stage(p, term_bytes + sizeof(double) * lut_doubles));
    DTerm* tbuf = (DTerm*)p->h_stage;
    std::vector<int> toff(nb), tcnt(nb);
    p->h_terms.assign((size_t)p->max_terms * nb, DTerm{});
    for (int j = 0; j < nb; ++j) {
      toff[j] = j * p->max_terms;
      tcnt[j] = (int)obs[c0 + j].diag.size();
      std::copy(obs[c0 + j].diag.begin(), obs[c0 + j].diag.end(), p->h_terms.begin() + toff[j]);
    }
    std::memcpy(tbuf, p->h_terms.data(), term_bytes);
    HIPCHK(hipMemcpyAsync(p->d_terms, tbuf, term_bytes, hipMemcpyHostToDevice, c->stream));
    if (ini) {
      // upload (complex128) and convert to the storage type
      const size_t N = 1ull << n;
      double* tmp = nullptr;
      HIPCHK(hipMalloc(&tmp, N * 16));
      hipError_t e = hipSuccess;
      for (int j = 0; j < nb && e == hipSuccess; ++j) {
        e = hipMemcpyAsync(tmp, ini + (size_t)(c0 + j) * 2 * N, N * 16, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
          e = simple_convert_in(p->dtype, tmp, (char*)p->d_ini_states + (size_t)j * p->S, n, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      }
      (void)hipFree(tmp);
      if (e != hipSuccess) return fail(QG_ERR_HIP, std::string("ini_state upload: ") + hipGetErrorString(e));
    }
    // Phase tables, layer by layer: t[k] = exp(-i g (k - W)) for g = +gamma
    // (dir 0) and -gamma (dir 1), stored [layer][circuit][dir].  A layer's
    // tables are filled and sent just before the first pass that reads them,
    // so in a streamed run the host's sincos work for later layers overlaps
    // the device's passes (about 2.5 ms for 512 circuits, 8 layers, W = 30).
    // cos is even and sin odd (exactly, in the math library), so one sincos
    // per |d| fills four entries.
    double* lut_host = phase == 1 ? (double*)(p->h_stage + term_bytes) : nullptr;
    const size_t lut_layer = (size_t)nb * 2 * lut_len * 2;  // doubles per layer
    int lut_filled = 0;
    auto lut_fill = [&](int upto) -> int {
      for (; phase == 1 && lut_filled < std::min(upto, L); ++lut_filled) {
        const int i = lut_filled;
        double* layer = lut_host + (size_t)i * lut_layer;
        for (int j = 0; j < nb; ++j) {
          const double g = gammas[(c0 + j) * L + i];
          double* t0 = layer + ((size_t)j * 2 + 0) * lut_len * 2;
          double* t1 = layer + ((size_t)j * 2 + 1) * lut_len * 2;
          for (int m = 0; m <= Wi; ++m) {
            const double cm = std::cos(g * (double)m), sm = std::sin(g * (double)m);
            // dir 0 at d = +m and d = -m; dir 1 is its mirror image
            t0[2 * (Wi + m)] = cm;
            t0[2 * (Wi + m) + 1] = -sm;
            t0[2 * (Wi - m)] = cm;
            t0[2 * (Wi - m) + 1] = sm;
            t1[2 * (Wi + m)] = cm;
            t1[2 * (Wi + m) + 1] = sm;
            t1[2 * (Wi - m)] = cm;
            t1[2 * (Wi - m) + 1] = -sm;
          }
        }
        HIPCHK(hipMemcpyAsync(p->d_lut + (size_t)i * lut_layer, layer, sizeof(double) * lut_layer,
                              hipMemcpyHostToDevice, c->stream));
      }
      return QG_OK;
    };
    auto lut_ptr = [&](int j, int i, int dir) -> const double* {
      return phase == 1 ? p->d_lut + (size_t)i * lut_layer + ((size_t)j * 2 + dir) * lut_len * 2 : nullptr;
    };
    auto hist = [&](int j, int i, int g) {
      return wsp(p, p->off_hist + (((size_t)j * L + i) * P + g) * p->S);
    };
    auto lam = [&](int j) { return wsp(p, p->off_lam + (size_t)j * 2 * p->S); };
    // every qubit's mixer gate of circuit j, layer i: forward and inverse
    std::vector<GateAB> gab((size_t)nb * L * 2);
    for (int j = 0; j < nb; ++j)
      for (int i = 0; i < L; ++i) {
        gab[((size_t)j * L + i) * 2] = gate_ab(betas[(c0 + j) * L + i], 1.0);
        gab[((size_t)j * L + i) * 2 + 1] = gate_ab(betas[(c0 + j) * L + i], -1.0);
      }
    // descriptors: one batched launch per pass, nb consecutive descriptors
    p->launches.clear();
    int nd = 0, row = 0;
    std::vector<std::vector<int>> out_rows((size_t)nb * per_out);
    std::vector<int> cnt(p->nrows, 0);
    Geo cur_geo;
    auto begin_pass = [&](const Geo& G0, bool two, int role) {
      bool tf = false;
      const Geo G = with_role(G0, p->dtype, role, &tf);
      cur_geo = G;
      Launch l;
      l.desc = nd;
      l.batch = nb;
      l.R = G.R;
      l.K = G.K;
      l.post = tf;
      l.threads = G.threads;
      l.grid = grid_for_pass(p, G, two, tf, phase);
      // a batched launch fills the device across circuits: fewer workgroups
      // per circuit, each running more tiles (amortises its prologue); up to
      // wg_per_cu per CU in all, beyond the resident count, so that the
      // circuits' tiles spread evenly over the CUs
      while (l.grid > 1 && (uint64_t)l.grid * nb > (uint64_t)p->wg_per_cu * p->ctx->num_cu) l.grid >>= 1;
      l.phase = phase;
      l.two = two;
      l.kind = 0;
      p->launches.push_back(l);
      return l.grid;
    };
    auto base_desc = [&](const Geo& G, int j) -> PassDesc& {
      PassDesc& d = p->h_desc[nd++];
      desc_geo(d, n, G);
      d.grid = p->gmax;
      d.init_tab = p->d_init;
      d.dterms = p->d_terms + toff[j];
      d.ndterms = tcnt[j];
      d.lut_off = Wi;
      d.part = p->d_part;
      return d;
    };
    const uint32_t phase_flag = phase == 1 ? F_PHASE_LUT : 0u;
    // With two or more passes per layer the cost phase of layer i + 1 closes
    // the last pass of layer i (the first layer's opens its first pass), so
    // the adjoint of that pass takes d/dgamma and undoes the phase in the
    // natural layout before its reduce-first reductions: psi never has to be
    // rotated back.  With one pass per layer the phase opens each pass and
    // the adjoint takes the general form.
    const bool phase_last = P >= 2;
    // passes go to the device as they are built (graph replay and profiled
    // runs launch the whole list at the end); each descriptor splits at most
    // its circuit's terms
    const bool streamed = !p->use_graph && !p->profile;
    if (streamed) {
      size_t per_pass = 0;
      for (int j = 0; j < nb; ++j) per_pass += (size_t)tcnt[j];
      CHK(stream_begin(p, per_pass * (size_t)P * L * (grad ? 2 : 1)));
    }
    // forward
    for (int i = 0; i < L; ++i) {
      // this layer's passes read the phase tables of layers i and i + 1
      CHK(lut_fill(i + 2));
      for (int g = 0; g < P; ++g) {
        const Geo& G = p->geo[g];
        const bool last = (i == L - 1) && (g == P - 1);
        const int gr = begin_pass(G, last, last ? ROLE_TWO_FP64 : ROLE_FWD);
        for (int j = 0; j < nb; ++j) {
          PassDesc& d = base_desc(cur_geo, j);
          uint32_t f = F_ROT_A | F_STORE_A | phase_flag;
          if (g == 0 && (i == 0 || !phase_last)) {
            f |= (i == 0 && !ini ? F_INIT_TAB : F_LOAD_A) | F_PHASE_PRE;
            d.phase_gamma = gammas[(c0 + j) * L + i];
            d.lut = lut_ptr(j, i, 0);
          } else {
            f |= F_LOAD_A;
          }
          if (phase_last && g == P - 1 && i + 1 < L) {
            f |= F_PHASE_POST;
            d.phase_gamma = gammas[(c0 + j) * L + i + 1];
            d.lut = lut_ptr(j, i + 1, 0);
          }
          d.a_in = (i == 0 && g == 0) ? (ini ? (const void*)((char*)p->d_ini_states + (size_t)j * p->S) : nullptr)
                                      : (g == 0 ? hist(j, i - 1, P - 1) : hist(j, i, g - 1));
          d.a_out = hist(j, i, g);
          const GateAB ab = gab[((size_t)j * L + i) * 2];
          for (int lb = G.own_lo; lb < G.own_hi; ++lb) {
            d.gtype[lb] = G_X;
            d.ga[lb] = ab.a;
            d.gb[lb] = ab.b;
          }
          if (last) {
            f |= F_OBS_DIAG | F_STORE_B;
            d.b_out = lam(j);
            d.row_base = row;
            out_rows[(size_t)j * per_out].push_back(row + G.K);
            for (int k = 0; k <= kMaxK; ++k) cnt[row + k] = gr;
            row += kMaxK + 1;
          }
          d.flags = f;
        }
        if (streamed && p->launches.size() - p->stream_next >= 2) CHK(stream_flush(p));
      }
    }
    if (grad) {
      for (int i = L - 1; i >= 0; --i) {
        for (int g = P - 1; g >= 0; --g) {
          const Geo& G = p->geo[g];
          const int gr = begin_pass(G, true, ROLE_REV_HIST);
          for (int j = 0; j < nb; ++j) {
            PassDesc& d = base_desc(cur_geo, j);
            // the history state is read once here (5345-5352 circuits/s
            // for the 512-circuit batch with ordinary loads, 5352-5378 with these)
            uint32_t f = F_LOAD_A | F_ROT_A | F_LOAD_B | F_ROT_B | F_STORE_B | F_REDUCE | F_NT_A | phase_flag;
            d.a_in = hist(j, i, g);
            d.b_in = lam(j);
            d.b_out = lam(j);
            d.row_base = row;
            const GateAB ab = gab[((size_t)j * L + i) * 2 + 1];
            for (int lb = G.own_lo; lb < G.own_hi; ++lb) {
              d.gtype[lb] = G_X;
              d.ga[lb] = ab.a;
              d.gb[lb] = ab.b;
              d.own[lb] = 1;
              out_rows[(size_t)j * per_out + 1 + 2 * i].push_back(row + lb);
            }
            if (!phase_last && g == 0) {
              f |= F_DRED_POST | F_PHASE_POST;
              d.phase_gamma = -gammas[(c0 + j) * L + i];
              d.lut = lut_ptr(j, i, 1);
              out_rows[(size_t)j * per_out + 2 + 2 * i].push_back(row + G.K);
            }
            if (phase_last && g == P - 1 && i + 1 < L) {
              // the forward pass ended with the phase of layer i + 1
              f |= F_PHASE_PRE | F_DRED_PRE;
              d.phase_gamma = -gammas[(c0 + j) * L + i + 1];
              d.lut = lut_ptr(j, i + 1, 1);
              out_rows[(size_t)j * per_out + 2 + 2 * (i + 1)].push_back(row + G.K);
            }
            if (phase_last && g == 0 && i == 0) {
              // the first layer's phase opened this pass: psi is rotated back
              // to the initial state times that phase (general form), and
              // lambda is not needed after it
              f = (f | F_DRED_POST) & ~F_STORE_B;
              out_rows[(size_t)j * per_out + 2].push_back(row + G.K);
            }
            d.flags = f;
            for (int k = 0; k <= kMaxK; ++k) cnt[row + k] = gr;
            row += kMaxK + 1;
          }
          if (streamed && p->launches.size() - p->stream_next >= 2) CHK(stream_flush(p));
        }
      }
    }
    std::vector<double> scale(out_rows.size(), 1.0);
    if (!grad) {
      for (int j = 0; j < nb; ++j)
        for (int k = 1; k < per_out; ++k) out_rows[(size_t)j * per_out + k].clear();
    }
    if (streamed) {
      CHK(stream_flush(p));
    } else {
      CHK(run_launches(p, nd));
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    CHK(collect_profile(p));
    CHK(upload_finalize(p, out_rows, cnt, scale));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    total_ms += ms;
    const int nout = nb * per_out;
    HIPCHK(hipMemcpyAsync(p->h_out, p->d_out, sizeof(double) * nout, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int j = 0; j < nb; ++j) {
      const double* o = p->h_out + (size_t)j * per_out;
      E_out[c0 + j] = o[0];
      if (grad)
        for (int k = 0; k < 2 * L; ++k) grad_out[(size_t)(c0 + j) * 2 * L + k] = o[1 + k];
    }
    p->last_state_ptr = hist(0, L - 1, P - 1);
    p->state_sign = false;
    for (int i = 0; i < L; ++i)
      if (n % 2 == 1) p->state_sign ^= gate_flips(betas[(size_t)c0 * L + i], 1.0);
  }
  p->last_ms = total_ms;
  return QG_OK;
}

extern "C" int qg_qaoa_grad_run(qg_plan* p, int B, const qg_term* terms, const int32_t* term_off,
                                const double* betas, const double* gammas, double* E_out, double* grad_out) {
  return qaoa_run(p, B, terms, term_off, betas, gammas, nullptr, E_out, grad_out, true);
}

extern "C" int qg_qaoa_grad_run_ini(qg_plan* p, int B, const qg_term* terms, const int32_t* term_off,
                                    const double* betas, const double* gammas, const double* ini_states,
                                    double* E_out, double* grad_out) {
  return qaoa_run(p, B, terms, term_off, betas, gammas, ini_states, E_out, grad_out, true);
}

extern "C" int qg_qaoa_expec_val(qg_plan* p, int B, const qg_term* terms, const int32_t* term_off,
                                 const double* betas, const double* gammas, double* E_out) {
  return qaoa_run(p, B, terms, term_off, betas, gammas, nullptr, E_out, nullptr, false);
}

extern "C" int qg_qaoa_expec_val_ini(qg_plan* p, int B, const qg_term* terms, const int32_t* term_off,
                                     const double* betas, const double* gammas, const double* ini_states,
                                     double* E_out) {
  return qaoa_run(p, B, terms, term_off, betas, gammas, ini_states, E_out, nullptr, false);
}

// ---------------------------------------------------------------------------
// device-resident State
// ---------------------------------------------------------------------------
struct qg_state {
  qg_ctx* ctx = nullptr;
  int n = 0, dtype = 0;
  void* v = nullptr;
  void* tmp = nullptr;
  uint32_t* d_map = nullptr;
  DTerm* d_terms = nullptr;
  DTerm* d_xy = nullptr;
  int max_terms = 0;
  double* d_tab = nullptr;
  double* d_dot = nullptr;
};

static int state_init(qg_state* s, char ini) {
  const int n = s->n;
  std::vector<double> tab(n + 1, 0.0);
  if (ini == '0') {
    tab[0] = 1.0;
  } else if (ini == '+') {
    for (int k = 0; k <= n; ++k) tab[k] = std::pow(2.0, -0.5 * n);
  } else {
    return fail(QG_ERR_INVALID, std::string("Invalid initialization format ") + ini + ".");
  }
  HIPCHK(hipMemcpyAsync(s->d_tab, tab.data(), sizeof(double) * (n + 1), hipMemcpyHostToDevice, s->ctx->stream));
  HIPCHK(simple_init(s->dtype, s->v, n, s->d_tab, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return QG_OK;
}

extern "C" int qg_state_destroy(qg_state* s) {
  if (!s) return QG_OK;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  (void)hipFree(s->v);
  (void)hipFree(s->tmp);
  (void)hipFree(s->d_map);
  (void)hipFree(s->d_terms);
  (void)hipFree(s->d_xy);
  (void)hipFree(s->d_tab);
  (void)hipFree(s->d_dot);
  delete s;
  return QG_OK;
}

extern "C" int qg_state_create(qg_ctx* ctx, int n, int dtype, char ini, qg_state** out) {
  if (!ctx || !out) return fail(QG_ERR_INVALID, "NULL argument");
  if (n < 1 || n > 31) return fail(QG_ERR_INVALID, "State: qubit number must be in [1, 31]");
  if (dtype != QG_C128 && dtype != QG_C64) return fail(QG_ERR_INVALID, "bad dtype");
  if (ini != '0' && ini != '+') return fail(QG_ERR_INVALID, std::string("Invalid initialization format ") + ini + ".");
  HIPCHK(hipSetDevice(ctx->device));
  qg_state* s = new qg_state();
  s->ctx = ctx;
  s->n = n;
  s->dtype = dtype;
  s->max_terms = std::max(64, n * n);
  const size_t bytes = (size_t)(1ull << n) * (dtype == QG_C64 ? 8 : 16);
  hipError_t e = hipMalloc(&s->v, bytes);
  if (e == hipSuccess) e = hipMalloc(&s->tmp, bytes);
  if (e == hipSuccess) e = hipMalloc(&s->d_map, sizeof(uint32_t) * n);
  if (e == hipSuccess) e = hipMalloc(&s->d_terms, sizeof(DTerm) * s->max_terms);
  if (e == hipSuccess) e = hipMalloc(&s->d_xy, sizeof(DTerm) * s->max_terms);
  if (e == hipSuccess) e = hipMalloc(&s->d_tab, sizeof(double) * (n + 1));
  if (e == hipSuccess) e = hipMalloc(&s->d_dot, sizeof(double) * (kDotBlocks + 1));
  if (e != hipSuccess) {
    qg_state_destroy(s);
    return fail(e == hipErrorOutOfMemory ? QG_ERR_NOMEM : QG_ERR_HIP, std::string("qg_state_create: ") + hipGetErrorString(e));
  }
  int r = state_init(s, ini);
  if (r != QG_OK) {
    qg_state_destroy(s);
    return r;
  }
  *out = s;
  return QG_OK;
}

extern "C" int qg_state_reset(qg_state* s, char ini) {
  if (!s) return fail(QG_ERR_INVALID, "NULL state");
  HIPCHK(hipSetDevice(s->ctx->device));
  return state_init(s, ini);
}

extern "C" int qg_state_set(qg_state* s, const double* vec) {
  if (!s || !vec) return fail(QG_ERR_INVALID, "NULL argument");
  HIPCHK(hipSetDevice(s->ctx->device));
  const size_t N = 1ull << s->n;
  double* tmp = nullptr;
  HIPCHK(hipMalloc(&tmp, N * 16));
  hipError_t e = hipMemcpyAsync(tmp, vec, N * 16, hipMemcpyHostToDevice, s->ctx->stream);
  if (e == hipSuccess) e = simple_convert_in(s->dtype, tmp, s->v, s->n, s->ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(QG_ERR_HIP, std::string("qg_state_set: ") + hipGetErrorString(e));
  return QG_OK;
}

extern "C" int qg_state_get(qg_state* s, double* vec) {
  if (!s || !vec) return fail(QG_ERR_INVALID, "NULL argument");
  HIPCHK(hipSetDevice(s->ctx->device));
  const size_t N = 1ull << s->n;
  double* tmp = nullptr;
  HIPCHK(hipMalloc(&tmp, N * 16));
  hipError_t e = simple_convert_out(s->dtype, s->v, tmp, s->n, s->ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(vec, tmp, N * 16, hipMemcpyDeviceToHost, s->ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(QG_ERR_HIP, std::string("qg_state_get: ") + hipGetErrorString(e));
  return QG_OK;
}

extern "C" int qg_state_rot(qg_state* s, int axis, double angle, int q, int deriv) {
  if (!s) return fail(QG_ERR_INVALID, "NULL state");
  if (axis < 0 || axis > 2) return fail(QG_ERR_INVALID, "Invalid axis " + std::to_string(axis));
  if (q < 0 || q >= s->n) return fail(QG_ERR_INVALID, "qubit index out of range");
  HIPCHK(hipSetDevice(s->ctx->device));
  const int type = axis == 0 ? G_X : axis == 1 ? G_Y : G_Z;
  double c = std::cos(0.5 * angle), sn = std::sin(0.5 * angle);
  if (deriv) {  // d/da (cos(a/2) I - i sin(a/2) P) = -sin(a/2)/2 I - i cos(a/2)/2 P
    const double c2 = -0.5 * sn, s2 = 0.5 * c;
    c = c2;
    sn = s2;
  }
  HIPCHK(simple_rot1(s->dtype, s->v, s->n, s->n - 1 - q, type, c, sn, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return QG_OK;
}

extern "C" int qg_state_rot_all(qg_state* s, int axis, double angle) {
  if (!s) return fail(QG_ERR_INVALID, "NULL state");
  for (int q = 0; q < s->n; ++q) CHK(qg_state_rot(s, axis, angle, q, 0));
  return QG_OK;
}

static int state_permute(qg_state* s, const std::vector<uint32_t>& cols) {
  HIPCHK(hipMemcpyAsync(s->d_map, cols.data(), sizeof(uint32_t) * s->n, hipMemcpyHostToDevice, s->ctx->stream));
  HIPCHK(simple_permute(s->dtype, s->v, s->tmp, s->n, s->d_map, s->ctx->stream));
  std::swap(s->v, s->tmp);
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return QG_OK;
}

extern "C" int qg_state_cnot_ladder(qg_state* s, int stacking) {
  if (!s) return fail(QG_ERR_INVALID, "NULL state");
  if (stacking != 0 && stacking != 1) return fail(QG_ERR_INVALID, "stacking must be 0 or 1");
  HIPCHK(hipSetDevice(s->ctx->device));
  // gather: (L psi)(y) = psi(L^{-1} y); L0^{-1} = L1 and L1^{-1} = L0
  std::vector<uint32_t> cols;
  ladder_map(s->n, stacking == 0 ? 1 : 0, cols);
  return state_permute(s, cols);
}

extern "C" int qg_state_cnot(qg_state* s, int control, int target) {
  if (!s) return fail(QG_ERR_INVALID, "NULL state");
  const int n = s->n;
  if (control < 0 || target < 0 || control >= n || target >= n || control == target)
    return fail(QG_ERR_INVALID, "Invalid CNOT indecies " + std::to_string(control) + " and " +
                                    std::to_string(target) + ", for " + std::to_string(n) + " qubits.");
  HIPCHK(hipSetDevice(s->ctx->device));
  std::vector<uint32_t> cols(n);
  const int bc = n - 1 - control, bt = n - 1 - target;
  for (int k = 0; k < n; ++k) cols[k] = (1u << k) | (k == bc ? (1u << bt) : 0u);
  return state_permute(s, cols);
}

extern "C" int qg_state_cz_ladder(qg_state* s) {
  if (!s) return fail(QG_ERR_INVALID, "NULL state");
  HIPCHK(hipSetDevice(s->ctx->device));
  HIPCHK(simple_diag_phase(s->dtype, s->v, s->n, s->d_terms, 0, 0.0, 1, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return QG_OK;
}

static int state_upload_obs(qg_state* s, const qg_term* terms, int nterms, ObsSplit& o) {
  CHK(split_observable(s->n, terms, nterms, o));
  if ((int)o.diag.size() > s->max_terms || (int)o.xy.size() > s->max_terms)
    return fail(QG_ERR_INVALID, "too many observable terms");
  if (!o.diag.empty())
    HIPCHK(hipMemcpyAsync(s->d_terms, o.diag.data(), sizeof(DTerm) * o.diag.size(), hipMemcpyHostToDevice,
                          s->ctx->stream));
  if (!o.xy.empty())
    HIPCHK(hipMemcpyAsync(s->d_xy, o.xy.data(), sizeof(DTerm) * o.xy.size(), hipMemcpyHostToDevice, s->ctx->stream));
  return QG_OK;
}

extern "C" int qg_state_exp_diag(qg_state* s, const qg_term* terms, int nterms, double angle) {
  if (!s) return fail(QG_ERR_INVALID, "NULL state");
  HIPCHK(hipSetDevice(s->ctx->device));
  ObsSplit o;
  CHK(state_upload_obs(s, terms, nterms, o));
  HIPCHK(simple_diag_phase(s->dtype, s->v, s->n, s->d_terms, (int)o.diag.size(), angle, 0, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return QG_OK;
}

extern "C" int qg_state_multiply_observable(qg_state* s, const qg_term* terms, int nterms) {
  if (!s) return fail(QG_ERR_INVALID, "NULL state");
  HIPCHK(hipSetDevice(s->ctx->device));
  ObsSplit o;
  CHK(state_upload_obs(s, terms, nterms, o));
  HIPCHK(simple_obs_apply(s->dtype, s->v, s->tmp, s->n, s->d_terms, (int)o.diag.size(), s->d_xy, (int)o.xy.size(),
                          s->ctx->stream));
  std::swap(s->v, s->tmp);
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return QG_OK;
}

extern "C" int qg_state_expec_val(qg_state* s, const qg_term* terms, int nterms, double* out) {
  if (!s || !out) return fail(QG_ERR_INVALID, "NULL argument");
  HIPCHK(hipSetDevice(s->ctx->device));
  ObsSplit o;
  CHK(state_upload_obs(s, terms, nterms, o));
  HIPCHK(simple_obs_apply(s->dtype, s->v, s->tmp, s->n, s->d_terms, (int)o.diag.size(), s->d_xy, (int)o.xy.size(),
                          s->ctx->stream));
  HIPCHK(simple_dot(s->dtype, s->v, s->tmp, s->n, 0, s->d_dot, s->d_dot + kDotBlocks, s->ctx->stream));
  HIPCHK(hipMemcpyAsync(out, s->d_dot + kDotBlocks, sizeof(double), hipMemcpyDeviceToHost, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return QG_OK;
}

extern "C" int qg_state_term_prob(qg_state* s, const qg_term* terms, int nterms, double* prob) {
  if (!s) return fail(QG_ERR_INVALID, "NULL state");
  return term_prob(s->ctx, s->dtype, s->v, s->n, terms, nterms, prob);
}

extern "C" int qg_state_copy(qg_state* dst, const qg_state* src) {
  if (!dst || !src) return fail(QG_ERR_INVALID, "NULL state");
  if (dst->n != src->n || dst->dtype != src->dtype)
    return fail(QG_ERR_INVALID, "qg_state_copy: states differ in qubit number or dtype");
  if (dst == src) return QG_OK;
  HIPCHK(hipSetDevice(dst->ctx->device));
  const size_t bytes = ((size_t)1 << dst->n) * (dst->dtype == QG_C64 ? 8 : 16);
  // both contexts' earlier work on the two vectors is finished first
  HIPCHK(hipStreamSynchronize(src->ctx->stream));
  HIPCHK(hipMemcpyAsync(dst->v, src->v, bytes, hipMemcpyDeviceToDevice, dst->ctx->stream));
  HIPCHK(hipStreamSynchronize(dst->ctx->stream));
  return QG_OK;
}

extern "C" int qg_state_sample_basis(qg_state* st, const int32_t* order, int shots, const double* u, int32_t* idx) {
  if (!st || !u || !idx) return fail(QG_ERR_INVALID, "NULL argument");
  if (shots < 1) return fail(QG_ERR_INVALID, "shots must be >= 1");
  HIPCHK(hipSetDevice(st->ctx->device));
  SampleBufs b;
  CHK(sample_setup(st->n, order, shots, u, b));
  HIPCHK(simple_sample_basis(st->dtype, st->v, st->n, b.order, b.u, shots, b.idx, b.mass, st->ctx->stream));
  HIPCHK(hipStreamSynchronize(st->ctx->stream));
  HIPCHK(hipMemcpy(idx, b.idx, (size_t)shots * 4, hipMemcpyDeviceToHost));
  return QG_OK;
}

extern "C" int qg_state_norm_error(qg_state* s, double* out) {
  if (!s || !out) return fail(QG_ERR_INVALID, "NULL argument");
  HIPCHK(hipSetDevice(s->ctx->device));
  double nrm2 = 0.0;
  HIPCHK(simple_dot(s->dtype, s->v, s->v, s->n, 1, s->d_dot, s->d_dot + kDotBlocks, s->ctx->stream));
  HIPCHK(hipMemcpyAsync(&nrm2, s->d_dot + kDotBlocks, sizeof(double), hipMemcpyDeviceToHost, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  *out = 1.0 - std::sqrt(nrm2);
  return QG_OK;
}

extern "C" int qg_state_scale(qg_state* s, double re, double im) {
  if (!s) return fail(QG_ERR_INVALID, "NULL state");
  HIPCHK(hipSetDevice(s->ctx->device));
  HIPCHK(simple_scale(s->dtype, s->v, s->n, re, im, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return QG_OK;
}

// Lays a McClean plan out again after a change of its shape parameters.
// Lays a McClean plan out again (new shape or tile options).  An owned
// workspace too small for the new layout is freed (the next run allocates);
// a supplied one is detached when detach_small is set, else it is an error.
static int relayout(qg_plan* p, bool detach_small = false) {
  HIPCHK(hipSetDevice(p->ctx->device));
  HIPCHK(hipStreamSynchronize(p->ctx->stream));
  drop_graph(p);
  CHK(mcclean_layout(p));
  if (p->ws && p->ws_bytes > p->ws_cap) {
    if (p->ws_owned) {
      (void)hipFree(p->ws);
    } else if (!detach_small) {
      return fail(QG_ERR_INVALID, "the new layout needs a larger workspace than the one supplied");
    }
    p->ws = nullptr;
    p->ws_cap = 0;
    p->ws_owned = false;
  }
  p->last_state_ptr = nullptr;
  return QG_OK;
}

extern "C" int qg_plan_reshape(qg_plan* p, int n, int L, int ckpt_every) {
  if (!p) return fail(QG_ERR_INVALID, "NULL plan");
  if (p->kind != PK_MCCLEAN || p->mc_batch) return fail(QG_ERR_UNSUPPORTED, "qg_plan_reshape applies to McClean plans");
  if (n < 1 || n > kMaxQubits) return fail(QG_ERR_INVALID, "qubit number must be in [1, 32]");
  if (L < 1) return fail(QG_ERR_INVALID, "layer number must be >= 1");
  if (ckpt_every < QG_HISTORY) return fail(QG_ERR_INVALID, "ckpt_every must be >= 0 or QG_HISTORY");
  p->n = n;
  p->L = L;
  p->ckpt_req = ckpt_every;
  return relayout(p, true);
}

extern "C" int qg_plan_set_graph(qg_plan* p, int on) {
  if (!p) return fail(QG_ERR_INVALID, "NULL plan");
  HIPCHK(hipSetDevice(p->ctx->device));
  HIPCHK(hipStreamSynchronize(p->ctx->stream));
  drop_graph(p);
  p->use_graph = on != 0;
  return QG_OK;
}

extern "C" int qg_plan_set_lds64(qg_plan* p, int mode) {
  if (!p) return fail(QG_ERR_INVALID, "NULL plan");
  if (mode < -1 || mode > 2) return fail(QG_ERR_INVALID, "mode must be -1, 0, 1 or 2");
  if (p->mc_batch) return fail(QG_ERR_UNSUPPORTED, "batch plans keep the default tile shape");
  if (p->kind != PK_MCCLEAN) return fail(QG_ERR_UNSUPPORTED, "FP64 tiles apply to McClean plans");
  p->lds64_req = mode;
  return relayout(p);
}

extern "C" int qg_plan_lds64(const qg_plan* p, int* out) {
  if (!p || !out) return fail(QG_ERR_INVALID, "NULL argument");
  *out = !p->lds64 ? 0 : p->lds64_adj ? 1 : 2;
  return QG_OK;
}

extern "C" int qg_plan_tune(qg_plan* p, int kmax, int cmin, int r_fwd, int r_rev) {
  if (!p) return fail(QG_ERR_INVALID, "NULL plan");
  if (kmax < 4 || kmax > kMaxK) return fail(QG_ERR_INVALID, "kmax must be in [4, 14]");
  if (kmax > 13 && p->dtype != QG_C64) return fail(QG_ERR_INVALID, "kmax 14 needs a QG_C64 plan");
  if (cmin < 0 || cmin >= kmax) return fail(QG_ERR_INVALID, "cmin must be in [0, kmax)");
  if (r_fwd != 4 || r_rev != 4) return fail(QG_ERR_INVALID, "r_fwd, r_rev must be 4 (16 amplitudes per thread)");
  if (p->mc_batch) return fail(QG_ERR_UNSUPPORTED, "batch plans keep the default tile shape");
  if (p->kind != PK_MCCLEAN) return fail(QG_ERR_UNSUPPORTED, "tuning applies to McClean plans");
  p->kmax = kmax;
  p->kmax_auto = false;
  p->cmin = cmin;
  p->r_fwd = r_fwd;
  p->r_rev = r_rev;
  return relayout(p);
}

extern "C" int qg_plan_set_fused(qg_plan* p, int on) {
  if (!p) return fail(QG_ERR_INVALID, "NULL plan");
  p->fused = on != 0;
  return QG_OK;
}

extern "C" int qg_plan_ckpt_every(qg_plan* p, int* out) {
  if (!p || !out) return fail(QG_ERR_INVALID, "NULL argument");
  *out = p->history ? 0 : p->ckpt;
  return QG_OK;
}
