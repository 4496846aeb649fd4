// Context management and the extern "C" surface of libsoliton.so.
//
// One context owns the device state of one relaxation: ping-pong Field and
// Velocity buffers, the three stage slopes l1..l3, per-tile partial sums, a
// DevState block with the arrest bookkeeping, and per-step history.  The
// theory-specific kernels are reached through a TheoryOps table.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/soliton.h"
#include "relax_kernels.h"

namespace ss {
const TheoryOps* ops_chiral_magnet();
const TheoryOps* ops_ginzburg_landau();
const TheoryOps* ops_baby_skyrme();
const TheoryOps* ops_anisotropic_sc();
const TheoryOps* ops_bec();
const TheoryOps* ops_mcsh();
const TheoryOps* ops_fs();
const TheoryOps* ops_st();
const TheoryOps* ops_lc();
}  // namespace ss

using namespace ss;

static thread_local std::string g_global_err;

// Counter-based uniform in [0, 1): splitmix64 of (seed, counter), top 53 bits.
__device__ __forceinline__ double hash_uniform(unsigned long long seed, unsigned long long ctr) {
  unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (ctr + 1ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// Components [c0, c1) += U(-amp, amp) or N(0, amp^2) (Box-Muller on two
// counters), then |first rn components| rescaled to rv where nonzero.
template <typename R>
__global__ void __launch_bounds__(256) k_noise(R* F, long long N, int c0, int c1, int gaussian, double amp,
                                               unsigned long long seed, int rn, double rv) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
    for (int c = c0; c < c1; ++c) {
      const unsigned long long ctr = 2ull * ((unsigned long long)c * (unsigned long long)N + (unsigned long long)i);
      double d;
      if (gaussian) {
        const double u1 = 1.0 - hash_uniform(seed, ctr), u2 = hash_uniform(seed, ctr + 1ull);
        d = amp * sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
      } else {
        d = amp * (2.0 * hash_uniform(seed, ctr) - 1.0);
      }
      F[(long long)c * N + i] = R((double)F[(long long)c * N + i] + d);
    }
    if (rn > 0) {
      double s2 = 0.0;
      for (int c = 0; c < rn; ++c) s2 += (double)F[(long long)c * N + i] * (double)F[(long long)c * N + i];
      const double sn = sqrt(s2);
      if (sn != 0.0)
        for (int c = 0; c < rn; ++c) F[(long long)c * N + i] = R((double)F[(long long)c * N + i] * (rv / sn));
    }
  }
}

static const TheoryOps* theory_ops(int id) {
  switch (id) {
    case SS_THEORY_CHIRAL_MAGNET: return ops_chiral_magnet();
    case SS_THEORY_GINZBURG_LANDAU: return ops_ginzburg_landau();
    case SS_THEORY_BABY_SKYRME: return ops_baby_skyrme();
    case SS_THEORY_ANISOTROPIC_SC: return ops_anisotropic_sc();
    case SS_THEORY_BEC: return ops_bec();
    case SS_THEORY_MAXWELL_CHERN_SIMONS_HIGGS: return ops_mcsh();
    case SS_THEORY_FERROMAGNETIC_SC: return ops_fs();
    case SS_THEORY_SPIN_TRIPLET: return ops_st();
    case SS_THEORY_LIQUID_CRYSTAL: return ops_lc();
    default: return nullptr;
  }
}

// ---------------------------------------------------------------------------
// Theory-independent kernels.

__global__ void k_begin_call(DevState* st, double prev_energy) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->prev_energy = prev_energy;
    st->eslot[0] = st->eslot[1] = prev_energy;  // the first kill test compares with the caller's energy
    st->kill = 0;
    st->step = 0;
  }
}

// Threads of the one-block finalize kernels of a step: four waves (a 1024-
// thread block measured 1 us slower per step on a 320^2 lattice, where these
// launches are on the critical path).
constexpr int kFinThreads = 256;

// Sum of per-tile energy partials (fixed order: Kahan per thread, then a
// shuffle tree), then the arrest decision of the previous step
// (core/integrator.py:338-343): kill iff killkinen and E_new > E_prev.
__global__ void __launch_bounds__(kFinThreads)
k_fin_energy(const double* part, int n, DevState* st, double* histE, int cap, int killkinen, int parity) {
  __shared__ double red[16];
  double s = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < n; i += kFinThreads) {
    const double y = part[i] - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
  }
  const double E = block_sum<kFinThreads>(s, red);
  if (threadIdx.x == 0) {
    st->last_energy = E;
    const int step = st->step;
    if (step > 0) {
      st->kill = (killkinen && E > st->prev_energy) ? 1 : 0;
      st->prev_energy = E;
      st->eslot[parity] = E;
      if (step - 1 < cap) histE[step - 1] = E;
    } else {
      st->kill = 0;
    }
  }
}

__global__ void __launch_bounds__(kFinThreads)
k_fin_err(const double* p0, const double* p1, int n, DevState* st, double* histErr, int cap) {
  __shared__ double red[16];
  double m0 = -1.0e300, m1 = 0.0;
  for (int i = threadIdx.x; i < n; i += kFinThreads) {
    m0 = nmax(m0, p0[i]);
    m1 = nmax(m1, p1[i]);
  }
  const double a = block_max<kFinThreads>(m0, red);
  const double b = block_max<kFinThreads>(m1, red);
  if (threadIdx.x == 0) {
    st->err_max = a;
    st->err_absmax = b;
    const int step = st->step;
    if (step < cap) histErr[step] = a;
    st->step = step + 1;
  }
}

// Two finalizes in one launch: err of step n (as k_fin_err, from the n_err
// partials p0, p1), then the energy of the field after it and the arrest
// decision (as k_fin_energy with the step counter already advanced, from the
// n_en partials p2; the per-thread Kahan order is k_fin_energy's, so the energy
// is the same to the bit).  The 4+1 schedule calls it after its sweep; the
// paired schedules after the 1+2 sweep of step n + 1.
__global__ void __launch_bounds__(kFinThreads)
k_fin_err_energy(const double* p0, const double* p1, int n_err, const double* p2, int n_en, DevState* st,
                 double* histErr, double* histE, int cap, int killkinen, int parity) {
  __shared__ double red[16];
  double m0 = -1.0e300, m1 = 0.0, s = 0.0, c = 0.0;
  const int n = n_err > n_en ? n_err : n_en;
  for (int i = threadIdx.x; i < n; i += kFinThreads) {
    if (i < n_err) {
      m0 = nmax(m0, p0[i]);
      m1 = nmax(m1, p1[i]);
    }
    if (i < n_en) {
      const double y = p2[i] - c;
      const double t = s + y;
      c = (t - s) - y;
      s = t;
    }
  }
  const double a = block_max<kFinThreads>(m0, red);
  const double b = block_max<kFinThreads>(m1, red);
  const double E = block_sum<kFinThreads>(s, red);
  if (threadIdx.x == 0) {
    st->err_max = a;
    st->err_absmax = b;
    const int step = st->step;
    if (step < cap) histErr[step] = a;
    st->step = step + 1;
    st->last_energy = E;
    st->kill = (killkinen && E > st->prev_energy) ? 1 : 0;
    st->prev_energy = E;
    st->eslot[parity] = E;
    if (step < cap) histE[step] = E;
  }
}

// Plain sum of partials into out[0] (observables).
__global__ void __launch_bounds__(1024) k_fin_sum(const double* part, int n, double* out) {
  __shared__ double red[16];
  double s = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) {
    const double y = part[i] - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
  }
  const double E = block_sum<1024>(s, red);
  if (threadIdx.x == 0) out[0] = E;
}

// Velocity kill of the last step of a call, applied in memory.
template <typename R>
__global__ void k_cond_zero(R* V, long long n, const DevState* st) {
  if (st->kill == 0) return;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    V[i] = R(0);
}

template <typename A, typename B>
__global__ void k_convert(const A* src, B* dst, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = B(src[i]);
}

// L2 normalisation of a field (core/integrator.py:295-321 with the theory's
// compute_norm / do_norm_kernel): sum |f|^2 gv over all components, then
// f /= sqrt(sum).  Partial sums per block, one-block finalize.
__global__ void __launch_bounds__(256) k_l2_partial(const double* G, long long n, double gv, double* part) {
  __shared__ double red[4];
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    s += G[i] * G[i] * gv;
  const double b = block_sum<256>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}

__global__ void __launch_bounds__(1024) k_fin_sqrt(const double* part, int n, double* out) {
  __shared__ double red[16];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) s += part[i];
  const double t = block_sum<1024>(s, red);
  if (threadIdx.x == 0) out[0] = sqrt(t);
}

__global__ void k_div_scalar(double* G, long long n, const double* s) {
  const double d = s[0];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    G[i] = G[i] / d;
}

// Compensated FP32 form: split an FP64 field into hi = fl32(F), lo = fl32(F - hi);
// join hi + lo back into FP64; fold lo into hi (to the plain FP32 form).
__global__ void k_split(const double* src, float* hi, float* lo, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double v = src[i];
    const float h = (float)v;
    hi[i] = h;
    lo[i] = (float)(v - (double)h);
  }
}
__global__ void k_join(const float* hi, const float* lo, double* dst, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = (double)hi[i] + (double)lo[i];
}
__global__ void k_fold(float* hi, const float* lo, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    hi[i] = (float)((double)hi[i] + (double)lo[i]);
}

// ---------------------------------------------------------------------------
// Library-held scratch of ss_energy_dev_noalloc: the per-tile energy partials
// live in this file-scope device array instead of memory the library
// allocates, so the call works where hipMalloc may not run (inside HIP graph
// capture, allocation-free serving loops).  Each GPU holds its own copy for
// the life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (the array's name is a host-side
// stand-in, not a device address).
constexpr int kScratchDoubles = 1 << 18;  // 2 MiB: energy partials of up to 262144 tiles
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static double* g_scratch[kMaxDevices] = {nullptr};
static std::mutex g_scratch_mu;

// Packed p_f entries 0..5 are the lattice geometry (core/params.py:247-248);
// the theories' coupling constants follow.
constexpr int kCorePF = 6;

// Resolves the scratch array's address on device ``dev`` (once per device,
// with that device current); 0 on success.
static int scratch_setup(int dev, std::string* err) {
  if (dev < 0 || dev >= kMaxDevices) {
    if (err) *err = "device id out of range for the static scratch table";
    return -1;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[dev]) return 0;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) {
    if (err) *err = "hipGetDevice failed";
    return -2;
  }
  if (cur != dev && hipSetDevice(dev) != hipSuccess) {
    if (err) *err = "hipSetDevice failed";
    return -2;
  }
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (cur != dev) hipSetDevice(cur);
  if (e != hipSuccess || !ptr) {
    if (err) *err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
    return -2;
  }
  g_scratch[dev] = (double*)ptr;
  return 0;
}

struct ss_ctx {
  int theory = 0;
  const TheoryOps* ops = nullptr;
  int prec = 0;
  int var_mode = 0;
  int dev = 0;  // the device current at creation
  int32_t pi[kMaxPI] = {0};
  int ni = 0;
  double pf[kMaxPF] = {0};
  int nfl = 0;
  alignas(16) unsigned char tp[512];
  Geom g{};
  hipStream_t stream = nullptr;
  void* F[2] = {nullptr, nullptr};
  void* V[2] = {nullptr, nullptr};
  void* l[3] = {nullptr, nullptr, nullptr};
  int cur = 0;
  int ntiles = 0;          // capacity of each partials row (the most tiles over the shape variants)
  int nt0 = 0;             // tiles of the default shape (energy, charge and pass-0 kernels)
  int tv[5] = {0, 0, 0, 0, 0};  // tile-shape variant of each fused pass (index = pass)
  int fuse23 = 0;          // passes 2 and 3 as one sweep (variant tv[2]: a strip, or a tile where two fields fit)
  int fuse41 = 0;          // pass 4 and the next step's pass 1 as one sweep (variant tv[4], at least 3)
  int pairs = 0;           // paired schedule: passes 1+2 (variant tv[1]) and 3+4 (tv[4]) as two sweeps
  void* l1alt = nullptr;   // with fuse41: l1 of the steps that start from parity 1
  float* Flo = nullptr;    // SS_PREC_FP32C: low words of the field (nf*N, updated in place by pass 4)
  double* part = nullptr;  // 3 x ntiles
  DevState* st = nullptr;
  double* hist = nullptr;  // 2 x hist_cap
  int hist_cap = 0;
  int last_nsteps = 0;
  double* dens = nullptr;
  double* scal = nullptr;  // scalar results on device
  DevState* host_st = nullptr;  // pinned
  // exact-adjoint scratch (allocated on first use): stage field, gradient,
  // the two transposed-stencil inputs, and per-block partials of the combine
  void* temp = nullptr;
  double* G = nullptr;
  double* Px = nullptr;
  double* Py = nullptr;
  double* xpart = nullptr;
  int xpart_n = 0;
  int graph_steps = 0;
  hipGraphExec_t graph[2] = {nullptr, nullptr};  // indexed by starting parity
  hipEvent_t ev_dev = nullptr;  // orders the device-pointer calls against the context's stream
  double* pgpart = nullptr;     // partials of ss_energy_param_grad (allocated on first use)
  size_t pgpart_n = 0;
  std::string err;
};

// Status codes: -1 invalid argument / unsupported request, -2 HIP runtime error.
static int vfail(ss_ctx* c, int code, const char* fmt, va_list ap) {
  char buf[512];
  vsnprintf(buf, sizeof(buf), fmt, ap);
  if (c) c->err = buf; else g_global_err = buf;
  return code;
}
static int fail(ss_ctx* c, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  const int rc = vfail(c, -1, fmt, ap);
  va_end(ap);
  return rc;
}
static int hipfail(ss_ctx* c, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  const int rc = vfail(c, -2, fmt, ap);
  va_end(ap);
  return rc;
}

#define HIPCHK(c, expr)                                                                     \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) return hipfail((c), "%s failed: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

static size_t elem_size(int prec) { return prec ? sizeof(float) : sizeof(double); }

static int resolve_geom(ss_ctx* c, const int32_t* pi, int ni, const double* pf, int nfl, Geom* g) {
  if (ni < 10 || nfl < 6) return fail(c, "packed parameters too short: need p_i[0..9], p_f[0..5] (got %d, %d)", ni, nfl);
  if (ni > kMaxPI || nfl > kMaxPF) return fail(c, "packed parameters too long (%d ints, %d doubles)", ni, nfl);
  g->X = pi[0];
  g->Y = pi[1];
  g->halo = pi[2];
  g->nf = pi[4];
  if (g->X < 1 || g->Y < 1) return fail(c, "xlen and ylen must be positive (got %d, %d)", g->X, g->Y);
  if (g->halo < 2) return fail(c, "halo must be >= 2 for the 5-point stencils (got %d)", g->halo);
  if (pi[3] != 2) return fail(c, "number_coordinates must be 2 (got %d)", pi[3]);
  g->N = (long long)g->X * g->Y;
  if ((long long)pi[5] != g->N) return fail(c, "dim_grid p_i[5]=%d does not equal xlen*ylen=%lld", pi[5], g->N);
  g->lsx = pf[2];
  g->lsy = pf[3];
  g->gv = pf[4];
  g->dt = pf[5];
  if (!(g->lsx > 0.0) || !(g->lsy > 0.0)) return fail(c, "lattice spacings must be positive");
  g->inv_lsx = 1.0 / g->lsx;
  g->inv_lsy = 1.0 / g->lsy;
  g->inv_lsx2 = 1.0 / (g->lsx * g->lsx);
  g->inv_lsy2 = 1.0 / (g->lsy * g->lsy);
  g->inv_2gv = 1.0 / (2.0 * g->gv);
  g->inv_2lsx2 = 1.0 / (2.0 * g->lsx * g->lsx);
  g->inv_2lsy2 = 1.0 / (2.0 * g->lsy * g->lsy);
  g->killkinen = pi[7];
  g->unit_mag = pi[9];
  return 0;
}

static void free_graphs(ss_ctx* c) {
  for (int i = 0; i < 2; ++i) {
    if (c->graph[i]) hipGraphExecDestroy(c->graph[i]);
    c->graph[i] = nullptr;
  }
}

static void free_fields(ss_ctx* c) {
  for (int i = 0; i < 2; ++i) {
    if (c->F[i]) hipFree(c->F[i]);
    if (c->V[i]) hipFree(c->V[i]);
    c->F[i] = c->V[i] = nullptr;
  }
  for (int i = 0; i < 3; ++i) {
    if (c->l[i]) hipFree(c->l[i]);
    c->l[i] = nullptr;
  }
  if (c->l1alt) hipFree(c->l1alt);
  c->l1alt = nullptr;
  if (c->Flo) hipFree(c->Flo);
  c->Flo = nullptr;
}

// The second l1 buffer of the fused 4+1 schedule: a step reads l1 of its
// starting parity and writes the next step's l1 into the other one.
static int ensure_l1alt(ss_ctx* c) {
  if (c->l1alt) return 0;
  const size_t bytes = (size_t)c->g.nf * (size_t)c->g.N * elem_size(c->prec);
  HIPCHK(c, hipMalloc(&c->l1alt, bytes));
  HIPCHK(c, hipMemsetAsync(c->l1alt, 0, bytes, c->stream));
  return 0;
}

static void* l1_buf(ss_ctx* c, int parity) { return (c->fuse41 && parity) ? c->l1alt : c->l[0]; }

// The fused 4+1 sweep is available (not in the compensated form, whose pass 4
// forms the new field from F and its low words): in the variant of pass 4 or
// in strips.
static bool can41(const ss_ctx* c) {
  return c->prec != SS_PREC_FP32C &&
         (c->ops->fuse41_variant(c->prec, c->tv[4], c->g) || c->ops->can_fuse41(c->prec, c->g));
}

// Shape variant of the fused 4+1 sweep: tv[4] where the sweep exists in it,
// else the first strip variant.
static int fused41_v(const ss_ctx* c) {
  return c->ops->fuse41_variant(c->prec, c->tv[4], c->g) ? c->tv[4] : 3;
}

// The fused 4+1 schedule is active (reference flow only).
static bool use41(const ss_ctx* c) {
  return c->fuse41 && c->var_mode == SS_VAR_REFERENCE && !c->ops->l2_normalize && c->prec != SS_PREC_FP32C;
}

// Most tiles of the 4+1 sweep whose finalize is folded into the next step's
// 2+3 sweep: every workgroup of that sweep reduces all of the partials again,
// which pays only where the lattice is small and the one-block finalize launch
// is a sizeable part of the step.
constexpr int kFoldMaxTiles = 512;

// The 4+1 schedule with each step's finalize folded into the next step's 2+3
// sweep (not in the timed schedule, which separates every launch).
static bool fold_mode(const ss_ctx* c, bool timed) {
  return use41(c) && c->fuse23 && !timed && c->ops->ntiles(c->prec, fused41_v(c), c->g) <= kFoldMaxTiles;
}

// The paired schedule is active (reference flow, plain FP32 or FP64 form, and
// both sweeps fit the first strip variant).
static bool use_pairs(const ss_ctx* c) {
  return c->pairs && c->var_mode == SS_VAR_REFERENCE && !c->ops->l2_normalize && c->prec != SS_PREC_FP32C &&
         c->ops->pairs_variant(c->prec, 3, c->g);
}
// The compensated form's paired schedule: its pass 1 (id 8) carries the low
// words in the second ring half, so only passes 3 and 4 pair (id 11).
static bool use_comp_pairs(const ss_ctx* c) {
  return c->pairs && c->var_mode == SS_VAR_REFERENCE && !c->ops->l2_normalize && c->prec == SS_PREC_FP32C &&
         c->ops->pairs_variant(c->prec, 3, c->g);
}
// Strip variant of a paired sweep: v where both sweeps run in it, else 3.
static int pairs_v(const ss_ctx* c, int v) { return c->ops->pairs_variant(c->prec, v, c->g) ? v : 3; }
// The compensated form's 1+2 sweep (id 12: F, Temp2 and the low words in a
// three-part ring) runs in strip variant v.
static bool comp12_in(const ss_ctx* c, int v) {
  return c->prec == SS_PREC_FP32C && (c->ops->pairs_variant(c->prec, v, c->g) & 2) != 0;
}
// The compensated paired schedule runs passes 1+2 as that sweep (in the first
// strip variant at least), not as pass 1 and pass 2.
static bool comp12_ok(const ss_ctx* c) { return use_comp_pairs(c) && comp12_in(c, 3); }
static int comp12_v(const ss_ctx* c, int v) { return comp12_in(c, v) ? v : 3; }

// Paired schedules (plain, or compensated with its 1+2 sweep) finalize the err
// of a step together with the next step's energy: three launches per step
// instead of four.  Not in timed runs, which finalize every step on its own.
static bool defer_err(const ss_ctx* c, bool timed) { return !timed && (use_pairs(c) || comp12_ok(c)); }
// Calls whose first step runs outside the graphs: the 4+1 schedule (its first
// step carries pass 1) and the deferred paired schedules (no err pending in
// their first step).
static bool lead_step(const ss_ctx* c, bool timed) { return use41(c) || defer_err(c, timed); }

// Shape variant of the fused 2+3 sweep: tv[2] where the sweep exists in it,
// else the first strip variant.
static int fused23_v(const ss_ctx* c) {
  return c->ops->fuse23_variant(c->prec, c->tv[2], c->g) ? c->tv[2] : 3;
}

// Pass 1 of the step: in the compensated form the streaming pass 1 (id 8),
// which also takes the energy of F + low words, in a strip variant.
static int pass1_id(const ss_ctx* c) { return c->prec == SS_PREC_FP32C ? 8 : 1; }
static int pass1_variant(const ss_ctx* c) {
  return (c->prec == SS_PREC_FP32C && c->tv[1] < 3) ? 3 : c->tv[1];
}

// Energy partials of the field F (of F + its low words in the compensated
// form) into c->part; *nparts = their number.
static int energy_partials(ss_ctx* c, const void* F, int* nparts) {
  if (c->prec == SS_PREC_FP32C) {
    PassBufs b{};
    b.F = F;
    b.Flo = c->Flo;
    b.lout = c->l[0];  // l1 of this field: scratch (every step recomputes it)
    b.part0 = c->part;
    b.st = c->st;
    const int v = pass1_variant(c);
    HIPCHK(c, c->ops->pass(c->prec, 8, v, b, c->tp, c->g, c->stream));
    *nparts = c->ops->ntiles(c->prec, v, c->g);
  } else {
    HIPCHK(c, c->ops->density(c->prec, -1, F, nullptr, c->part, c->tp, c->g, c->stream));
    *nparts = c->nt0;
  }
  return 0;
}

// The variants the passes run in: a fused sweep or the compensated pass 1
// falls back to the first strip variant where the chosen one lacks it; the
// stored variants are set to what runs, so ss_get_tile_variants reports it.
static void settle_variants(ss_ctx* c) {
  if (c->pairs) {
    c->tv[1] = c->prec == SS_PREC_FP32C ? comp12_v(c, c->tv[1]) : pairs_v(c, c->tv[1]);
    c->tv[4] = pairs_v(c, c->tv[4]);
  }
  if (c->fuse23) c->tv[2] = fused23_v(c);
  if (c->fuse41) c->tv[4] = fused41_v(c);
  c->tv[1] = pass1_variant(c);
}

static int alloc_low_words(ss_ctx* c) {
  const size_t bytes = (size_t)c->g.nf * (size_t)c->g.N * sizeof(float);
  if (!c->Flo) HIPCHK(c, hipMalloc(&c->Flo, bytes));
  HIPCHK(c, hipMemsetAsync(c->Flo, 0, bytes, c->stream));
  return 0;
}

static int alloc_fields(ss_ctx* c) {
  const size_t bytes = (size_t)c->g.nf * (size_t)c->g.N * elem_size(c->prec);
  for (int i = 0; i < 2; ++i) {
    HIPCHK(c, hipMalloc(&c->F[i], bytes));
    HIPCHK(c, hipMalloc(&c->V[i], bytes));
    HIPCHK(c, hipMemsetAsync(c->F[i], 0, bytes, c->stream));
    HIPCHK(c, hipMemsetAsync(c->V[i], 0, bytes, c->stream));
  }
  for (int i = 0; i < 3; ++i) {
    HIPCHK(c, hipMalloc(&c->l[i], bytes));
    HIPCHK(c, hipMemsetAsync(c->l[i], 0, bytes, c->stream));
  }
  if (c->fuse41 && ensure_l1alt(c)) return -1;
  if (c->prec == SS_PREC_FP32C && alloc_low_words(c)) return -1;
  c->cur = 0;
  return 0;
}

static int alloc_partials(ss_ctx* c) {
  int nt = 0;
  for (int v = 0; v < c->ops->num_shapes; ++v) {
    const int n = c->ops->ntiles(c->prec, v, c->g);
    nt = n > nt ? n : nt;
  }
  c->nt0 = c->ops->ntiles(c->prec, 0, c->g);
  if (nt > c->ntiles || !c->part) {
    if (c->part) hipFree(c->part);
    c->part = nullptr;
    HIPCHK(c, hipMalloc(&c->part, sizeof(double) * 3 * (size_t)nt));
    c->ntiles = nt;
  }
  return 0;
}

static int ensure_hist(ss_ctx* c, int n) {
  if (n <= c->hist_cap) return 0;
  int cap = c->hist_cap ? c->hist_cap : 1024;
  while (cap < n) cap *= 2;
  if (c->hist) hipFree(c->hist);
  c->hist = nullptr;
  HIPCHK(c, hipMalloc(&c->hist, sizeof(double) * 2 * (size_t)cap));
  c->hist_cap = cap;
  free_graphs(c);  // graphs captured the old history pointer
  return 0;
}

static PassBufs bufs_for(ss_ctx* c, int parity) {
  PassBufs b{};
  b.F = c->F[parity];
  b.V = c->V[parity];
  b.l1 = l1_buf(c, parity);
  b.l2 = c->l[1];
  b.l3 = c->l[2];
  b.Fout = c->F[parity ^ 1];
  b.Vout = c->V[parity ^ 1];
  b.st = c->st;
  b.Flo = c->Flo;
  return b;
}

static int launch_step_exact(ss_ctx* c, int parity);
static int launch_step_l2(ss_ctx* c, int parity);

// Finalize of a 4+1 sweep that produced the field of parity ``parity``: err of
// the step, energy of the new field, arrest decision.
static int launch_fin41(ss_ctx* c, int parity) {
  const int nt = c->ops->ntiles(c->prec, fused41_v(c), c->g);
  hipLaunchKernelGGL(k_fin_err_energy, dim3(1), dim3(kFinThreads), 0, c->stream, (const double*)c->part,
                     (const double*)(c->part + c->ntiles), nt, (const double*)(c->part + 2 * c->ntiles), nt, c->st,
                     c->hist + c->hist_cap, c->hist, c->hist_cap, c->g.killkinen, parity);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// Launches of one step from buffer parity ``parity``.  With ``ev`` non-null an
// event is recorded before the first and after every launch (7 events: pass 1,
// energy finalize, pass 2 or 2+3, pass 3, pass 4 or 4+1, err finalize).
//
// Separate schedule: pass 1 (l1, energy of F) -> arrest decision of the
// previous step -> passes 2, 3 -> pass 4 -> err.  Fused 4+1 schedule: pass 1
// runs only in the first step of a call (``first``); every step ends with the
// 4+1 sweep, which also gives l1 and the energy of the new field, so the
// arrest decision follows it directly.
// The finalize that follows a 1+2 sweep whose energy partials are in p2 (n_en
// of them): with the err of the previous step still pending (deferred
// schedule, not the first step of a call) both in one launch, else the
// energy's alone.
static int launch_fin_after12(ss_ctx* c, bool pending, int n_err, const double* p2, int n_en, int parity) {
  if (pending)
    hipLaunchKernelGGL(k_fin_err_energy, dim3(1), dim3(kFinThreads), 0, c->stream, (const double*)c->part,
                       (const double*)(c->part + c->ntiles), n_err, p2, n_en, c->st, c->hist + c->hist_cap, c->hist,
                       c->hist_cap, c->g.killkinen, parity);
  else
    hipLaunchKernelGGL(k_fin_energy, dim3(1), dim3(kFinThreads), 0, c->stream, p2, n_en, c->st, c->hist,
                       c->hist_cap, c->g.killkinen, parity);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// The err finalize of a 3+4 sweep (partials p0, p1; n of them).
static int launch_fin_err(ss_ctx* c, int n) {
  hipLaunchKernelGGL(k_fin_err, dim3(1), dim3(kFinThreads), 0, c->stream, (const double*)c->part,
                     (const double*)(c->part + c->ntiles), n, c->st, c->hist + c->hist_cap, c->hist_cap);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// Paired schedule: passes 1+2 (l1 and the energy of F; l2 from Temp2 = F +
// dt V/2, V read before the arrest decision) -> the decision -> passes 3+4
// (after an arrest with l2 = l1, which is the l2 of V = 0) -> err.  Events as
// in launch_step, passes 2 and 3 empty.  Deferred (defer_err): the err of the
// step is finalized with the next step's energy, after its 1+2 sweep (``first``:
// no err pending, the first step of a call), and by do_steps after the last
// step of a call; the timed run finalizes every step on its own.
static int launch_step_pairs(ss_ctx* c, int parity, hipEvent_t* ev, bool first) {
  PassBufs b = bufs_for(c, parity);
  double* p0 = c->part;
  double* p1 = c->part + c->ntiles;
  double* p2 = c->part + 2 * c->ntiles;
  const bool defer = defer_err(c, ev != nullptr);
  const int v12 = pairs_v(c, c->tv[1]), v34 = pairs_v(c, c->tv[4]);
  const int n12 = c->ops->ntiles(c->prec, v12, c->g), n34 = c->ops->ntiles(c->prec, v34, c->g);
  int e = 0;
  if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
  b.lout = l1_buf(c, parity);
  b.lout2 = c->l[1];
  b.part0 = defer ? p2 : p0;
  HIPCHK(c, c->ops->pass(c->prec, 9, v12, b, c->tp, c->g, c->stream));
  if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
  if (launch_fin_after12(c, defer && !first, n34, b.part0, n12, parity)) return -1;
  if (ev)
    for (int k = 0; k < 3; ++k) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
  b.lout = nullptr;
  b.lout2 = nullptr;
  b.part0 = p0;
  b.part1 = p1;
  HIPCHK(c, c->ops->pass(c->prec, 10, v34, b, c->tp, c->g, c->stream));
  if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
  if (!defer && launch_fin_err(c, n34)) return -1;
  if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
  return 0;
}

// Compensated form, paired.  Where the three-part ring fits: the
// compensated 1+2 sweep (F, its low words and V in; l1 with the energy of
// hi + lo, and l2 from Temp2 = F + dt V/2 with V read before the arrest
// decision) -> the decision -> compensated 3+4 (after an arrest with
// l2 = l1) -> err.  Elsewhere: pass 1 (low words, energy of hi + lo) -> the
// decision -> pass 2 (V arrested) -> compensated 3+4 -> err.
static int launch_step_comp_pairs(ss_ctx* c, int parity, hipEvent_t* ev, bool first) {
  PassBufs b = bufs_for(c, parity);
  double* p0 = c->part;
  double* p1 = c->part + c->ntiles;
  double* p2 = c->part + 2 * c->ntiles;
  const int v1 = pass1_variant(c), v34 = pairs_v(c, c->tv[4]);
  int e = 0;
  if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
  if (comp12_ok(c)) {  // err finalize deferred as in launch_step_pairs
    const bool defer = defer_err(c, ev != nullptr);
    const int v12 = comp12_v(c, c->tv[1]);
    const int n12 = c->ops->ntiles(c->prec, v12, c->g), n34 = c->ops->ntiles(c->prec, v34, c->g);
    b.lout = l1_buf(c, parity);
    b.lout2 = c->l[1];
    b.part0 = defer ? p2 : p0;
    HIPCHK(c, c->ops->pass(c->prec, 12, v12, b, c->tp, c->g, c->stream));
    if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
    if (launch_fin_after12(c, defer && !first, n34, b.part0, n12, parity)) return -1;
    if (ev)
      for (int k = 0; k < 3; ++k) HIPCHK(c, hipEventRecord(ev[e++], c->stream));  // no separate passes 2, 3
    b.lout = nullptr;
    b.lout2 = nullptr;
    b.part0 = p0;
    b.part1 = p1;
    HIPCHK(c, c->ops->pass(c->prec, 11, v34, b, c->tp, c->g, c->stream));
    if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
    if (!defer && launch_fin_err(c, n34)) return -1;
    if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
    return 0;
  }
  b.lout = l1_buf(c, parity);
  b.part0 = p0;
  HIPCHK(c, c->ops->pass(c->prec, 8, v1, b, c->tp, c->g, c->stream));
  if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
  hipLaunchKernelGGL(k_fin_energy, dim3(1), dim3(kFinThreads), 0, c->stream, (const double*)p0,
                     c->ops->ntiles(c->prec, v1, c->g), c->st, c->hist, c->hist_cap, c->g.killkinen, parity);
  HIPCHK(c, hipGetLastError());
  if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
  b.lout = c->l[1];
  HIPCHK(c, c->ops->pass(c->prec, 2, c->tv[2], b, c->tp, c->g, c->stream));
  if (ev)
    for (int k = 0; k < 2; ++k) HIPCHK(c, hipEventRecord(ev[e++], c->stream));  // no separate pass 3
  b.lout = nullptr;
  b.part1 = p1;
  HIPCHK(c, c->ops->pass(c->prec, 11, v34, b, c->tp, c->g, c->stream));
  if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
  hipLaunchKernelGGL(k_fin_err, dim3(1), dim3(kFinThreads), 0, c->stream, (const double*)p0, (const double*)p1,
                     c->ops->ntiles(c->prec, v34, c->g), c->st, c->hist + c->hist_cap, c->hist_cap);
  HIPCHK(c, hipGetLastError());
  if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
  return 0;
}

static int launch_step(ss_ctx* c, int parity, hipEvent_t* ev, bool first) {
  if (c->var_mode == SS_VAR_EXACT_ADJOINT || c->ops->l2_normalize) {
    if (ev) return fail(c, "pass timing is available for the fused reference flow only");
    return c->ops->l2_normalize ? launch_step_l2(c, parity) : launch_step_exact(c, parity);
  }
  if (use_pairs(c)) return launch_step_pairs(c, parity, ev, first);
  if (use_comp_pairs(c)) return launch_step_comp_pairs(c, parity, ev, first);
  const bool f41 = use41(c);
  const bool fold = fold_mode(c, ev != nullptr);
  PassBufs b = bufs_for(c, parity);
  double* p0 = c->part;
  double* p1 = c->part + c->ntiles;
  double* p2 = c->part + 2 * c->ntiles;
  double* histE = c->hist;
  double* histErr = c->hist + c->hist_cap;
  int e = 0;
  if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
  if (!f41 || first) {
    b.lout = l1_buf(c, parity);
    b.part0 = p0;
    const int v1 = pass1_variant(c);
    HIPCHK(c, c->ops->pass(c->prec, pass1_id(c), v1, b, c->tp, c->g, c->stream));
    if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
    hipLaunchKernelGGL(k_fin_energy, dim3(1), dim3(kFinThreads), 0, c->stream, (const double*)p0,
                       c->ops->ntiles(c->prec, v1, c->g), c->st, histE, c->hist_cap, c->g.killkinen, parity);
    HIPCHK(c, hipGetLastError());
    if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
  } else if (ev) {
    HIPCHK(c, hipEventRecord(ev[e++], c->stream));
    HIPCHK(c, hipEventRecord(ev[e++], c->stream));
  }
  if (c->fuse23) {
    b.lout = c->l[1];
    b.lout2 = c->l[2];
    if (fold && !first) {  // the previous step's finalize (its partials are those of the last 4+1 sweep)
      b.fin_p0 = p0;
      b.fin_p1 = p1;
      b.fin_p2 = p2;
      b.fin_n = c->ops->ntiles(c->prec, fused41_v(c), c->g);
      b.fin_st = c->st;
      b.histE = histE;
      b.histErr = histErr;
      b.hist_cap = c->hist_cap;
      b.killkinen = c->g.killkinen;
      b.parity = parity;
    }
    HIPCHK(c, c->ops->pass(c->prec, 5, fused23_v(c), b, c->tp, c->g, c->stream));
    b.fin_n = 0;
    if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
    if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));  // no separate pass 3
  } else {
    b.lout = c->l[1];
    HIPCHK(c, c->ops->pass(c->prec, 2, c->tv[2], b, c->tp, c->g, c->stream));
    if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
    b.lout = c->l[2];
    HIPCHK(c, c->ops->pass(c->prec, 3, c->tv[3], b, c->tp, c->g, c->stream));
    if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
  }
  b.lout2 = nullptr;
  b.part0 = p0;
  b.part1 = p1;
  if (f41) {
    const int v4 = fused41_v(c);
    b.lout = l1_buf(c, parity ^ 1);
    b.part2 = p2;
    HIPCHK(c, c->ops->pass(c->prec, 6, v4, b, c->tp, c->g, c->stream));
    if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
    if (fold) return 0;  // finalized by the next step's 2+3 sweep, or at the end of the call
    if (launch_fin41(c, parity ^ 1)) return -1;
    if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
    return 0;
  }
  b.lout = nullptr;
  HIPCHK(c, c->ops->pass(c->prec, c->prec == SS_PREC_FP32C ? 7 : 4, c->tv[4], b, c->tp, c->g, c->stream));
  if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
  hipLaunchKernelGGL(k_fin_err, dim3(1), dim3(kFinThreads), 0, c->stream, (const double*)p0, (const double*)p1,
                     c->ops->ntiles(c->prec, c->tv[4], c->g), c->st, histErr, c->hist_cap);
  HIPCHK(c, hipGetLastError());
  if (ev) HIPCHK(c, hipEventRecord(ev[e++], c->stream));
  return 0;
}

static int ensure_exact(ss_ctx* c) {
  if (c->prec != SS_PREC_FP64)
    return fail(c, c->ops->l2_normalize ? "the L2-normalised flow of %s needs FP64 storage"
                                        : "the exact-adjoint variation needs FP64 storage", c->ops->name);
  const size_t nd = (size_t)c->g.nf * (size_t)c->g.N;
  if (!c->G) {
    HIPCHK(c, hipMalloc(&c->temp, nd * sizeof(double)));
    HIPCHK(c, hipMalloc(&c->G, nd * sizeof(double)));
    HIPCHK(c, hipMalloc(&c->Px, nd * sizeof(double)));
    HIPCHK(c, hipMalloc(&c->Py, nd * sizeof(double)));
  }
  const int nb = (int)((c->g.N + kThreads - 1) / kThreads);
  if (!c->xpart || c->xpart_n < nb) {
    if (c->xpart) hipFree(c->xpart);
    c->xpart = nullptr;
    HIPCHK(c, hipMalloc(&c->xpart, 2 * sizeof(double) * (size_t)nb));
    c->xpart_n = nb;
  }
  return 0;
}

// One step of the unfused exact-adjoint flow: the same RK4 stages, kill and
// error as the fused passes, with the variation replaced by
// (1/gv) dE_h/dphi (theory flow convention applied).
static int launch_step_exact(ss_ctx* c, int parity) {
  if (ensure_exact(c)) return -1;
  PassBufs b = bufs_for(c, parity);
  HIPCHK(c, c->ops->density(c->prec, -1, b.F, nullptr, c->part, c->tp, c->g, c->stream));
  hipLaunchKernelGGL(k_fin_energy, dim3(1), dim3(kFinThreads), 0, c->stream, (const double*)c->part, c->nt0, c->st,
                     c->hist, c->hist_cap, c->g.killkinen, parity);
  HIPCHK(c, hipGetLastError());
  const double ca[4] = {0.0, 0.5, 0.5, 1.0};
  const double cb[4] = {0.0, 0.0, 0.5, 0.5};
  const void* lprev[4] = {nullptr, nullptr, c->l[0], c->l[1]};
  void* lout[4] = {c->l[0], c->l[1], c->l[2], c->temp};
  const long long n = (long long)c->g.nf * c->g.N;
  const int nb = (int)((n + 255) / 256);
  for (int st = 0; st < 4; ++st) {
    hipLaunchKernelGGL((stage_field<double>), dim3(nb < 4096 ? nb : 4096), dim3(256), 0, c->stream,
                       (double*)c->temp, (const double*)b.F, (const double*)b.V, (const double*)lprev[st], ca[st],
                       cb[st], st > 0 ? 1 : 0, (const DevState*)c->st, c->g, (double*)nullptr);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, c->ops->adjoint(c->prec, c->temp, c->G, c->Px, c->Py, c->temp, lout[st], 1, c->tp, c->g, c->stream));
  }
  b.part0 = c->xpart;
  b.part1 = c->xpart + c->xpart_n;
  int nblk = 0;
  HIPCHK(c, c->ops->combine(c->prec, b, c->G, c->tp, c->g, c->stream, &nblk));
  hipLaunchKernelGGL(k_fin_err, dim3(1), dim3(kFinThreads), 0, c->stream, (const double*)b.part0, (const double*)b.part1,
                     nblk, c->st, c->hist + c->hist_cap, c->hist_cap);
  HIPCHK(c, hipGetLastError());
  return 0;
}

// Grid of the L2 normalisation's partial sums (at most 1024 blocks, and no
// more than the partial buffer holds).
static int l2_blocks(const ss_ctx* c) {
  const long long n = (long long)c->g.nf * c->g.N;
  int nb = (int)((n + 255) / 256);
  if (nb > 1024) nb = 1024;
  if (nb > c->xpart_n) nb = c->xpart_n;
  return nb;
}

// The norm from the nb partial sums in c->xpart, and G divided by it.
static int l2_finish(ss_ctx* c, double* G, int nb) {
  const long long n = (long long)c->g.nf * c->g.N;
  hipLaunchKernelGGL(k_fin_sqrt, dim3(1), dim3(1024), 0, c->stream, (const double*)c->xpart, nb, c->scal + 1);
  hipLaunchKernelGGL(k_div_scalar, dim3(nb), dim3(256), 0, c->stream, G, n, (const double*)(c->scal + 1));
  HIPCHK(c, hipGetLastError());
  return 0;
}

static int l2_normalize_buf(ss_ctx* c, double* G) {
  const long long n = (long long)c->g.nf * c->g.N;
  const int nb = l2_blocks(c);
  hipLaunchKernelGGL(k_l2_partial, dim3(nb), dim3(256), 0, c->stream, (const double*)G, n, c->g.gv, c->xpart);
  return l2_finish(c, G, nb);
}

// One step of the unfused L2-normalised reference flow (Bose-Einstein
// condensate): every stage field and the combined field are normalised as
// core/integrator.py:295-337 does.  FP64 storage only.
static int launch_step_l2(ss_ctx* c, int parity) {
  if (ensure_exact(c)) return -1;
  PassBufs b = bufs_for(c, parity);
  HIPCHK(c, c->ops->density(c->prec, -1, b.F, nullptr, c->part, c->tp, c->g, c->stream));
  hipLaunchKernelGGL(k_fin_energy, dim3(1), dim3(kFinThreads), 0, c->stream, (const double*)c->part, c->nt0, c->st,
                     c->hist, c->hist_cap, c->g.killkinen, parity);
  HIPCHK(c, hipGetLastError());
  const double ca[4] = {0.0, 0.5, 0.5, 1.0};
  const double cb[4] = {0.0, 0.0, 0.5, 0.5};
  const void* lprev[4] = {nullptr, nullptr, c->l[0], c->l[1]};
  // each stage field is formed together with its partial sums of |.|^2 gv
  // (on the normalisation's grid, so in its order), then normalised
  const int nb = l2_blocks(c);
  for (int st = 0; st < 4; ++st) {
    hipLaunchKernelGGL((stage_field<double>), dim3(nb), dim3(256), 0, c->stream, (double*)c->temp,
                       (const double*)b.F, (const double*)b.V, (const double*)lprev[st], ca[st], cb[st],
                       st > 0 ? 1 : 0, (const DevState*)c->st, c->g, c->xpart);
    HIPCHK(c, hipGetLastError());
    if (l2_finish(c, (double*)c->temp, nb)) return -1;
    PassBufs v{};
    v.F = c->temp;
    v.st = c->st;
    if (st < 3) {
      // l_{st+1} = -dt EG of the normalised stage field, written by the
      // variation pass itself (pass 1 on that field: the same EG, times -dt);
      // its energy partials land in c->part, whose energy this step's
      // finalize has already read
      v.lout = c->l[st];
      v.part0 = c->part;
      HIPCHK(c, c->ops->pass(c->prec, 1, 0, v, c->tp, c->g, c->stream));
    } else {  // stage 4: EG itself, which the combine takes
      v.lout = c->G;
      HIPCHK(c, c->ops->pass(c->prec, 0, 0, v, c->tp, c->g, c->stream));
    }
  }
  b.part0 = c->xpart;
  b.part1 = c->xpart + c->xpart_n;
  int nblk = 0;
  HIPCHK(c, c->ops->combine(c->prec, b, c->G, c->tp, c->g, c->stream, &nblk));
  hipLaunchKernelGGL(k_fin_err, dim3(1), dim3(kFinThreads), 0, c->stream, (const double*)b.part0, (const double*)b.part1,
                     nblk, c->st, c->hist + c->hist_cap, c->hist_cap);
  HIPCHK(c, hipGetLastError());
  return l2_normalize_buf(c, (double*)b.Fout);
}

// After the last step of a call: E(F_final), the last kill decision, and the
// kill applied to V in memory.  In the fused 4+1 schedule the last step has
// already given E(F_final) and the decision.
static int launch_finish(ss_ctx* c, int nsteps) {
  if (!(use41(c) && nsteps > 0)) {
    int np = 0;
    if (energy_partials(c, c->F[c->cur], &np)) return -1;
    hipLaunchKernelGGL(k_fin_energy, dim3(1), dim3(kFinThreads), 0, c->stream, (const double*)c->part, np, c->st,
                       c->hist, c->hist_cap, c->g.killkinen, c->cur);
    HIPCHK(c, hipGetLastError());
  }
  const long long n = (long long)c->g.nf * c->g.N;
  if (c->prec)
    hipLaunchKernelGGL(k_cond_zero<float>, dim3(1024), dim3(256), 0, c->stream, (float*)c->V[c->cur], n, c->st);
  else
    hipLaunchKernelGGL(k_cond_zero<double>, dim3(1024), dim3(256), 0, c->stream, (double*)c->V[c->cur], n, c->st);
  HIPCHK(c, hipGetLastError());
  return 0;
}

static int build_graph(ss_ctx* c, int parity) {
  hipGraph_t gr = nullptr;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int s = 0; s < c->graph_steps && rc == 0; ++s) rc = launch_step(c, parity ^ (s & 1), nullptr, false);
  hipError_t e = hipStreamEndCapture(c->stream, &gr);
  if (rc) {
    if (gr) hipGraphDestroy(gr);
    return rc;
  }
  if (e != hipSuccess) return hipfail(c, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&c->graph[parity], gr, nullptr, nullptr, 0);
  hipGraphDestroy(gr);
  if (e != hipSuccess) return hipfail(c, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
  // move the executable graph to the device now rather than at its first launch
  HIPCHK(c, hipGraphUpload(c->graph[parity], c->stream));
  return 0;
}

extern "C" {

int ss_version(void) { return 100; }

const char* ss_global_error(void) { return g_global_err.c_str(); }

int ss_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* ss_last_error(ss_ctx* c) { return c ? c->err.c_str() : g_global_err.c_str(); }

int ss_ctx_create(int theory_id, const int32_t* p_i, int n_i, const double* p_f, int n_f, int precision,
                  int variation_mode, ss_ctx** out) {
  if (!out) return fail(nullptr, "ss_ctx_create: out is NULL");
  *out = nullptr;
  const TheoryOps* ops = theory_ops(theory_id);
  if (!ops) return fail(nullptr, "unknown or unbuilt theory id %d", theory_id);
  if (precision != SS_PREC_FP64 && precision != SS_PREC_FP32 && precision != SS_PREC_FP32C)
    return fail(nullptr, "bad precision %d", precision);
  if (variation_mode != SS_VAR_REFERENCE && variation_mode != SS_VAR_EXACT_ADJOINT)
    return fail(nullptr, "bad variation mode %d", variation_mode);
  if (variation_mode == SS_VAR_EXACT_ADJOINT && precision != SS_PREC_FP64)
    return fail(nullptr, "the exact-adjoint variation needs FP64 storage");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return hipfail(nullptr, "no HIP device available");
  ss_ctx* c = new ss_ctx();
  c->theory = theory_id;
  c->ops = ops;
  c->prec = precision;
  c->var_mode = variation_mode;
  auto bail = [&](int rc) {
    g_global_err = c->err;
    ss_ctx_destroy(c);
    return rc;
  };
  if (hipGetDevice(&c->dev) != hipSuccess) {
    hipfail(c, "hipGetDevice failed");
    return bail(-2);
  }
  {  // set-up of the allocation-free entry point's scratch on this device
    std::string e;
    const int rc = scratch_setup(c->dev, &e);
    if (rc) {
      fail(c, "%s", e.c_str());
      return bail(rc);
    }
  }
  if (resolve_geom(c, p_i, n_i, p_f, n_f, &c->g)) return bail(-1);
  if (c->g.nf != ops->nf) {
    fail(c, "theory %s is built for %d fields (number_total_fields=%d)", ops->name, ops->nf, c->g.nf);
    return bail(-1);
  }
  if (precision == SS_PREC_FP32C && !ops->can_comp(precision, c->g)) {
    fail(c, "the compensated FP32 form is not available for theory %s at this size", ops->name);
    return bail(-1);
  }
  char msg[256] = {0};
  if (ops->make_tp(p_i, n_i, p_f, n_f, c->tp, msg, sizeof(msg))) {
    fail(c, "%s", msg);
    return bail(-1);
  }
  memcpy(c->pi, p_i, sizeof(int32_t) * n_i);
  c->ni = n_i;
  memcpy(c->pf, p_f, sizeof(double) * n_f);
  c->nfl = n_f;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    hipfail(c, "hipStreamCreate failed");
    return bail(-2);
  }
  settle_variants(c);
  int rc = alloc_fields(c);
  if (rc == 0) rc = alloc_partials(c);
  if (rc) return bail(rc);
  if (hipMalloc(&c->st, sizeof(DevState)) != hipSuccess || hipMalloc(&c->scal, 8 * sizeof(double)) != hipSuccess ||
      hipHostMalloc(&c->host_st, sizeof(DevState), hipHostMallocDefault) != hipSuccess) {
    hipfail(c, "allocation of state failed");
    return bail(-2);
  }
  if (hipMemsetAsync(c->st, 0, sizeof(DevState), c->stream) != hipSuccess) {
    hipfail(c, "hipMemsetAsync of state failed");
    return bail(-2);
  }
  rc = ensure_hist(c, 1024);
  if (rc) return bail(rc);
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    hipfail(c, "initial synchronize failed");
    return bail(-2);
  }
  *out = c;
  return 0;
}

int ss_ctx_destroy(ss_ctx* c) {
  if (!c) return 0;
  if (c->stream) hipStreamSynchronize(c->stream);
  free_graphs(c);
  free_fields(c);
  if (c->part) hipFree(c->part);
  if (c->st) hipFree(c->st);
  if (c->hist) hipFree(c->hist);
  if (c->dens) hipFree(c->dens);
  if (c->scal) hipFree(c->scal);
  if (c->host_st) hipHostFree(c->host_st);
  if (c->temp) hipFree(c->temp);
  if (c->G) hipFree(c->G);
  if (c->Px) hipFree(c->Px);
  if (c->Py) hipFree(c->Py);
  if (c->xpart) hipFree(c->xpart);
  if (c->ev_dev) hipEventDestroy(c->ev_dev);
  if (c->pgpart) hipFree(c->pgpart);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

int ss_set_params(ss_ctx* c, const int32_t* p_i, int n_i, const double* p_f, int n_f) {
  if (!c) return fail(nullptr, "null context");
  Geom g{};
  if (resolve_geom(c, p_i, n_i, p_f, n_f, &g)) return -1;
  if (g.X != c->g.X || g.Y != c->g.Y || g.nf != c->g.nf)
    return fail(c, "set_params cannot change xlen, ylen or number_total_fields");
  char msg[256] = {0};
  alignas(16) unsigned char tp[512];
  if (c->ops->make_tp(p_i, n_i, p_f, n_f, tp, msg, sizeof(msg))) return fail(c, "%s", msg);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  memcpy(c->tp, tp, sizeof(tp));
  c->g = g;
  memcpy(c->pi, p_i, sizeof(int32_t) * n_i);
  c->ni = n_i;
  memcpy(c->pf, p_f, sizeof(double) * n_f);
  c->nfl = n_f;
  free_graphs(c);
  return 0;
}

int ss_set_variation_mode(ss_ctx* c, int mode) {
  if (!c) return fail(nullptr, "null context");
  if (mode != SS_VAR_REFERENCE && mode != SS_VAR_EXACT_ADJOINT) return fail(c, "bad variation mode %d", mode);
  if (mode == SS_VAR_EXACT_ADJOINT && c->prec != SS_PREC_FP64)
    return fail(c, "the exact-adjoint variation needs FP64 storage");
  c->var_mode = mode;
  free_graphs(c);
  return 0;
}

int ss_set_precision(ss_ctx* c, int prec) {
  if (!c) return fail(nullptr, "null context");
  if (prec != SS_PREC_FP64 && prec != SS_PREC_FP32 && prec != SS_PREC_FP32C) return fail(c, "bad precision %d", prec);
  if (prec == c->prec) return 0;
  if (prec != SS_PREC_FP64 && c->var_mode == SS_VAR_EXACT_ADJOINT)
    return fail(c, "the exact-adjoint variation needs FP64 storage");
  if (prec == SS_PREC_FP32C && !c->ops->can_comp(prec, c->g))
    return fail(c, "the compensated FP32 form is not available for theory %s at this size", c->ops->name);
  const long long n = (long long)c->g.nf * c->g.N;
  if (prec != SS_PREC_FP64 && c->prec != SS_PREC_FP64) {
    // FP32 <-> compensated FP32: the same storage; the low words start at
    // zero, or are folded into the field (rounded to FP32) and dropped.
    free_graphs(c);
    if (prec == SS_PREC_FP32C) {
      c->prec = prec;
      if (alloc_low_words(c)) return -1;
    } else {
      hipLaunchKernelGGL(k_fold, dim3(2048), dim3(256), 0, c->stream, (float*)c->F[c->cur], (const float*)c->Flo, n);
      HIPCHK(c, hipGetLastError());
      HIPCHK(c, hipStreamSynchronize(c->stream));
      hipFree(c->Flo);
      c->Flo = nullptr;
      c->prec = prec;
    }
    if (c->fuse41 && !can41(c)) c->fuse41 = 0;
    settle_variants(c);
    return 0;
  }
  const size_t nb = (size_t)n * elem_size(prec);
  void* nF = nullptr;
  void* nV = nullptr;
  float* nLo = nullptr;
  HIPCHK(c, hipMalloc(&nF, nb));
  HIPCHK(c, hipMalloc(&nV, nb));
  if (prec == SS_PREC_FP32C) HIPCHK(c, hipMalloc(&nLo, nb));
  if (prec == SS_PREC_FP64) {
    if (c->prec == SS_PREC_FP32C)
      hipLaunchKernelGGL(k_join, dim3(2048), dim3(256), 0, c->stream, (const float*)c->F[c->cur], (const float*)c->Flo,
                         (double*)nF, n);
    else
      hipLaunchKernelGGL((k_convert<float, double>), dim3(2048), dim3(256), 0, c->stream, (const float*)c->F[c->cur],
                         (double*)nF, n);
    hipLaunchKernelGGL((k_convert<float, double>), dim3(2048), dim3(256), 0, c->stream, (const float*)c->V[c->cur], (double*)nV, n);
  } else {
    if (prec == SS_PREC_FP32C)
      hipLaunchKernelGGL(k_split, dim3(2048), dim3(256), 0, c->stream, (const double*)c->F[c->cur], (float*)nF, nLo, n);
    else
      hipLaunchKernelGGL((k_convert<double, float>), dim3(2048), dim3(256), 0, c->stream,
                         (const double*)c->F[c->cur], (float*)nF, n);
    hipLaunchKernelGGL((k_convert<double, float>), dim3(2048), dim3(256), 0, c->stream, (const double*)c->V[c->cur], (float*)nV, n);
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  free_graphs(c);
  free_fields(c);
  c->prec = prec;
  if (c->fuse23 && !c->ops->fuse23_variant(prec, c->tv[2], c->g) && !c->ops->can_fuse23(prec, c->g)) c->fuse23 = 0;
  if (c->fuse41 && !can41(c)) c->fuse41 = 0;
  settle_variants(c);
  if (alloc_fields(c)) return -1;
  HIPCHK(c, hipMemcpyAsync(c->F[0], nF, nb, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->V[0], nV, nb, hipMemcpyDeviceToDevice, c->stream));
  if (nLo) HIPCHK(c, hipMemcpyAsync(c->Flo, nLo, nb, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  hipFree(nF);
  hipFree(nV);
  if (nLo) hipFree(nLo);
  c->cur = 0;
  return alloc_partials(c);
}

static int host_copy(ss_ctx* c, int buf, double* host, size_t n, bool up) {
  if (!c) return fail(nullptr, "null context");
  const size_t want = (size_t)c->g.nf * (size_t)c->g.N;
  if (n != want) return fail(c, "buffer size %zu does not equal nf*xlen*ylen = %zu", n, want);
  if (!host) return fail(c, "null host pointer");
  void* dev;
  if (buf == SS_BUF_FIELD) dev = c->F[c->cur];
  else if (buf == SS_BUF_VELOCITY) dev = c->V[c->cur];
  else return fail(c, "unknown buffer %d", buf);
  if (c->prec == SS_PREC_FP64) {
    if (up) HIPCHK(c, hipMemcpyAsync(dev, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    else HIPCHK(c, hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  } else if (c->prec == SS_PREC_FP32C && buf == SS_BUF_FIELD) {
    // compensated form: the field is hi + lo
    std::vector<float> hi(n), lo(n);
    if (up) {
      for (size_t i = 0; i < n; ++i) {
        hi[i] = (float)host[i];
        lo[i] = (float)(host[i] - (double)hi[i]);
      }
      HIPCHK(c, hipMemcpyAsync(dev, hi.data(), n * sizeof(float), hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipMemcpyAsync(c->Flo, lo.data(), n * sizeof(float), hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
    } else {
      HIPCHK(c, hipMemcpyAsync(hi.data(), dev, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipMemcpyAsync(lo.data(), c->Flo, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      for (size_t i = 0; i < n; ++i) host[i] = (double)hi[i] + (double)lo[i];
    }
  } else {
    std::vector<float> tmp(n);
    if (up) {
      for (size_t i = 0; i < n; ++i) tmp[i] = (float)host[i];
      HIPCHK(c, hipMemcpyAsync(dev, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
    } else {
      HIPCHK(c, hipMemcpyAsync(tmp.data(), dev, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      for (size_t i = 0; i < n; ++i) host[i] = (double)tmp[i];
    }
  }
  return 0;
}

int ss_upload(ss_ctx* c, int buf, const double* host, size_t n) { return host_copy(c, buf, (double*)host, n, true); }
int ss_download(ss_ctx* c, int buf, double* host, size_t n) { return host_copy(c, buf, host, n, false); }

int ss_use_graphs(ss_ctx* c, int steps) {
  if (!c) return fail(nullptr, "null context");
  if (steps < 0) return fail(c, "steps_per_graph must be >= 0");
  free_graphs(c);
  c->graph_steps = steps ? ((steps + 1) & ~1) : 0;
  return 0;
}

int ss_prepare_graphs(ss_ctx* c, int nsteps) {
  if (!c) return fail(nullptr, "null context");
  if (nsteps < 0) return fail(c, "nsteps must be >= 0");
  if (ensure_hist(c, nsteps + 1)) return -1;
  if ((c->var_mode == SS_VAR_EXACT_ADJOINT || c->ops->l2_normalize) && ensure_exact(c)) return -1;
  if (c->graph_steps <= 0) return 0;
  for (int p = 0; p < 2; ++p)
    if (!c->graph[p] && build_graph(c, p)) return -1;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

static int do_steps(ss_ctx* c, double prev_energy, int nsteps, hipEvent_t* ev, double* acc_ms) {
  if (!c) return fail(nullptr, "null context");
  if (nsteps < 0) return fail(c, "nsteps must be >= 0");
  if (ensure_hist(c, nsteps + 1)) return -1;
  // the exact-adjoint and L2-normalised flows allocate their scratch on first
  // use: do it here, never inside a graph capture
  if ((c->var_mode == SS_VAR_EXACT_ADJOINT || c->ops->l2_normalize) && nsteps > 0 && ensure_exact(c)) return -1;
  hipLaunchKernelGGL(k_begin_call, dim3(1), dim3(64), 0, c->stream, c->st, prev_energy);
  HIPCHK(c, hipGetLastError());
  int done = 0;
  const bool timed = ev != nullptr;
  if (lead_step(c, timed) && nsteps > 0) {
    // the first step (4+1: carries pass 1; deferred pairs: no err pending) stays outside the graphs
    if (launch_step(c, c->cur, ev, true)) return -1;
    if (ev) {
      HIPCHK(c, hipEventSynchronize(ev[6]));
      for (int k = 0; k < 6; ++k) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
        acc_ms[k] += ms;
      }
    }
    c->cur ^= 1;
    done = 1;
  }
  if (!ev && c->graph_steps > 0 && nsteps - done >= c->graph_steps) {
    while (nsteps - done >= c->graph_steps) {
      if (!c->graph[c->cur] && build_graph(c, c->cur)) return -1;
      HIPCHK(c, hipGraphLaunch(c->graph[c->cur], c->stream));
      done += c->graph_steps;  // even: parity unchanged
    }
  }
  for (; done < nsteps; ++done) {
    if (launch_step(c, c->cur, ev, !lead_step(c, timed))) return -1;
    if (ev) {
      HIPCHK(c, hipEventSynchronize(ev[6]));
      for (int k = 0; k < 6; ++k) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
        acc_ms[k] += ms;
      }
    }
    c->cur ^= 1;
  }
  if (fold_mode(c, ev != nullptr) && nsteps > 0 && launch_fin41(c, c->cur)) return -1;
  if (defer_err(c, timed) && nsteps > 0 &&
      launch_fin_err(c, c->ops->ntiles(c->prec, pairs_v(c, c->tv[4]), c->g)))  // the last step's err
    return -1;
  if (launch_finish(c, nsteps)) return -1;
  c->last_nsteps = nsteps;
  return 0;
}

int ss_step(ss_ctx* c, double prev_energy, int nsteps, double* e_new, double* err_max, double* err_abs) {
  if (do_steps(c, prev_energy, nsteps, nullptr, nullptr)) return -1;
  HIPCHK(c, hipMemcpyAsync(c->host_st, c->st, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (e_new) *e_new = c->host_st->last_energy;
  if (err_max) *err_max = c->host_st->err_max;
  if (err_abs) *err_abs = c->host_st->err_absmax;
  return 0;
}

int ss_time_passes(ss_ctx* c, double prev_energy, int nsteps, double* avg_ms) {
  if (!c) return fail(nullptr, "null context");
  if (!avg_ms || nsteps < 1) return fail(c, "ss_time_passes needs nsteps >= 1 and an output array");
  hipEvent_t ev[7];
  for (int k = 0; k < 7; ++k) HIPCHK(c, hipEventCreate(&ev[k]));
  double acc[6] = {0, 0, 0, 0, 0, 0};
  int rc = do_steps(c, prev_energy, nsteps, ev, acc);
  if (rc == 0) {
    hipStreamSynchronize(c->stream);
    for (int k = 0; k < 6; ++k) avg_ms[k] = acc[k] / nsteps;
  }
  for (int k = 0; k < 7; ++k) hipEventDestroy(ev[k]);
  return rc;
}

int ss_set_tile_variants(ss_ctx* c, const int* variants) {
  if (!c) return fail(nullptr, "null context");
  if (!variants) return fail(c, "null variant array");
  for (int p = 0; p < 4; ++p)
    if (variants[p] < 0 || variants[p] >= c->ops->num_shapes)
      return fail(c, "shape variant of pass %d is %d (valid: 0 .. %d)", p + 1, variants[p], c->ops->num_shapes - 1);
  for (int p = 0; p < 4; ++p) c->tv[p + 1] = variants[p];
  settle_variants(c);
  free_graphs(c);  // captured graphs hold the old launch shapes
  return 0;
}

int ss_get_tile_variants(ss_ctx* c, int* variants, int* rows) {
  if (!c) return fail(nullptr, "null context");
  for (int p = 0; p < 4; ++p) {
    if (variants) variants[p] = c->tv[p + 1];
    if (rows) rows[p] = c->ops->shape_rows(c->prec, c->tv[p + 1]);
  }
  return 0;
}

// Times each fused pass in every tile-shape variant on the live buffers and
// keeps the fastest per pass.  Only scratch outputs are written (l_k, the
// partials and the next-parity F, V), so the state of the relaxation is
// unchanged.
int ss_autotune_tiles(ss_ctx* c, int reps, int* chosen, double* ms) {
  if (!c) return fail(nullptr, "null context");
  if (reps < 1) return fail(c, "reps must be >= 1");
  if (c->var_mode == SS_VAR_EXACT_ADJOINT || c->ops->l2_normalize)
    return fail(c, "tile autotuning applies to the fused reference flow only");
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  PassBufs b = bufs_for(c, c->cur);
  b.part0 = c->part;
  b.part1 = c->part + c->ntiles;
  // the compensated pass 4 updates the low words in place: time it on a copy
  float* lo_scratch = nullptr;
  if (c->prec == SS_PREC_FP32C) {
    const size_t lb = (size_t)c->g.nf * (size_t)c->g.N * sizeof(float);
    hipError_t e = hipMalloc(&lo_scratch, lb);
    if (e == hipSuccess) e = hipMemcpyAsync(lo_scratch, c->Flo, lb, hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) {
      if (lo_scratch) hipFree(lo_scratch);
      hipEventDestroy(e0);
      hipEventDestroy(e1);
      return hipfail(c, "scratch for the compensated pass 4 failed: %s", hipGetErrorString(e));
    }
    b.Flo = lo_scratch;
  }
  int pick[5] = {0, 0, 0, 0, 0};
  float bestp[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  int rc = 0;
  for (int p = 1; p <= 4 && rc == 0; ++p) {
    b.lout = p <= 3 ? c->l[p - 1] : nullptr;
    const bool comp = c->prec == SS_PREC_FP32C;
    const int pid = (comp && p == 4) ? 7 : (comp && p == 1) ? 8 : p;
    const int v0 = (comp && p == 1) ? 3 : 0;  // the compensated pass 1 streams
    float best = 0.f;
    for (int v = 0; v < c->ops->num_shapes && rc == 0; ++v) {
      if (v < v0) {
        if (ms) ms[(p - 1) * c->ops->num_shapes + v] = 0.f;
        continue;
      }
      hipError_t e = c->ops->pass(c->prec, pid, v, b, c->tp, c->g, c->stream);  // warm-up
      if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
      for (int r = 0; r < reps && e == hipSuccess; ++r) e = c->ops->pass(c->prec, pid, v, b, c->tp, c->g, c->stream);
      if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float t = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
      if (e != hipSuccess) {
        rc = hipfail(c, "tile autotuning of pass %d variant %d failed: %s", p, v, hipGetErrorString(e));
        break;
      }
      t /= reps;
      if (ms) ms[(p - 1) * c->ops->num_shapes + v] = t;
      if (v == v0 || t < best) {
        best = t;
        pick[p] = v;
      }
    }
    bestp[p] = best;
  }
  const float sep23 = bestp[2] + bestp[3];
  // passes 2 + 3 fused (streaming variants only) against the best separate pair
  float best23 = 0.f;
  int pick23 = -1;
  {
    b.lout = c->l[1];
    b.lout2 = c->l[2];
    for (int v = 0; v < c->ops->num_shapes && rc == 0; ++v) {
      if (!c->ops->fuse23_variant(c->prec, v, c->g)) {
        if (ms) ms[4 * c->ops->num_shapes + v] = 0.0;
        continue;
      }
      hipError_t e = c->ops->pass(c->prec, 5, v, b, c->tp, c->g, c->stream);
      if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
      for (int r = 0; r < reps && e == hipSuccess; ++r) e = c->ops->pass(c->prec, 5, v, b, c->tp, c->g, c->stream);
      if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float t = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
      if (e != hipSuccess) {
        rc = hipfail(c, "tile autotuning of fused passes 2+3 variant %d failed: %s", v, hipGetErrorString(e));
        break;
      }
      t /= reps;
      if (ms) ms[4 * c->ops->num_shapes + v] = t;
      if (pick23 < 0 || t < best23) {
        best23 = t;
        pick23 = v;
      }
    }
  }
  // pass 4 fused with the next step's pass 1 against the best separate pair;
  // the next step's l1 goes to the l1 buffer of the other parity (scratch)
  const float sep41 = bestp[4] + bestp[1];
  float best41 = 0.f;
  int pick41 = -1;
  bool any41 = false;
  for (int v = 0; v < c->ops->num_shapes && c->prec != SS_PREC_FP32C; ++v)
    any41 = any41 || c->ops->fuse41_variant(c->prec, v, c->g);
  if (rc == 0 && any41) rc = ensure_l1alt(c);
  if (rc == 0 && any41) {
    b.lout = (b.l1 == c->l[0]) ? c->l1alt : c->l[0];
    b.lout2 = nullptr;
    b.part2 = c->part + 2 * c->ntiles;
    for (int v = 0; v < c->ops->num_shapes && rc == 0; ++v) {
      if (!c->ops->fuse41_variant(c->prec, v, c->g)) {
        if (ms) ms[5 * c->ops->num_shapes + v] = 0.0;
        continue;
      }
      hipError_t e = c->ops->pass(c->prec, 6, v, b, c->tp, c->g, c->stream);
      if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
      for (int r = 0; r < reps && e == hipSuccess; ++r) e = c->ops->pass(c->prec, 6, v, b, c->tp, c->g, c->stream);
      if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float t = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
      if (e != hipSuccess) {
        rc = hipfail(c, "tile autotuning of fused passes 4+1 variant %d failed: %s", v, hipGetErrorString(e));
        break;
      }
      t /= reps;
      if (ms) ms[5 * c->ops->num_shapes + v] = t;
      if (pick41 < 0 || t < best41) {
        best41 = t;
        pick41 = v;
      }
    }
  }
  // the paired schedule: passes 1+2 (id 9) and 3+4 (id 10) in the strip
  // variants where both run; their outputs are the l buffers and the other
  // parity's field, which the state does not hold
  float best12 = 0.f, best34 = 0.f;
  int pick12 = -1, pick34 = -1;
  // (compensated form: its 1+2 sweep, id 12, where the three-part ring fits,
  // and its 3+4 sweep, id 11)
  const bool compf = c->prec == SS_PREC_FP32C;
  const bool any_pairs = c->ops->pairs_variant(c->prec, 3, c->g);
  for (int k = 0; k < 2 && rc == 0; ++k) {
    const int pid = k == 0 ? (compf ? 12 : 9) : compf ? 11 : 10;
    b.lout = k == 0 ? c->l[0] : nullptr;
    b.lout2 = k == 0 ? c->l[1] : nullptr;
    for (int v = 0; v < c->ops->num_shapes && rc == 0; ++v) {
      if (!any_pairs || !c->ops->pairs_variant(c->prec, v, c->g) || (compf && k == 0 && !comp12_in(c, v))) {
        if (ms) ms[(6 + k) * c->ops->num_shapes + v] = 0.0;
        continue;
      }
      hipError_t e = c->ops->pass(c->prec, pid, v, b, c->tp, c->g, c->stream);
      if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
      for (int r = 0; r < reps && e == hipSuccess; ++r) e = c->ops->pass(c->prec, pid, v, b, c->tp, c->g, c->stream);
      if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float t = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
      if (e != hipSuccess) {
        rc = hipfail(c, "tile autotuning of paired sweep %d variant %d failed: %s", pid, v, hipGetErrorString(e));
        break;
      }
      t /= reps;
      if (ms) ms[(6 + k) * c->ops->num_shapes + v] = t;
      float& best = k == 0 ? best12 : best34;
      int& pick = k == 0 ? pick12 : pick34;
      if (pick < 0 || t < best) {
        best = t;
        pick = v;
      }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (lo_scratch) {
    hipStreamSynchronize(c->stream);
    hipFree(lo_scratch);
  }
  if (rc) return rc;
  for (int p = 1; p <= 4; ++p) {
    c->tv[p] = pick[p];
    if (chosen) chosen[p - 1] = pick[p];
  }
  // a fused sweep replaces the separate passes only when at least 3 % faster
  c->pairs = 0;
  c->fuse23 = 0;
  if (pick23 >= 0 && best23 < 0.97f * sep23) {
    c->fuse23 = 1;
    c->tv[2] = pick23;
  }
  c->fuse41 = 0;
  if (pick41 >= 0 && best41 < 0.97f * sep41) {
    c->fuse41 = 1;
    c->tv[4] = pick41;
  }
  // the paired schedule replaces all of these when at least 3 % faster per step
  const float sched = (c->fuse41 ? best41 : sep41) + (c->fuse23 ? best23 : sep23);
  if (!compf && pick12 >= 0 && pick34 >= 0 && best12 + best34 < 0.97f * sched) {
    c->fuse23 = c->fuse41 = 0;
    c->pairs = 1;
    c->tv[1] = pick12;
    c->tv[4] = pick34;
    if (chosen) {
      chosen[0] = pick12;
      chosen[3] = pick34;
    }
  }
  // compensated form: pass 1 + (2+3 or 2, 3) + pass 4, against the paired
  // schedule, which is the 1+2 sweep + 3+4 where the three-part ring fits
  // (else pass 1 + pass 2 + 3+4)
  if (compf && pick34 >= 0) {
    const float sepc = bestp[1] + (c->fuse23 ? best23 : sep23) + bestp[4];
    const float pairc = pick12 >= 0 ? best12 + best34 : bestp[1] + bestp[2] + best34;
    if (pairc < 0.97f * sepc) {
      c->fuse23 = 0;
      c->pairs = 1;
      c->tv[2] = pick[2];
      c->tv[4] = pick34;
      if (pick12 >= 0) c->tv[1] = pick12;
      if (chosen) {
        if (pick12 >= 0) chosen[0] = pick12;
        chosen[1] = pick[2];
        chosen[3] = pick34;
      }
    }
  }
  settle_variants(c);
  free_graphs(c);
  return 0;
}

int ss_set_fused_stages(ss_ctx* c, int flags) {
  if (!c) return fail(nullptr, "null context");
  flags &= ~SS_FUSE_COMP12;  // reported by ss_get_fused_stages, implied by SS_FUSE_PAIRS
  if (flags & ~(SS_FUSE_23 | SS_FUSE_41 | SS_FUSE_PAIRS)) return fail(c, "unknown fused-stage flags 0x%x", flags);
  if ((flags & SS_FUSE_PAIRS) && (flags & (SS_FUSE_23 | SS_FUSE_41)))
    return fail(c, "the paired schedule replaces the 2+3 and 4+1 sweeps: set it alone");
  if ((flags & SS_FUSE_PAIRS) &&
      !(c->ops->pairs_variant(c->prec, 3, c->g) && c->var_mode == SS_VAR_REFERENCE && !c->ops->l2_normalize))
    return fail(c, "passes 1+2 and 3+4 of theory %s cannot run as paired sweeps in this form and size",
                c->ops->name);
  if ((flags & SS_FUSE_23) && !c->ops->fuse23_variant(c->prec, c->tv[2], c->g) && !c->ops->can_fuse23(c->prec, c->g))
    return fail(c, "passes 2 and 3 of theory %s cannot share one sweep at this precision and size", c->ops->name);
  if ((flags & SS_FUSE_41) && !can41(c))
    return fail(c, "pass 4 and pass 1 of theory %s cannot share one sweep at this precision and size", c->ops->name);
  if ((flags & SS_FUSE_41) && ensure_l1alt(c)) return -1;
  c->fuse23 = (flags & SS_FUSE_23) ? 1 : 0;
  c->fuse41 = (flags & SS_FUSE_41) ? 1 : 0;
  c->pairs = (flags & SS_FUSE_PAIRS) ? 1 : 0;
  settle_variants(c);
  free_graphs(c);
  return 0;
}

int ss_get_fused_stages(ss_ctx* c, int* flags) {
  if (!c) return fail(nullptr, "null context");
  if (flags)
    *flags = (c->fuse23 ? SS_FUSE_23 : 0) | (c->fuse41 ? SS_FUSE_41 : 0) | (c->pairs ? SS_FUSE_PAIRS : 0) |
             (comp12_ok(c) ? SS_FUSE_COMP12 : 0);
  return 0;
}

int ss_step_history(ss_ctx* c, double* energies, double* errs, int n) {
  if (!c) return fail(nullptr, "null context");
  if (n < 0 || n > c->last_nsteps) return fail(c, "history has %d entries (asked %d)", c->last_nsteps, n);
  if (n == 0) return 0;
  if (energies) HIPCHK(c, hipMemcpyAsync(energies, c->hist, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  if (errs) HIPCHK(c, hipMemcpyAsync(errs, c->hist + c->hist_cap, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

static int reduce_kind(ss_ctx* c, int kind, double* out) {
  if (!c) return fail(nullptr, "null context");
  if (!out) return fail(c, "null output");
  if (kind >= c->ops->ncharges) return fail(c, "theory %s has %d charges (asked %d)", c->ops->name, c->ops->ncharges, kind);
  int np = c->nt0;
  if (kind < 0) {
    if (energy_partials(c, c->F[c->cur], &np)) return -1;
  } else {
    HIPCHK(c, c->ops->density(c->prec, kind, c->F[c->cur], nullptr, c->part, c->tp, c->g, c->stream));
  }
  hipLaunchKernelGGL(k_fin_sum, dim3(1), dim3(1024), 0, c->stream, (const double*)c->part, np, c->scal);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(out, c->scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int ss_energy(ss_ctx* c, double* e) { return reduce_kind(c, -1, e); }

int ss_charge(ss_ctx* c, int kind, double* q) {
  if (c && kind < 0) return fail(c, "charge kind must be >= 0");
  return reduce_kind(c, kind, q);
}

int ss_num_charges(ss_ctx* c) { return c ? c->ops->ncharges : 0; }

int ss_density(ss_ctx* c, int kind, double* host_out, size_t n) {
  if (!c) return fail(nullptr, "null context");
  if (n != (size_t)c->g.N) return fail(c, "density needs xlen*ylen = %lld values (got %zu)", c->g.N, n);
  if (kind >= c->ops->ncharges || kind < -1) return fail(c, "bad density kind %d", kind);
  if (!c->dens) HIPCHK(c, hipMalloc(&c->dens, sizeof(double) * (size_t)c->g.N));
  HIPCHK(c, c->ops->density(c->prec, kind, c->F[c->cur], c->dens, c->part, c->tp, c->g, c->stream));
  HIPCHK(c, hipMemcpyAsync(host_out, c->dens, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int ss_init_field(ss_ctx* c, const ss_init_desc* d) {
  if (!c) return fail(nullptr, "null context");
  if (!d) return fail(c, "null init description");
  if (d->mode < 0 || d->mode > 3 || !((c->ops->init_modes() >> d->mode) & 1u))
    return fail(c, "init mode %d is not available on the device for theory %s", d->mode, c->ops->name);
  if (d->ansatz < 0 || d->ansatz > 2) return fail(c, "bad ansatz %d (0 Bloch, 1 Neel, 2 anti)", d->ansatz);
  if (d->nsol < 0 || (d->nsol > 0 && !d->solitons)) return fail(c, "bad soliton list");
  InitArgs a{};
  a.mode = d->mode;
  a.ansatz = d->ansatz;
  a.skN = d->skyrmion_number;
  a.rot = d->rotation;
  a.scale = d->scale;
  a.vortex_number = d->vortex_number;
  for (int k = 0; k < 4; ++k) a.p[k] = d->extra[k];
  a.nsol = d->nsol;
  for (int k = 0; k < c->nfl && k < kMaxPF; ++k) a.pf[k] = c->pf[k];
  double* sol = nullptr;
  if (d->nsol > 0) {
    HIPCHK(c, hipMalloc(&sol, sizeof(double) * 3 * (size_t)d->nsol));
    const hipError_t e = hipMemcpyAsync(sol, d->solitons, sizeof(double) * 3 * (size_t)d->nsol,
                                        hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) {
      hipFree(sol);
      return hipfail(c, "copy of the soliton list failed: %s", hipGetErrorString(e));
    }
  }
  a.sol = sol;
  hipError_t e = c->ops->init(c->prec, a, c->F[c->cur], c->g, c->stream);
  if (e == hipSuccess)
    e = hipMemsetAsync(c->V[c->cur], 0, elem_size(c->prec) * (size_t)c->g.nf * (size_t)c->g.N, c->stream);
  if (e == hipSuccess && c->Flo)
    e = hipMemsetAsync(c->Flo, 0, sizeof(float) * (size_t)c->g.nf * (size_t)c->g.N, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (sol) hipFree(sol);
  if (e != hipSuccess) return hipfail(c, "device initialisation failed: %s", hipGetErrorString(e));
  return 0;
}

int ss_add_noise(ss_ctx* c, int comp_begin, int comp_end, int gaussian, double amplitude, uint64_t seed,
                 int renorm_ncomp, double renorm_norm) {
  if (!c) return fail(nullptr, "null context");
  if (comp_begin < 0 || comp_end > c->g.nf || comp_begin > comp_end) return fail(c, "bad component range");
  if (renorm_ncomp < 0 || renorm_ncomp > c->g.nf) return fail(c, "bad renormalised component count");
  const long long nb = (c->g.N + 255) / 256;
  const dim3 grid((unsigned)(nb < 8192 ? nb : 8192));
  if (c->prec)
    hipLaunchKernelGGL((k_noise<float>), grid, dim3(256), 0, c->stream, (float*)c->F[c->cur], c->g.N, comp_begin,
                       comp_end, gaussian, amplitude, (unsigned long long)seed, renorm_ncomp, renorm_norm);
  else
    hipLaunchKernelGGL((k_noise<double>), grid, dim3(256), 0, c->stream, (double*)c->F[c->cur], c->g.N, comp_begin,
                       comp_end, gaussian, amplitude, (unsigned long long)seed, renorm_ncomp, renorm_norm);
  HIPCHK(c, hipGetLastError());
  if (c->Flo) HIPCHK(c, hipMemsetAsync(c->Flo, 0, sizeof(float) * (size_t)c->g.nf * (size_t)c->g.N, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int ss_aux_components(ss_ctx* c, int kind) {
  if (!c || kind < 0 || kind >= kNumAuxKinds) return 0;
  return c->ops->aux_components(kind);
}

int ss_aux_density(ss_ctx* c, int kind, int comp, double* host_out, size_t n) {
  if (!c) return fail(nullptr, "null context");
  if (n != (size_t)c->g.N) return fail(c, "aux density needs xlen*ylen = %lld values (got %zu)", c->g.N, n);
  if (!host_out) return fail(c, "null host pointer");
  if (kind < 0 || kind >= kNumAuxKinds) return fail(c, "bad aux kind %d", kind);
  const int nc = c->ops->aux_components(kind);
  if (comp < 0 || comp >= nc)
    return fail(c, "theory %s has %d components of aux kind %d (asked %d)", c->ops->name, nc, kind, comp);
  if (!c->dens) HIPCHK(c, hipMalloc(&c->dens, sizeof(double) * (size_t)c->g.N));
  HIPCHK(c, c->ops->density(c->prec, kAuxBase + 8 * kind + comp, c->F[c->cur], c->dens, c->part, c->tp, c->g,
                            c->stream));
  HIPCHK(c, hipMemcpyAsync(host_out, c->dens, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

static int variation_into(ss_ctx* c, const void* field, void* out, int mode, hipStream_t st) {
  if (mode == SS_VAR_EXACT_ADJOINT) {
    if (ensure_exact(c)) return -1;
    HIPCHK(c, c->ops->adjoint(c->prec, field, (double*)out, c->Px, c->Py, nullptr, nullptr, 0, c->tp, c->g, st));
    return 0;
  }
  if (mode != SS_VAR_REFERENCE) return fail(c, "bad variation mode %d", mode);
  PassBufs b{};
  b.F = field;
  b.lout = out;
  b.st = c->st;
  HIPCHK(c, c->ops->pass(c->prec, 0, 0, b, c->tp, c->g, st));
  return 0;
}

int ss_variation(ss_ctx* c, int mode, double* host_eg, size_t n) {
  if (!c) return fail(nullptr, "null context");
  const size_t want = (size_t)c->g.nf * (size_t)c->g.N;
  if (n != want) return fail(c, "variation needs nf*xlen*ylen = %zu values (got %zu)", want, n);
  void* scratch = c->l[2];
  if (variation_into(c, c->F[c->cur], scratch, mode, c->stream)) return -1;
  if (c->prec == SS_PREC_FP64) {
    HIPCHK(c, hipMemcpyAsync(host_eg, scratch, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> tmp(n);
    HIPCHK(c, hipMemcpyAsync(tmp.data(), scratch, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; ++i) host_eg[i] = tmp[i];
  }
  return 0;
}

// The device-pointer calls run on the caller's stream but use the context's
// scratch (partials, exact-adjoint buffers): the caller's stream first waits
// for the context's stream, and the context's stream then waits for the call,
// so they are ordered against ss_step and against each other even when issued
// from different streams.
static int dev_call_begin(ss_ctx* c, hipStream_t st) {
  if (c->prec != SS_PREC_FP64)
    return fail(c, "device-pointer calls take FP64 fields: the context is in precision %d", c->prec);
  if (!c->ev_dev) HIPCHK(c, hipEventCreateWithFlags(&c->ev_dev, hipEventDisableTiming));
  if (st != c->stream) {
    HIPCHK(c, hipEventRecord(c->ev_dev, c->stream));
    HIPCHK(c, hipStreamWaitEvent(st, c->ev_dev, 0));
  }
  return 0;
}
static int dev_call_end(ss_ctx* c, hipStream_t st) {
  if (st != c->stream) {
    HIPCHK(c, hipEventRecord(c->ev_dev, st));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_dev, 0));
  }
  return 0;
}

int ss_variation_dev(ss_ctx* c, const void* field_dev, void* grad_dev, int mode, void* stream) {
  if (!c) return fail(nullptr, "null context");
  if (!field_dev || !grad_dev) return fail(c, "null device pointer");
  hipStream_t st = (hipStream_t)stream;  // NULL is the null stream, as for any HIP call
  if (mode == SS_VAR_EXACT_ADJOINT && ensure_exact(c)) return -1;  // allocates before any ordering
  if (dev_call_begin(c, st)) return -1;
  if (variation_into(c, field_dev, grad_dev, mode, st)) return -1;
  return dev_call_end(c, st);
}

int ss_energy_dev(ss_ctx* c, const void* field_dev, double* energy_dev, void* stream) {
  if (!c) return fail(nullptr, "null context");
  if (!field_dev || !energy_dev) return fail(c, "null device pointer");
  hipStream_t st = (hipStream_t)stream;  // NULL is the null stream, as for any HIP call
  if (dev_call_begin(c, st)) return -1;
  HIPCHK(c, c->ops->density(c->prec, -1, field_dev, nullptr, c->part, c->tp, c->g, st));
  hipLaunchKernelGGL(k_fin_sum, dim3(1), dim3(1024), 0, st, (const double*)c->part, c->nt0, energy_dev);
  HIPCHK(c, hipGetLastError());
  return dev_call_end(c, st);
}

__global__ void k_fill_zero(double* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.0;
}

int ss_energy_param_grad(ss_ctx* c, const void* field_dev, double* grad_dev, int n, void* stream) {
  if (!c) return fail(nullptr, "null context");
  if (!field_dev || !grad_dev) return fail(c, "null device pointer");
  if (n != c->nfl) return fail(c, "ss_energy_param_grad: n must be the packed p_f length %d (got %d)", c->nfl, n);
  const int np = c->ops->param_grad_count();
  if (np <= kCorePF) return fail(c, "parameter derivatives of the energy are not available for theory %s", c->ops->name);
  if (np > c->nfl) return fail(c, "theory %s packs %d p_f entries, the context holds %d", c->ops->name, np, c->nfl);
  const size_t need = (size_t)(np - kCorePF) * (size_t)c->ops->param_grad_tiles(c->g);
  if (c->pgpart_n < need) {
    if (c->pgpart) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      hipFree(c->pgpart);
    }
    c->pgpart = nullptr;
    c->pgpart_n = 0;
    HIPCHK(c, hipMalloc(&c->pgpart, need * sizeof(double)));
    c->pgpart_n = need;
  }
  hipStream_t st = (hipStream_t)stream;
  if (dev_call_begin(c, st)) return -1;
  hipLaunchKernelGGL(k_fill_zero, dim3(1), dim3(64), 0, st, grad_dev, n);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, c->ops->param_grad((const double*)field_dev, c->pgpart, grad_dev + kCorePF, c->pf, kCorePF,
                               np - kCorePF, c->g, st));
  return dev_call_end(c, st);
}

int ss_energy_dev_noalloc(ss_ctx* c, const void* field_dev, double* energy_dev, void* stream) {
  if (!c) return fail(nullptr, "null context");
  if (!field_dev || !energy_dev) return fail(c, "null device pointer");
  if (c->prec != SS_PREC_FP64)
    return fail(c, "device-pointer calls take FP64 fields: the context is in precision %d", c->prec);
  if (c->nt0 > kScratchDoubles)
    return fail(c, "ss_energy_dev_noalloc: the lattice needs %d energy partials, the static scratch holds %d "
                   "(%d x %d sites is too large for it; use ss_energy_dev)",
                c->nt0, kScratchDoubles, c->g.X, c->g.Y);
  double* part = (c->dev >= 0 && c->dev < kMaxDevices) ? g_scratch[c->dev] : nullptr;
  if (!part) {  // resolved by ss_ctx_create; kept as a check
    std::string e;
    const int rc = scratch_setup(c->dev, &e);
    if (rc) return fail(c, "%s", e.c_str());
    part = g_scratch[c->dev];
  }
  hipStream_t st = (hipStream_t)stream;  // NULL is the null stream, as for any HIP call
  HIPCHK(c, c->ops->density(c->prec, -1, field_dev, nullptr, part, c->tp, c->g, st));
  hipLaunchKernelGGL(k_fin_sum, dim3(1), dim3(1024), 0, st, (const double*)part, c->nt0, energy_dev);
  HIPCHK(c, hipGetLastError());
  return 0;
}

int ss_synchronize(ss_ctx* c) {
  if (!c) return fail(nullptr, "null context");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int ss_memory_bytes(ss_ctx* c, size_t* bytes) {
  if (!c || !bytes) return fail(c, "null argument");
  const size_t f = (size_t)c->g.nf * (size_t)c->g.N * elem_size(c->prec);
  *bytes = (c->l1alt ? 8 : 7) * f + (c->Flo ? f : 0) + sizeof(double) * 3 * (size_t)c->ntiles +
           sizeof(double) * 2 * (size_t)c->hist_cap + (c->dens ? sizeof(double) * (size_t)c->g.N : 0);
  return 0;
}

}  // extern "C"

// ===========================================================================
// Batched form: ss_batch holds many independent lattices of one theory.

struct ss_batch {
  int theory = 0;
  const TheoryOps* ops = nullptr;
  int prec = 0;
  int ninst = 0;
  int ni = 0, nfl = 0;
  hipStream_t stream = nullptr;
  std::vector<Geom> geoms;
  std::vector<long long> offs;
  std::vector<int> tile_begin, ntiles_i;
  std::vector<unsigned char> tps;  // theory parameters of every instance, 512 bytes each
  long long total = 0;  // elements per field buffer
  int ntiles = 0;
  int ntcap = 0;        // capacity of the tile table and of each partials row (the most tiles over the variants)
  int tv = 0;           // shape variant (0..2 LDS tiles, 3..8 streaming strips, as for a single lattice)
  int fuse23 = 0;       // passes 2 and 3 in one sweep where the variant has it
  int pairs = 0;        // the step as the paired sweeps 1+2 and 3+4 where the variant has them
  int sx = 0;           // streaming variants: rows per strip
  void* F[2] = {nullptr, nullptr};
  void* V[2] = {nullptr, nullptr};
  void* l[3] = {nullptr, nullptr, nullptr};
  double* part = nullptr;
  int* tile_inst = nullptr;
  void* insts = nullptr;
  BatchState* st = nullptr;
  BatchState* host_st = nullptr;
  int graph_steps = 0;
  double graph_tol = -1.0;
  hipGraphExec_t graph = nullptr;
  std::string err;
};

static int bfail(ss_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (b) b->err = buf; else g_global_err = buf;
  return code;
}

#define BHIP(b, expr)                                                                              \
  do {                                                                                             \
    hipError_t _e = (expr);                                                                        \
    if (_e != hipSuccess) return bfail((b), -2, "%s failed: %s", #expr, hipGetErrorString(_e));   \
  } while (0)

// Tile table and instance descriptors of the batch for tile variant ``v``
// (rows per tile: the variant's), uploaded; the graph is dropped.
static int batch_tiles(ss_batch* b, int v) {
  const int isz = b->ops->batch_inst_size();
  std::vector<unsigned char> host_insts((size_t)isz * b->ninst);
  std::vector<int> tinst;
  int sx = 0, tyw = TY;
  b->ops->batch_shape(b->prec, v, &sx, &tyw);
  const int tx = sx ? sx : b->ops->shape_rows(b->prec, v);  // rows per strip or per tile
  b->tile_begin.clear();
  b->ntiles_i.clear();
  b->ntiles = 0;
  for (int i = 0; i < b->ninst; ++i) {
    const Geom& g = b->geoms[i];
    const int nty = (g.Y + tyw - 1) / tyw;
    const int nt = ((g.X + tx - 1) / tx) * nty;
    b->tile_begin.push_back(b->ntiles);
    b->ntiles_i.push_back(nt);
    b->ops->batch_fill(host_insts.data() + (size_t)isz * i, g, b->tps.data() + (size_t)512 * i, b->offs[i],
                       b->ntiles, nt, nty);
    for (int k = 0; k < nt; ++k) tinst.push_back(i);
    b->ntiles += nt;
  }
  if (b->ntiles > b->ntcap) return bfail(b, -1, "tile table of variant %d exceeds its capacity", v);
  if (b->graph) hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  BHIP(b, hipMemcpyAsync(b->tile_inst, tinst.data(), sizeof(int) * tinst.size(), hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipMemcpyAsync(b->insts, host_insts.data(), host_insts.size(), hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  b->tv = v;
  b->sx = sx;
  return 0;
}

static BatchBufs batch_bufs(ss_batch* b) {
  BatchBufs B{};
  for (int k = 0; k < 2; ++k) {
    B.F[k] = b->F[k];
    B.V[k] = b->V[k];
  }
  for (int k = 0; k < 3; ++k) B.l[k] = b->l[k];
  B.part0 = b->part;
  B.part1 = b->part + b->ntcap;
  B.tile_inst = b->tile_inst;
  B.insts = b->insts;
  B.st = b->st;
  B.ntiles = b->ntiles;
  B.per_xcd = (b->ntiles + 7) / 8;
  B.sx = b->sx;
  return B;
}

static bool batch_use23(const ss_batch* b) { return b->fuse23 && b->ops->batch_fuse23(b->prec, b->tv); }
static bool batch_use_pairs(const ss_batch* b) { return b->pairs && b->ops->batch_pairs(b->prec, b->tv); }

// Paired batch step: passes 1+2 of every active instance (V before its arrest
// decision) -> the decisions -> passes 3+4 (l2 = l1 where arrested) -> err.
static int batch_launch_step(ss_batch* b, double tol) {
  const BatchBufs B = batch_bufs(b);
  if (batch_use_pairs(b)) {
    BHIP(b, b->ops->batch_pass(b->prec, 7, b->tv, B, b->stream));
    BHIP(b, b->ops->batch_fin(0, b->ninst, B, 0, 0.0, b->stream));
    BHIP(b, b->ops->batch_pass(b->prec, 8, b->tv, B, b->stream));
    BHIP(b, b->ops->batch_fin(1, b->ninst, B, 0, tol, b->stream));
    return 0;
  }
  BHIP(b, b->ops->batch_pass(b->prec, 1, b->tv, B, b->stream));
  BHIP(b, b->ops->batch_fin(0, b->ninst, B, 0, 0.0, b->stream));
  if (batch_use23(b)) {
    BHIP(b, b->ops->batch_pass(b->prec, 6, b->tv, B, b->stream));
  } else {
    BHIP(b, b->ops->batch_pass(b->prec, 2, b->tv, B, b->stream));
    BHIP(b, b->ops->batch_pass(b->prec, 3, b->tv, B, b->stream));
  }
  BHIP(b, b->ops->batch_pass(b->prec, 4, b->tv, B, b->stream));
  BHIP(b, b->ops->batch_fin(1, b->ninst, B, 0, tol, b->stream));
  return 0;
}

static int batch_read_state(ss_batch* b) {
  BHIP(b, hipMemcpyAsync(b->host_st, b->st, sizeof(BatchState) * b->ninst, hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return 0;
}

extern "C" {

int ss_batch_create(int theory_id, int ninst, const int32_t* p_i_all, int n_i, const double* p_f_all, int n_f,
                    int precision, ss_batch** out) {
  if (!out) return bfail(nullptr, -1, "ss_batch_create: out is NULL");
  *out = nullptr;
  const TheoryOps* ops = theory_ops(theory_id);
  if (!ops) return bfail(nullptr, -1, "unknown or unbuilt theory id %d", theory_id);
  if (ninst < 1) return bfail(nullptr, -1, "ninst must be >= 1");
  if (ops->l2_normalize)
    return bfail(nullptr, -1, "theory %s L2-normalises its stage fields; the batched form does not", ops->name);
  if (precision != SS_PREC_FP64 && precision != SS_PREC_FP32) return bfail(nullptr, -1, "bad precision %d", precision);
  ss_batch* b = new ss_batch();
  b->theory = theory_id;
  b->ops = ops;
  b->prec = precision;
  b->ninst = ninst;
  b->ni = n_i;
  b->nfl = n_f;
  auto bail = [&](int rc) {
    g_global_err = b->err;
    ss_batch_destroy(b);
    return rc;
  };
  b->tps.assign((size_t)512 * ninst, 0);
  for (int i = 0; i < ninst; ++i) {
    ss_ctx tmp;
    Geom g{};
    const int32_t* pi = p_i_all + (size_t)i * n_i;
    const double* pf = p_f_all + (size_t)i * n_f;
    if (resolve_geom(&tmp, pi, n_i, pf, n_f, &g)) {
      bfail(b, -1, "instance %d: %s", i, tmp.err.c_str());
      return bail(-1);
    }
    if (g.nf != ops->nf) {
      bfail(b, -1, "instance %d: theory %s is built for %d fields", i, ops->name, ops->nf);
      return bail(-1);
    }
    // the batched strip kernels address an instance's arrays with 32-bit byte
    // offsets below 2^31 from its own base, as the single-lattice strips do
    // (FP64, the widest the batch may convert to)
    if (!((long long)g.nf * g.N * (long long)sizeof(double) < (1LL << 31))) {
      bfail(b, -1, "instance %d: %d x %d x %d values exceed the batched form's limit (8*nf*xlen*ylen < 2^31)", i,
            g.nf, g.X, g.Y);
      return bail(-1);
    }
    char msg[256] = {0};
    if (ops->make_tp(pi, n_i, pf, n_f, b->tps.data() + (size_t)512 * i, msg, sizeof(msg))) {
      bfail(b, -1, "instance %d: %s", i, msg);
      return bail(-1);
    }
    b->geoms.push_back(g);
    b->offs.push_back(b->total);
    b->total += (long long)g.nf * g.N;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    bfail(b, -2, "no HIP device available");
    return bail(-2);
  }
  // tile-table capacity over every variant in both storage precisions, so
  // that ss_batch_set_precision needs no new table
  for (int pr = SS_PREC_FP64; pr <= SS_PREC_FP32; ++pr)
    for (int v = 0; v < ops->num_shapes; ++v) {
      int sx = 0, tyw = TY;
      ops->batch_shape(pr, v, &sx, &tyw);
      const int tx = sx ? sx : ops->shape_rows(pr, v);
      int nt = 0;
      for (const Geom& g : b->geoms) nt += ((g.X + tx - 1) / tx) * ((g.Y + tyw - 1) / tyw);
      b->ntcap = nt > b->ntcap ? nt : b->ntcap;
    }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    bfail(b, -2, "hipStreamCreate failed");
    return bail(-2);
  }
  const size_t bytes = (size_t)b->total * elem_size(precision);
  bool ok = true;
  for (int k = 0; k < 2 && ok; ++k)
    ok = hipMalloc(&b->F[k], bytes) == hipSuccess && hipMalloc(&b->V[k], bytes) == hipSuccess;
  for (int k = 0; k < 3 && ok; ++k) ok = hipMalloc(&b->l[k], bytes) == hipSuccess;
  ok = ok && hipMalloc(&b->part, sizeof(double) * 2 * (size_t)b->ntcap) == hipSuccess;
  ok = ok && hipMalloc(&b->tile_inst, sizeof(int) * (size_t)b->ntcap) == hipSuccess;
  ok = ok && hipMalloc(&b->insts, (size_t)ops->batch_inst_size() * ninst) == hipSuccess;
  ok = ok && hipMalloc(&b->st, sizeof(BatchState) * ninst) == hipSuccess;
  ok = ok && hipHostMalloc(&b->host_st, sizeof(BatchState) * ninst, hipHostMallocDefault) == hipSuccess;
  if (!ok) {
    bfail(b, -2, "device allocation of the batch failed (%zu bytes per field buffer)", bytes);
    return bail(-2);
  }
  for (int k = 0; k < 2 && ok; ++k)
    ok = hipMemsetAsync(b->F[k], 0, bytes, b->stream) == hipSuccess &&
         hipMemsetAsync(b->V[k], 0, bytes, b->stream) == hipSuccess;
  for (int k = 0; k < 3 && ok; ++k) ok = hipMemsetAsync(b->l[k], 0, bytes, b->stream) == hipSuccess;
  ok = ok && hipMemsetAsync(b->st, 0, sizeof(BatchState) * ninst, b->stream) == hipSuccess;
  ok = ok && hipStreamSynchronize(b->stream) == hipSuccess;
  if (!ok) {
    bfail(b, -2, "initialisation of the batch failed");
    return bail(-2);
  }
  if (batch_tiles(b, 0)) return bail(-2);
  *out = b;
  return 0;
}

int ss_batch_destroy(ss_batch* b) {
  if (!b) return 0;
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->graph) hipGraphExecDestroy(b->graph);
  for (int k = 0; k < 2; ++k) {
    if (b->F[k]) hipFree(b->F[k]);
    if (b->V[k]) hipFree(b->V[k]);
  }
  for (int k = 0; k < 3; ++k)
    if (b->l[k]) hipFree(b->l[k]);
  if (b->part) hipFree(b->part);
  if (b->tile_inst) hipFree(b->tile_inst);
  if (b->insts) hipFree(b->insts);
  if (b->st) hipFree(b->st);
  if (b->host_st) hipHostFree(b->host_st);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

const char* ss_batch_last_error(ss_batch* b) { return b ? b->err.c_str() : g_global_err.c_str(); }

int ss_batch_num_tiles(ss_batch* b) { return b ? b->ntiles : 0; }

static int batch_copy(ss_batch* b, int inst, int buf, double* host, size_t n, bool up) {
  if (!b) return bfail(nullptr, -1, "null batch");
  if (inst < 0 || inst >= b->ninst) return bfail(b, -1, "instance %d out of range [0, %d)", inst, b->ninst);
  const Geom& g = b->geoms[inst];
  const size_t want = (size_t)g.nf * (size_t)g.N;
  if (n != want) return bfail(b, -1, "instance %d needs %zu values (got %zu)", inst, want, n);
  if (!host) return bfail(b, -1, "null host pointer");
  if (batch_read_state(b)) return -2;
  const int cur = b->host_st[inst].cur;
  void* base;
  if (buf == SS_BUF_FIELD) base = b->F[cur];
  else if (buf == SS_BUF_VELOCITY) base = b->V[cur];
  else return bfail(b, -1, "unknown buffer %d", buf);
  const size_t es = elem_size(b->prec);
  char* dev = (char*)base + (size_t)b->offs[inst] * es;
  if (b->prec == SS_PREC_FP64) {
    if (up) BHIP(b, hipMemcpyAsync(dev, host, n * es, hipMemcpyHostToDevice, b->stream));
    else BHIP(b, hipMemcpyAsync(host, dev, n * es, hipMemcpyDeviceToHost, b->stream));
    BHIP(b, hipStreamSynchronize(b->stream));
  } else {
    std::vector<float> tmp(n);
    if (up) {
      for (size_t k = 0; k < n; ++k) tmp[k] = (float)host[k];
      BHIP(b, hipMemcpyAsync(dev, tmp.data(), n * es, hipMemcpyHostToDevice, b->stream));
      BHIP(b, hipStreamSynchronize(b->stream));
    } else {
      BHIP(b, hipMemcpyAsync(tmp.data(), dev, n * es, hipMemcpyDeviceToHost, b->stream));
      BHIP(b, hipStreamSynchronize(b->stream));
      for (size_t k = 0; k < n; ++k) host[k] = tmp[k];
    }
  }
  return 0;
}

int ss_batch_upload(ss_batch* b, int inst, int buf, const double* host, size_t n) {
  return batch_copy(b, inst, buf, (double*)host, n, true);
}
int ss_batch_download(ss_batch* b, int inst, int buf, double* host, size_t n) {
  return batch_copy(b, inst, buf, host, n, false);
}

int ss_batch_set_precision(ss_batch* b, int precision) {
  if (!b) return bfail(nullptr, -1, "null batch");
  if (precision != SS_PREC_FP64 && precision != SS_PREC_FP32)
    return bfail(b, -1, "the batched form stores FP64 or FP32 fields (got precision %d)", precision);
  if (precision == b->prec) return 0;
  BHIP(b, hipStreamSynchronize(b->stream));
  const size_t n = (size_t)b->total;
  const size_t bytes = n * elem_size(precision);
  void* nf[2] = {nullptr, nullptr};
  void* nv[2] = {nullptr, nullptr};
  void* nl[3] = {nullptr, nullptr, nullptr};
  bool ok = true;
  for (int k = 0; k < 2 && ok; ++k) ok = hipMalloc(&nf[k], bytes) == hipSuccess && hipMalloc(&nv[k], bytes) == hipSuccess;
  for (int k = 0; k < 3 && ok; ++k) ok = hipMalloc(&nl[k], bytes) == hipSuccess;
  if (!ok) {
    for (int k = 0; k < 2; ++k) {
      if (nf[k]) hipFree(nf[k]);
      if (nv[k]) hipFree(nv[k]);
    }
    for (int k = 0; k < 3; ++k)
      if (nl[k]) hipFree(nl[k]);
    return bfail(b, -2, "device allocation for the precision switch failed (%zu bytes per buffer)", bytes);
  }
  const dim3 grid(4096), blk(256);
  for (int k = 0; k < 2; ++k) {  // both parities: every instance keeps its own buffer parity
    if (precision == SS_PREC_FP64) {
      hipLaunchKernelGGL((k_convert<float, double>), grid, blk, 0, b->stream, (const float*)b->F[k], (double*)nf[k],
                         (long long)n);
      hipLaunchKernelGGL((k_convert<float, double>), grid, blk, 0, b->stream, (const float*)b->V[k], (double*)nv[k],
                         (long long)n);
    } else {
      hipLaunchKernelGGL((k_convert<double, float>), grid, blk, 0, b->stream, (const double*)b->F[k], (float*)nf[k],
                         (long long)n);
      hipLaunchKernelGGL((k_convert<double, float>), grid, blk, 0, b->stream, (const double*)b->V[k], (float*)nv[k],
                         (long long)n);
    }
  }
  BHIP(b, hipGetLastError());
  BHIP(b, hipStreamSynchronize(b->stream));
  for (int k = 0; k < 2; ++k) {
    hipFree(b->F[k]);
    hipFree(b->V[k]);
    b->F[k] = nf[k];
    b->V[k] = nv[k];
  }
  for (int k = 0; k < 3; ++k) {  // the stage scratch is recomputed by every step
    hipFree(b->l[k]);
    b->l[k] = nl[k];
  }
  b->prec = precision;
  return batch_tiles(b, b->tv);  // rows per tile depend on the precision; drops the graph
}

int ss_batch_begin(ss_batch* b) {
  if (!b) return bfail(nullptr, -1, "null batch");
  const BatchBufs B = batch_bufs(b);
  BHIP(b, b->ops->batch_pass(b->prec, 5, b->tv, B, b->stream));
  BHIP(b, b->ops->batch_fin(0, b->ninst, B, 2, 0.0, b->stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return 0;
}

int ss_batch_use_graphs(ss_batch* b, int steps) {
  if (!b) return bfail(nullptr, -1, "null batch");
  if (steps < 0) return bfail(b, -1, "steps_per_graph must be >= 0");
  if (b->graph) hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  b->graph_steps = steps;
  return 0;
}

int ss_batch_step(ss_batch* b, int nsteps, double tol, int check_every, int* n_active) {
  if (!b) return bfail(nullptr, -1, "null batch");
  if (nsteps < 0) return bfail(b, -1, "nsteps must be >= 0");
  if (check_every < 1) check_every = nsteps > 0 ? nsteps : 1;
  if (b->graph_steps > 0 && (!b->graph || b->graph_tol != tol)) {
    if (b->graph) hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    hipGraph_t gr = nullptr;
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int s = 0; s < b->graph_steps && rc == 0; ++s) rc = batch_launch_step(b, tol);
    hipError_t e = hipStreamEndCapture(b->stream, &gr);
    if (rc) {
      if (gr) hipGraphDestroy(gr);
      return rc;
    }
    if (e != hipSuccess) return bfail(b, -2, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&b->graph, gr, nullptr, nullptr, 0);
    hipGraphDestroy(gr);
    if (e != hipSuccess) return bfail(b, -2, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    b->graph_tol = tol;
  }
  int done = 0, since = 0, active = b->ninst;
  while (done < nsteps) {
    int did;
    if (b->graph && nsteps - done >= b->graph_steps) {
      BHIP(b, hipGraphLaunch(b->graph, b->stream));
      did = b->graph_steps;
    } else {
      if (batch_launch_step(b, tol)) return -2;
      did = 1;
    }
    done += did;
    since += did;
    if (since >= check_every || done >= nsteps) {
      since = 0;
      if (batch_read_state(b)) return -2;
      active = 0;
      for (int i = 0; i < b->ninst; ++i) active += b->host_st[i].active ? 1 : 0;
      if (active == 0) break;
    }
  }
  if (n_active) *n_active = active;
  return 0;
}

int ss_batch_state(ss_batch* b, double* energies, double* err_absmax, int* steps, int* converged_step) {
  if (!b) return bfail(nullptr, -1, "null batch");
  const BatchBufs B = batch_bufs(b);
  BHIP(b, b->ops->batch_pass(b->prec, 5, b->tv, B, b->stream));
  BHIP(b, b->ops->batch_fin(0, b->ninst, B, 1, 0.0, b->stream));
  if (batch_read_state(b)) return -2;
  for (int i = 0; i < b->ninst; ++i) {
    if (energies) energies[i] = b->host_st[i].last_energy;
    if (err_absmax) err_absmax[i] = b->host_st[i].err_absmax;
    if (steps) steps[i] = b->host_st[i].step;
    if (converged_step) converged_step[i] = b->host_st[i].converged_step;
  }
  return 0;
}

int ss_batch_memory_bytes(ss_batch* b, size_t* bytes) {
  if (!b || !bytes) return bfail(b, -1, "null argument");
  *bytes = 7 * (size_t)b->total * elem_size(b->prec) + sizeof(double) * 2 * (size_t)b->ntcap;
  return 0;
}

// Times the four RK4 passes of the batch in each tile variant on the live
// buffers (inactive instances are skipped, as in a step) and keeps the
// fastest.  The passes write only l_k, the partials and the next-parity F, V,
// so the relaxation state is unchanged.
int ss_batch_autotune_tiles(ss_batch* b, int reps, int* chosen, double* ms) {
  if (!b) return bfail(nullptr, -1, "null batch");
  if (reps < 1) return bfail(b, -1, "reps must be >= 1");
  hipEvent_t e0, e1;
  BHIP(b, hipEventCreate(&e0));
  BHIP(b, hipEventCreate(&e1));
  int best_v = 0, best_f = 0;
  float best = 0.f;
  int rc = 0;
  const int nv = b->ops->num_shapes;
  for (int v = 0; v < nv && rc == 0; ++v) {
    rc = batch_tiles(b, v);
    if (rc) break;
    const BatchBufs B = batch_bufs(b);
    // f = 0: passes 1, 2, 3, 4; f = 1: passes 1, 2+3, 4 (where the variant has the sweep);
    // f = 2: the paired sweeps 1+2, 3+4 (strip variants)
    for (int f = 0; f < 3 && rc == 0; ++f) {
      if ((f == 1 && !b->ops->batch_fuse23(b->prec, v)) || (f == 2 && !b->ops->batch_pairs(b->prec, v))) {
        if (ms) ms[f * nv + v] = 0.0;
        continue;
      }
      const int seq[3][4] = {{1, 2, 3, 4}, {1, 6, 4, 0}, {7, 8, 0, 0}};
      hipError_t e = hipSuccess;
      for (int k = 0; k < 4 && seq[f][k] && e == hipSuccess; ++k)  // warm-up
        e = b->ops->batch_pass(b->prec, seq[f][k], v, B, b->stream);
      if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
      for (int r = 0; r < reps && e == hipSuccess; ++r)
        for (int k = 0; k < 4 && seq[f][k] && e == hipSuccess; ++k)
          e = b->ops->batch_pass(b->prec, seq[f][k], v, B, b->stream);
      if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float t = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
      if (e != hipSuccess) {
        rc = bfail(b, -2, "batch tile autotuning of variant %d%s failed: %s", v, f ? " (fused 2+3)" : "",
                   hipGetErrorString(e));
        break;
      }
      t /= reps;
      if (ms) ms[f * nv + v] = t;
      // a paired schedule must win by 3 % (its passes 1+2 read V before the decision)
      if ((v == 0 && f == 0) || (f == 2 ? t < 0.97f * best : t < best)) {
        best = t;
        best_v = v;
        best_f = f;
      }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) return rc;
  if (batch_tiles(b, best_v)) return -2;
  b->fuse23 = best_f == 1;
  b->pairs = best_f == 2;
  if (chosen) *chosen = best_v;
  return 0;
}

int ss_batch_set_tile_variant(ss_batch* b, int variant) {
  if (!b) return bfail(nullptr, -1, "null batch");
  if (variant < 0 || variant >= b->ops->num_shapes)
    return bfail(b, -1, "batch shape variant is %d (valid: 0 .. %d)", variant, b->ops->num_shapes - 1);
  return batch_tiles(b, variant);
}

int ss_batch_set_fused_stages(ss_batch* b, int flags) {
  if (!b) return bfail(nullptr, -1, "null batch");
  if (flags & ~(SS_FUSE_23 | SS_FUSE_PAIRS)) return bfail(b, -1, "unknown batch fused-stage flags 0x%x", flags);
  if ((flags & SS_FUSE_PAIRS) && (flags & SS_FUSE_23))
    return bfail(b, -1, "the paired schedule replaces the 2+3 sweep: set it alone");
  if ((flags & SS_FUSE_PAIRS) && !b->ops->batch_pairs(b->prec, b->tv))
    return bfail(b, -1, "passes 1+2 and 3+4 of theory %s cannot run as paired sweeps in shape variant %d",
                 b->ops->name, b->tv);
  if ((flags & SS_FUSE_23) && !b->ops->batch_fuse23(b->prec, b->tv))
    return bfail(b, -1, "passes 2 and 3 of theory %s cannot share one sweep in shape variant %d at this precision",
                 b->ops->name, b->tv);
  b->fuse23 = (flags & SS_FUSE_23) ? 1 : 0;
  b->pairs = (flags & SS_FUSE_PAIRS) ? 1 : 0;
  if (b->graph) {
    hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  return 0;
}

int ss_batch_get_fused_stages(ss_batch* b, int* flags) {
  if (!b) return bfail(nullptr, -1, "null batch");
  if (flags) *flags = (batch_use23(b) ? SS_FUSE_23 : 0) | (batch_use_pairs(b) ? SS_FUSE_PAIRS : 0);
  return 0;
}

int ss_batch_tile_variant(ss_batch* b, int* variant, int* rows) {
  if (!b) return bfail(nullptr, -1, "null batch");
  if (variant) *variant = b->tv;
  if (rows) *rows = b->sx ? b->sx : b->ops->shape_rows(b->prec, b->tv);
  return 0;
}

}  // extern "C"
