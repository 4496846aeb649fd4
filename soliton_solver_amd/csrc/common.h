// Shared device-side definitions of the relaxation core: lattice geometry,
// stencil coefficients, the LDS tile shape and block reductions.
//
// Field layout in HBM is the reference's host layout (core/utils.py:16-42):
// component-major planes, each plane xlen rows of ylen contiguous values
// (y fastest): idx(a, x, y) = a*xlen*ylen + x*ylen + y.  Planes are
// structure-of-arrays, so one wavefront reads 64 consecutive y of one plane.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ss {

// 4th-order central coefficients (core/derivatives.py:51,98,106-110).
#define SS_C12 (1.0 / 12.0)
#define SS_C23 (2.0 / 3.0)
#define SS_C43 (4.0 / 3.0)
#define SS_C52 (5.0 / 2.0)

// Reference packing limits: the longest theory packing is 20 ints / 27 doubles.
constexpr int kMaxPI = 32;
constexpr int kMaxPF = 48;

// Lattice geometry and step size, resolved on the host from p_i / p_f.
struct Geom {
  int X, Y;        // xlen, ylen
  int halo;        // p_i[2]; stencil radius is always 2, halo only sets the band
  int nf;          // number_total_fields
  long long N;     // X*Y
  double lsx, lsy, gv, dt;
  double inv_lsx, inv_lsy;      // 1/lsx, 1/lsy
  double inv_lsx2, inv_lsy2;    // 1/(lsx*lsx), 1/(lsy*lsy)
  double inv_2gv;               // 1/(2 gv)
  double inv_2lsx2, inv_2lsy2;  // 1/(2 lsx lsx), 1/(2 lsy lsy)
  int killkinen, unit_mag;
};

// Where a site sits relative to the reference's derivative masks
// (core/derivatives.py:50,54,97,101,105) and the boundary band.
struct SiteInfo {
  int x, y;
  bool inx;    // halo-1 < x < X-halo: d/dx, d2/dx2 are computed
  bool iny;    // halo-1 < y < Y-halo
  bool inxy;   // halo < x < X-halo and halo < y < Y-halo: mixed derivative
  bool band;   // x < halo or x > X-halo-1 or y < halo or y > Y-halo-1
};

__device__ __forceinline__ SiteInfo site_info(const Geom& g, int x, int y) {
  SiteInfo s;
  s.x = x;
  s.y = y;
  s.inx = (x > g.halo - 1) && (x < g.X - g.halo);
  s.iny = (y > g.halo - 1) && (y < g.Y - g.halo);
  s.inxy = (x > g.halo) && (y > g.halo) && (x < g.X - g.halo) && (y < g.Y - g.halo);
  s.band = !(s.inx && s.iny);
  return s;
}

// Tile shape: TY consecutive y (one wavefront wide) by TX rows of x, plus a
// 2-site halo on each side for the 5-point stencils.
// Auxiliary per-site quantities of the reference's Simulation
// (core/simulation.py:211-386): |psi|^2 of the order parameter(s), the cosine
// of the phase difference, the supercurrent and the magnetic field.  A theory
// reports how many components it has of each (0: not defined for it).
enum AuxKind { kAuxHiggs = 0, kAuxHiggs1 = 1, kAuxHiggs2 = 2, kAuxPhaseDifference = 3, kAuxSupercurrent = 4,
               kAuxMagneticField = 5, kNumAuxKinds = 6 };
// density_pass kind codes >= kAuxBase select aux kind (code - kAuxBase) / 8, component (code - kAuxBase) % 8.
constexpr int kAuxBase = 1000;

constexpr int TY = 64;
constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

// Device-resident state of one relaxation: arrest bookkeeping and results.
struct DevState {
  double prev_energy;   // energy the next kill test compares against
  double last_energy;   // most recent total energy
  double err_max;       // signed max of the stage-4 energy gradient
  double err_absmax;    // max |stage-4 energy gradient|
  int kill;             // velocity of the current step is treated as zero
  int step;             // steps done in the current call
  int pad[2];
  double eslot[2];      // energy of the last field finalized in buffer parity p (read by folded finalizes)
};

// ---------------------------------------------------------------------------
// Block reductions: DPP/shuffle inside a wavefront, then LDS across waves.
// Deterministic for a fixed launch shape.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// NaN-propagating maximum (IEEE 754-2019 maximum): NaN in either operand gives
// NaN.  Every error reduction uses it, so a non-finite energy gradient cannot
// vanish from max |EG| the way it does under fmax (maxNum drops a NaN operand),
// and a blown-up field never reads as converged.
__device__ __forceinline__ double nmax(double a, double b) { return __builtin_elementwise_maximum(a, b); }
__device__ __forceinline__ float nmax(float a, float b) { return __builtin_elementwise_maximum(a, b); }

// Per-thread signed and absolute maxima of the stage-4 energy gradient over
// the sites one thread visits, folded into the thread's FP64 accumulators at
// the end.  FP32 values are compared in FP32 (exact, and gfx950's
// v_maximum3_f32 propagates NaN in one instruction, |e| as an operand
// modifier); FP64 values with fmax plus an unordered-compare flag (one compare
// per value, where a NaN-propagating FP64 maximum costs three more
// instructions).
template <typename R>
struct ErrMax;
template <>
struct ErrMax<float> {
  float m0 = -__builtin_inff(), m1 = 0.f;
  __device__ __forceinline__ void add(float e) {
    m0 = nmax(m0, e);
    m1 = nmax(m1, __builtin_fabsf(e));
  }
  __device__ __forceinline__ void fold(double& acc0, double& acc1) const {
    acc0 = nmax(acc0, (double)m0);
    acc1 = nmax(acc1, (double)m1);
  }
};
template <>
struct ErrMax<double> {
  double m0 = -1.0e300, m1 = 0.0;
  bool bad = false;
  __device__ __forceinline__ void add(double e) {
    m0 = fmax(m0, e);
    m1 = fmax(m1, __builtin_fabs(e));
    bad |= (e != e);
  }
  __device__ __forceinline__ void fold(double& acc0, double& acc1) const {
    const double nan = __builtin_nan("");
    acc0 = nmax(acc0, bad ? nan : m0);
    acc1 = nmax(acc1, bad ? nan : m1);
  }
};

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = nmax(v, __shfl_xor(v, off, 64));
  return v;
}

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) s += red[i];
  }
  return s;  // valid in thread 0
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double* red) {
  v = wave_max(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = -1.0e300;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) s = nmax(s, red[i]);
  }
  return s;  // valid in thread 0
}

// XCD-aware block -> tile mapping.  Workgroups are dispatched round-robin
// over the 8 XCDs (each with its own L2); this remap gives every XCD a
// contiguous band of tiles so halo re-reads between neighbouring tiles stay
// in one L2.  Returns -1 for the padding blocks of the last band.
__device__ __forceinline__ int xcd_tile(int bid, int ntiles, int per_xcd) {
  const int t = (bid & 7) * per_xcd + (bid >> 3);
  return t < ntiles ? t : -1;
}

}  // namespace ss
