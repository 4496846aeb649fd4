// Fused arrested-Newton RK4 passes, templated on the theory, the storage
// precision R and the pass number.
//
// The reference runs one RK4 step as 4 x (stage kernel + gradient kernel with
// global derivative scratch) + a combine + an energy kernel + 2 host
// reductions (core/integrator.py:217-348).  Here one step is four passes over
// HBM; every pass rebuilds its stage field on the fly from (F, V, l_k) into an
// LDS tile with a 2-site halo, evaluates the stencils and the theory's
// variation from LDS, and writes only what later passes need:
//
//   pass 1: Temp = F                         -> l1 = -dt EG,  energy partials of F
//   pass 2: Temp = F + 0.5 dt V              -> l2
//   pass 3: Temp = F + 0.5 dt (V + 0.5 l1)   -> l3
//   pass 4: Temp = F + dt (V + 0.5 l2)       -> l4 (registers), combine:
//           V' = V + (l1 + 2 l2 + 2 l3 + l4)/6, F' = F + (k1 + 2 k2 + 2 k3 + k4)/6
//           with k1 = dt V, k2 = dt (V + l1/2), k3 = dt (V + l2/2), k4 = dt (V + l3),
//           then the unit-norm constraint; signed max and max |EG| partials.
//
// These are the reference's stage formulas (integrator.py:50-96, 98-189) with
// k_s recomputed from V and l_{s-1} instead of stored.  The velocity kill of
// step n (E(F_{n+1}) > E_prev, integrator.py:339-343) is decided on the device
// after pass 1 of step n+1, which computes E(F_{n+1}); passes 2-4 then read V
// as zero, exactly as the reference's zeroed V would be read.
#pragma once

#include "common.h"
#include "stencil.h"

namespace ss {

struct PassBufs {
  const void* F;
  const void* V;
  const void* l1;
  const void* l2;
  const void* l3;
  void* lout;      // pass 0: EG; passes 1-3: l_pass
  void* Fout;      // pass 4
  void* Vout;      // pass 4
  double* part0;   // pass 1: energy partial per tile; pass 4: signed max partial
  double* part1;   // pass 4: max |EG| partial
  const DevState* st;
};

struct Launch {
  int ntiles, ntiles_y, per_xcd, grid;
};

template <typename R, int NF>
inline Launch make_launch(const Geom& g) {
  using S = TileShape<R, NF>;
  Launch L;
  const int ntx = (g.X + S::TX - 1) / S::TX;
  L.ntiles_y = (g.Y + TY - 1) / TY;
  L.ntiles = ntx * L.ntiles_y;
  L.per_xcd = (L.ntiles + 7) / 8;
  L.grid = 8 * L.per_xcd;
  return L;
}

template <class TH, typename R, int PASS>
__global__ void __launch_bounds__(kThreads)
stage_pass(PassBufs b, typename TH::TP tp, Geom g, Launch L) {
  using S = TileShape<R, TH::NF>;
  constexpr int NF = TH::NF;
  __shared__ Tile<R, NF> tile;
  __shared__ double red[kWaves];

  const int t = xcd_tile(blockIdx.x, L.ntiles, L.per_xcd);
  if (t < 0) return;  // whole block exits together, before any barrier
  const int tx = t / L.ntiles_y, ty = t - tx * L.ntiles_y;
  const int x0 = tx * S::TX, y0 = ty * TY;

  const R* __restrict__ F = (const R*)b.F;
  const R* __restrict__ V = (const R*)b.V;
  const R* __restrict__ l1 = (const R*)b.l1;
  const R* __restrict__ l2 = (const R*)b.l2;
  const R* __restrict__ l3 = (const R*)b.l3;
  const bool kill = (PASS >= 2) ? (b.st->kill != 0) : false;
  const R dt = R(g.dt);
  const R half = R(0.5);

  load_tile<R, NF>(tile, g, x0, y0, [&](int a, long long off) -> R {
    if (PASS <= 1) return F[off];
    const R v = kill ? R(0) : V[off];
    if (PASS == 2) return F[off] + half * (dt * v);
    if (PASS == 3) return F[off] + half * (dt * (v + half * l1[off]));
    return F[off] + dt * (v + half * l2[off]);
  });
  __syncthreads();

  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double acc0 = (PASS == 4) ? -1.0e300 : 0.0;
  double acc1 = 0.0;
#pragma unroll
  for (int k = 0; k < S::kRowsPerWave; ++k) {
    const int lx = w * S::kRowsPerWave + k, ly = lane;
    const int x = x0 + lx, y = y0 + ly;
    if (x < g.X && y < g.Y) {
      const SiteInfo s = site_info(g, x, y);
      R f[NF], d1[NF][2], d2[NF][3], eg[NF];
      tile_derivs<R, R, NF>(tile, lx, ly, g, s, f, d1, d2);
      TH::template variation<R>(tp, g, s, f, d1, d2, eg);
      const long long off = (long long)x * g.Y + y;
      if (PASS == 0) {
        R* out = (R*)b.lout;
#pragma unroll
        for (int a = 0; a < NF; ++a) out[off + a * g.N] = eg[a];
      } else if (PASS <= 3) {
        R* out = (R*)b.lout;
#pragma unroll
        for (int a = 0; a < NF; ++a) out[off + a * g.N] = -dt * eg[a];
      }
      if (PASS == 1) {
        double fd[NF], d1d[NF][2];
        tile_d1<double, R, NF>(tile, lx, ly, g, s, fd, d1d);
        acc0 += TH::template energy<double>(tp, g, s, fd, d1d) * g.gv;
      }
      if (PASS == 4) {
        R Fn[NF], Vn[NF];
        const R sixth = R(1.0 / 6.0);
#pragma unroll
        for (int a = 0; a < NF; ++a) {
          const long long o = off + a * g.N;
          const R Fc = F[o];
          const R Vc = kill ? R(0) : V[o];
          const R a1 = l1[o], a2 = l2[o], a3 = l3[o];
          const R a4 = -dt * eg[a];
          const R k1 = dt * Vc;
          const R k2 = dt * (Vc + half * a1);
          const R k3 = dt * (Vc + half * a2);
          const R k4 = dt * (Vc + a3);
          Vn[a] = Vc + sixth * (a1 + R(2) * a2 + R(2) * a3 + a4);
          Fn[a] = Fc + sixth * (k1 + R(2) * k2 + R(2) * k3 + k4);
          const double e = (double)eg[a];
          acc0 = fmax(acc0, e);
          acc1 = fmax(acc1, fabs(e));
        }
        if (TH::kUnitConstraint && g.unit_mag) {
          TH::template normalize<R>(Fn);
          TH::template project<R>(Vn, Fn);
        }
        R* Fo = (R*)b.Fout;
        R* Vo = (R*)b.Vout;
#pragma unroll
        for (int a = 0; a < NF; ++a) {
          Fo[off + a * g.N] = Fn[a];
          Vo[off + a * g.N] = Vn[a];
        }
      }
    }
  }
  if (PASS == 1) {
    const double s = block_sum<kThreads>(acc0, red);
    if (threadIdx.x == 0) b.part0[t] = s;
  } else if (PASS == 4) {
    const double m0 = block_max<kThreads>(acc0, red);
    if (threadIdx.x == 0) b.part0[t] = m0;
    const double m1 = block_max<kThreads>(acc1, red);
    if (threadIdx.x == 0) b.part1[t] = m1;
  }
}

// Energy (kind < 0) or charge density ``kind`` of F, times gv: block partial
// sums, and optionally the per-site density in reference order.
template <class TH, typename R>
__global__ void __launch_bounds__(kThreads)
density_pass(const void* Fv, double* dens, double* part, int kind, typename TH::TP tp, Geom g, Launch L) {
  using S = TileShape<R, TH::NF>;
  constexpr int NF = TH::NF;
  __shared__ Tile<R, NF> tile;
  __shared__ double red[kWaves];
  const int t = xcd_tile(blockIdx.x, L.ntiles, L.per_xcd);
  if (t < 0) return;
  const int tx = t / L.ntiles_y, ty = t - tx * L.ntiles_y;
  const int x0 = tx * S::TX, y0 = ty * TY;
  const R* __restrict__ F = (const R*)Fv;
  load_tile<R, NF>(tile, g, x0, y0, [&](int a, long long off) -> R { return F[off]; });
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < S::kRowsPerWave; ++k) {
    const int lx = w * S::kRowsPerWave + k, ly = lane;
    const int x = x0 + lx, y = y0 + ly;
    if (x < g.X && y < g.Y) {
      const SiteInfo s = site_info(g, x, y);
      double fd[NF], d1d[NF][2];
      tile_d1<double, R, NF>(tile, lx, ly, g, s, fd, d1d);
      const double v = (kind < 0 ? TH::template energy<double>(tp, g, s, fd, d1d)
                                 : TH::template charge<double>(kind, tp, g, s, fd, d1d)) * g.gv;
      if (dens) dens[(long long)x * g.Y + y] = v;
      acc += v;
    }
  }
  const double s = block_sum<kThreads>(acc, red);
  if (threadIdx.x == 0) part[t] = s;
}

// Host-side launchers for one theory.  ``prec`` 0 = FP64 storage, 1 = FP32.
template <class TH>
struct TheoryLaunchers {
  using TP = typename TH::TP;

  static int ntiles(int prec, const Geom& g) {
    return prec ? make_launch<float, TH::NF>(g).ntiles : make_launch<double, TH::NF>(g).ntiles;
  }

  template <typename R>
  static hipError_t pass_r(int pass, const PassBufs& b, const TP& tp, const Geom& g, hipStream_t st) {
    const Launch L = make_launch<R, TH::NF>(g);
    dim3 grid(L.grid), block(kThreads);
    switch (pass) {
      case 0: hipLaunchKernelGGL((stage_pass<TH, R, 0>), grid, block, 0, st, b, tp, g, L); break;
      case 1: hipLaunchKernelGGL((stage_pass<TH, R, 1>), grid, block, 0, st, b, tp, g, L); break;
      case 2: hipLaunchKernelGGL((stage_pass<TH, R, 2>), grid, block, 0, st, b, tp, g, L); break;
      case 3: hipLaunchKernelGGL((stage_pass<TH, R, 3>), grid, block, 0, st, b, tp, g, L); break;
      case 4: hipLaunchKernelGGL((stage_pass<TH, R, 4>), grid, block, 0, st, b, tp, g, L); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }

  static hipError_t pass(int prec, int pass, const PassBufs& b, const void* tp, const Geom& g, hipStream_t st) {
    return prec ? pass_r<float>(pass, b, *(const TP*)tp, g, st) : pass_r<double>(pass, b, *(const TP*)tp, g, st);
  }

  template <typename R>
  static hipError_t density_r(int kind, const void* F, double* dens, double* part, const TP& tp, const Geom& g,
                              hipStream_t st) {
    const Launch L = make_launch<R, TH::NF>(g);
    hipLaunchKernelGGL((density_pass<TH, R>), dim3(L.grid), dim3(kThreads), 0, st, F, dens, part, kind, tp, g, L);
    return hipGetLastError();
  }

  static hipError_t density(int prec, int kind, const void* F, double* dens, double* part, const void* tp,
                            const Geom& g, hipStream_t st) {
    return prec ? density_r<float>(kind, F, dens, part, *(const TP*)tp, g, st)
                : density_r<double>(kind, F, dens, part, *(const TP*)tp, g, st);
  }

  static int make_tp(const int32_t* pi, int ni, const double* pf, int nfl, void* tp, char* msg, int cap) {
    static_assert(sizeof(TP) <= 512, "theory parameter block too large");
    return TH::make(pi, ni, pf, nfl, (TP*)tp, msg, cap);
  }
};

// Table of entry points one theory exposes to the context code.
struct TheoryOps {
  const char* name;
  int nf;
  int ncharges;
  int (*make_tp)(const int32_t*, int, const double*, int, void*, char*, int);
  int (*ntiles)(int, const Geom&);
  hipError_t (*pass)(int, int, const PassBufs&, const void*, const Geom&, hipStream_t);
  hipError_t (*density)(int, int, const void*, double*, double*, const void*, const Geom&, hipStream_t);
};

#define SS_DEFINE_THEORY_OPS(TH, fn)                                                        \
  const ss::TheoryOps* fn() {                                                               \
    static const ss::TheoryOps ops = {TH::kName, TH::NF, TH::kNumCharges,                   \
                                      &ss::TheoryLaunchers<TH>::make_tp,                    \
                                      &ss::TheoryLaunchers<TH>::ntiles,                     \
                                      &ss::TheoryLaunchers<TH>::pass,                       \
                                      &ss::TheoryLaunchers<TH>::density};                   \
    return &ops;                                                                            \
  }

}  // namespace ss
