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

#include <string.h>

#include "common.h"
#include "dual.h"
#include "stencil.h"
#include "stream.h"
#include "init.h"

namespace ss {

struct Launch {
  int ntiles, ntiles_y, per_xcd, grid;
};

template <typename R, int NF, int TXR = 0>
inline Launch make_launch(const Geom& g) {
  using S = TileShape<R, NF, TXR>;
  Launch L;
  const int ntx = (g.X + S::TX - 1) / S::TX;
  L.ntiles_y = (g.Y + TY - 1) / TY;
  L.ntiles = ntx * L.ntiles_y;
  L.per_xcd = (L.ntiles + 7) / 8;
  L.grid = 8 * L.per_xcd;
  return L;
}

// Components of the LDS tile of a pass: PASS 23 (passes 2 and 3 in one sweep)
// holds Temp2 in components 0..NF-1 and Temp3 in NF..2NF-1; PASS 41 (pass 4
// and pass 1 of the next step) Temp4 and the new field F'.
template <class TH, int PASS>
constexpr int tile_nf() {
  return (PASS == 23 || PASS == 41) ? 2 * TH::NF : TH::NF;
}

// A two-field tile (PASS 23 or 41) with TX rows fits in 64 KiB of LDS.
template <class TH, typename R, int TX>
constexpr bool fused_tile_fits() {
  return (TX + 4) * 2 * TH::NF * (TY + 4) * (int)sizeof(R) <= 65536;
}

// One tile of one pass: stage field into LDS, stencils + variation from LDS,
// outputs for the sites of the tile, and the block's partial (energy sum for
// pass 1, signed max / max |EG| for pass 4) written to part0[t] / part1[t].
// PASS 23: both stage fields of passes 2 and 3 from one set of loads, l2 ->
// lout and l3 -> lout2 (the rows are swept once per stage field).  PASS 41:
// Temp4 and F' (new_field, pointwise in F, V, l1, l2, l3) for the tile and its
// halo from one set of loads; pass 4 on Temp4 (F', V' -> Fout, Vout, err
// partials to part0/part1), then pass 1 of the next step on F' (l1 -> lout,
// energy partial -> part2).
// Shared by the single-lattice kernel and the batched kernel.
template <class TH, typename R, int PASS, int TXR = 0, bool COMP = false>
__device__ __forceinline__ void stage_tile(Tile<R, tile_nf<TH, PASS>(), TXR>& tile, double* red, const PassBufs& b,
                                           const typename TH::TP& tp, const Geom& g, int x0, int y0, bool kill,
                                           int t) {
  using S = TileShape<R, TH::NF, TXR>;
  constexpr int NF = TH::NF;
  constexpr int NFT = tile_nf<TH, PASS>();
  const R* __restrict__ F = (const R*)b.F;
  const R* __restrict__ V = (const R*)b.V;
  const R* __restrict__ l1 = (const R*)b.l1;
  const R* __restrict__ l2 = (const R*)b.l2;
  const R* __restrict__ l3 = (const R*)b.l3;
  const R dt = R(g.dt);
  const R half = R(0.5);

  constexpr int RW = S::kRowsPerWave;
  constexpr int RY = S::RY, TXr = S::TX;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int y = y0 + lane;

  // Phase A: the interior sites this thread computes (rows w*RW..w*RW+RW-1 of
  // column ``lane``), one coalesced row of 64 values per wave and component.
  // Pass 4 keeps V and l2 of its own sites in registers for the combine (the
  // 4+1 sweep also l1 and l3).
  constexpr bool kP4 = PASS == 4 || PASS == 41;
  R Vk[kP4 ? RW : 1][NF], L2k[kP4 ? RW : 1][NF];
  R L1k[PASS == 41 ? RW : 1][NF], L3k[PASS == 41 ? RW : 1][NF];
  if constexpr (PASS == 41) {
#pragma unroll
    for (int k = 0; k < RW; ++k) {
      const int lx = w * RW + k, x = x0 + lx;
      const bool in = (x < g.X) && (y < g.Y);
      const long long off = in ? (long long)x * g.Y + y : 0;
      R f[NF], T4[NF], Fn[NF];
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        const long long o = off + (long long)a * g.N;
        f[a] = in ? F[o] : R(0);
        Vk[k][a] = (in && !kill) ? V[o] : R(0);
        L1k[k][a] = in ? l1[o] : R(0);
        L2k[k][a] = in ? l2[o] : R(0);
        L3k[k][a] = in ? l3[o] : R(0);
      }
      new_field<TH, R>(tp, g, in, f, Vk[k], L1k[k], L2k[k], L3k[k], dt, T4, Fn);
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        tile.v[a][lx + 2][lane + 2] = T4[a];
        tile.v[NF + a][lx + 2][lane + 2] = Fn[a];
      }
    }
  } else {
#pragma unroll
  for (int k = 0; k < RW; ++k) {
    const int lx = w * RW + k, x = x0 + lx;
    const bool in = (x < g.X) && (y < g.Y);
    const long long off = (long long)x * g.Y + y;
#pragma unroll
    for (int a = 0; a < NF; ++a) {
      const long long o = off + (long long)a * g.N;
      R val = R(0), val3 = R(0);
      if (PASS == 4) {
        Vk[k][a] = R(0);
        L2k[k][a] = R(0);
      }
      if (in) {
        if (PASS <= 1) {
          val = F[o];
        } else {
          const R fo = F[o];
          const R v = kill ? R(0) : V[o];
          if (PASS == 2) {
            val = fo + half * (dt * v);
          } else if (PASS == 3) {
            val = fo + half * (dt * (v + half * l1[o]));
          } else if (PASS == 23) {
            val = fo + half * (dt * v);
            val3 = fo + half * (dt * (v + half * l1[o]));
          } else {
            const R q = l2[o];
            val = fo + dt * (v + half * q);
            Vk[k][a] = v;
            L2k[k][a] = q;
          }
        }
      }
      tile.v[a][lx + 2][lane + 2] = val;
      if constexpr (PASS == 23) tile.v[NF + a][lx + 2][lane + 2] = val3;
    }
  }
  }
  // Phase B: the 2-site halo ring (2 rows above and below, 2 columns either side).
  constexpr int kRing = 4 * RY + 4 * TXr;
  for (int idx = threadIdx.x; idx < kRing; idx += kThreads) {
    int r, c;
    if (idx < 2 * RY) {
      r = idx / RY;
      c = idx - r * RY;
    } else if (idx < 4 * RY) {
      const int j = idx - 2 * RY;
      r = TXr + 2 + j / RY;
      c = j - (j / RY) * RY;
    } else {
      const int j = idx - 4 * RY;
      r = 2 + (j >> 2);
      const int q = j & 3;
      c = q < 2 ? q : RY - 4 + q;
    }
    const int gx = x0 - 2 + r, gy = y0 - 2 + c;
    const bool in = (gx >= 0) && (gx < g.X) && (gy >= 0) && (gy < g.Y);
    const long long off = (long long)gx * g.Y + gy;
    if constexpr (PASS == 41) {
      R f[NF], v[NF], a1[NF], a2[NF], a3[NF], T4[NF], Fn[NF];
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        const long long o = (in ? off : 0) + (long long)a * g.N;
        f[a] = in ? F[o] : R(0);
        v[a] = (in && !kill) ? V[o] : R(0);
        a1[a] = in ? l1[o] : R(0);
        a2[a] = in ? l2[o] : R(0);
        a3[a] = in ? l3[o] : R(0);
      }
      new_field<TH, R>(tp, g, in, f, v, a1, a2, a3, dt, T4, Fn);
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        tile.v[a][r][c] = T4[a];
        tile.v[NF + a][r][c] = Fn[a];
      }
      continue;
    }
#pragma unroll
    for (int a = 0; a < NF; ++a) {
      const long long o = off + (long long)a * g.N;
      R val = R(0), val3 = R(0);
      if (in) {
        if (PASS <= 1) {
          val = F[o];
        } else {
          const R v = kill ? R(0) : V[o];
          if (PASS == 2) val = F[o] + half * (dt * v);
          else if (PASS == 3) val = F[o] + half * (dt * (v + half * l1[o]));
          else if (PASS == 23) {
            val = F[o] + half * (dt * v);
            val3 = F[o] + half * (dt * (v + half * l1[o]));
          } else val = F[o] + dt * (v + half * l2[o]);
        }
      }
      tile.v[a][r][c] = val;
      if constexpr (PASS == 23) tile.v[NF + a][r][c] = val3;
    }
  }
  __syncthreads();

  // Phase C: march down the thread's rows with an x-window in registers (PASS
  // 23 and 41: once per field of the tile, components ao .. ao+NF-1; PASS 41
  // runs pass 4 on the first and pass 1 on the second).
  const int cy = lane + 2;
  double acc0 = kP4 ? -1.0e300 : 0.0;
  double acc1 = 0.0, acc2 = 0.0;
  ErrMax<R> em;  // pass 4: the stage-4 error of this thread's sites
#pragma unroll
  for (int stg = 0; stg < NFT / NF; ++stg) {
  const int ao = stg * NF;
  R win[NF][5];
#pragma unroll
  for (int a = 0; a < NF; ++a) {
#pragma unroll
    for (int j = 0; j < 4; ++j) win[a][j] = tile.v[ao + a][w * RW + j][cy];
  }
#pragma unroll
  for (int k = 0; k < RW; ++k) {
    const int lx = w * RW + k, cx = lx + 2;
    const int x = x0 + lx;
#pragma unroll
    for (int a = 0; a < NF; ++a) win[a][4] = tile.v[ao + a][cx + 2][cy];
    if (x < g.X && y < g.Y) {
      const SiteInfo s = site_info(g, x, y);
      R f[NF], d1[NF][2], d2[NF][3], eg[NF];
#pragma unroll
      for (int a = 0; a < NF; ++a)
        window_derivs<R, R, NFT>(tile, ao + a, win[a], cx, cy, g, s, f[a], d1[a], d2[a]);
      TH::template variation<R>(tp, g, s, f, d1, d2, eg);
      const long long off = (long long)x * g.Y + y;
      if (PASS == 0) {
        R* out = (R*)b.lout;
#pragma unroll
        for (int a = 0; a < NF; ++a) out[off + a * g.N] = eg[a];
      } else if (PASS <= 3 || PASS == 23 || (PASS == 41 && stg == 1)) {
        R* out = (R*)((PASS == 23 && stg) ? b.lout2 : b.lout);
#pragma unroll
        for (int a = 0; a < NF; ++a) out[off + a * g.N] = -dt * eg[a];
      }
      if (PASS == 1 || (PASS == 41 && stg == 1)) {
        // FP64 energy density from the same window values (no LDS re-read).
        double fd[NF], d1d[NF][2];
#pragma unroll
        for (int a = 0; a < NF; ++a)
          window_d1<double, R, NFT>(tile, ao + a, win[a], cx, cy, g, s, fd[a], d1d[a]);
        const double e1 = TH::template energy<double>(tp, g, s, fd, d1d) * g.gv;
        if (PASS == 41)
          acc2 += e1;
        else
          acc0 += e1;
      }
      if (PASS == 4 || (PASS == 41 && stg == 0)) {
        // RK4 combine.  With Temp4 = F + dt (V + l2/2) at the centre (f, from
        // LDS), F' = F + (k1 + 2 k2 + 2 k3 + k4)/6 = Temp4 + dt (l1 - 2 l2 + l3)/6,
        // so F is not re-read; V and l2 are the registers kept in phase A.
        // COMP (FP32): the new field is formed by compensated_field from F and
        // its low words instead.  PASS 41: F' is the tile's second field
        // (normalised by new_field), l1 and l3 the registers of phase A.
        R Fn[NF], Vn[NF];
        double Fd[COMP ? NF : 1];  // COMP: F + increment in FP64
        const R sixth = R(1.0 / 6.0);
        const R dt6 = dt * sixth;
#pragma unroll
        for (int a = 0; a < NF; ++a) {
          const long long o = off + a * g.N;
          const R Vc = Vk[k][a];
          const R a2 = L2k[k][a];
          const R a1 = (PASS == 41) ? L1k[k][a] : l1[o], a3 = (PASS == 41) ? L3k[k][a] : l3[o];
          const R a4 = -dt * eg[a];
          Vn[a] = Vc + sixth * (a1 + R(2) * a2 + R(2) * a3 + a4);
          if constexpr (COMP)
            Fd[a] = compensated_sum(F[o], b.Flo[o], dt * (Vc + half * a2) + dt6 * (a1 - R(2) * a2 + a3));
          else if constexpr (PASS == 41)
            Fn[a] = tile.v[NF + a][cx][cy];
          else
            Fn[a] = f[a] + dt6 * (a1 - R(2) * a2 + a3);
          em.add(eg[a]);
        }
        if constexpr (COMP) {
          static_assert(sizeof(R) == 4, "compensated combine: FP32 only");
          float lon[NF];
          compensated_split<TH>(tp, g, Fd, Fn, lon);
#pragma unroll
          for (int a = 0; a < NF; ++a) b.Flo[off + a * g.N] = lon[a];
          if (TH::kUnitConstraint && g.unit_mag) TH::template project<R>(tp, Vn, Fn);
        } else if (TH::kUnitConstraint && g.unit_mag) {
          if (PASS != 41) TH::template normalize<R>(tp, Fn);
          TH::template project<R>(tp, Vn, Fn);
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
#pragma unroll
    for (int a = 0; a < NF; ++a) {
#pragma unroll
      for (int j = 0; j < 4; ++j) win[a][j] = win[a][j + 1];
    }
  }
  }  // stage fields
  if (kP4) em.fold(acc0, acc1);
  if (PASS == 1) {
    const double s = block_sum<kThreads>(acc0, red);
    if (threadIdx.x == 0) b.part0[t] = s;
  } else if (kP4) {
    const double m0 = block_max<kThreads>(acc0, red);
    if (threadIdx.x == 0) b.part0[t] = m0;
    const double m1 = block_max<kThreads>(acc1, red);
    if (threadIdx.x == 0) b.part1[t] = m1;
    if (PASS == 41) {
      const double e = block_sum<kThreads>(acc2, red);
      if (threadIdx.x == 0) b.part2[t] = e;
    }
  }
}

template <class TH, typename R, int PASS, int TXR = 0, bool COMP = false>
__global__ void __launch_bounds__(kThreads)
stage_pass(PassBufs b, typename TH::TP tp, Geom g, Launch L) {
  using S = TileShape<R, TH::NF, TXR>;
  __shared__ Tile<R, tile_nf<TH, PASS>(), TXR> tile;
  __shared__ double red[kWaves];
  const int t = xcd_tile(blockIdx.x, L.ntiles, L.per_xcd);
  if (t < 0) return;  // whole block exits together, before any barrier
  const int tx = t / L.ntiles_y, ty = t - tx * L.ntiles_y;
  bool kill;
  if constexpr (PASS == 23)  // the previous step's finalize folded in, or its decision
    kill = b.fin_n > 0 ? fold_finalize<kThreads>(b, t == 0, red) : (b.st->kill != 0);
  else
    kill = (PASS >= 2) ? (b.st->kill != 0) : false;
  stage_tile<TH, R, PASS, TXR, COMP>(tile, red, b, tp, g, tx * S::TX, ty * TY, kill, t);
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
      double v;
      if (kind < 0) {
        tile_d1<double, R, NF>(tile, lx, ly, g, s, fd, d1d);
        v = TH::template energy<double>(tp, g, s, fd, d1d) * g.gv;
      } else if (kind >= kAuxBase) {
        double d2d[NF][3];
        tile_derivs<double, R, NF>(tile, lx, ly, g, s, fd, d1d, d2d);
        v = TH::template aux<double>((kind - kAuxBase) >> 3, (kind - kAuxBase) & 7, tp, g, s, fd, d1d, d2d);
      } else {
        double d2d[NF][3];
        tile_derivs<double, R, NF>(tile, lx, ly, g, s, fd, d1d, d2d);
        v = TH::template charge<double>(kind, tp, g, s, fd, d1d, d2d) * g.gv;
      }
      if (dens) dens[(long long)x * g.Y + y] = v;
      acc += v;
    }
  }
  const double s = block_sum<kThreads>(acc, red);
  if (threadIdx.x == 0) part[t] = s;
}

// Packed parameters by value (kernel argument).
struct PackedPF {
  double v[kMaxPF];
};

// Theories whose coupling constants are built by fill_tp in any scalar type
// (kParamGrad): the energy can be differentiated with respect to p_f.
template <class T, class = void>
struct HasParamGrad {
  static constexpr bool value = false;
};
template <class T>
struct HasParamGrad<T, decltype((void)T::kParamGrad)> {
  static constexpr bool value = T::kParamGrad;
};

// dE_h/dp_f[j] for j = j0 .. j0+nj-1 at the FP64 field Fv: forward-mode dual
// numbers seeded at p_f[j] through the theory's own parameter derivation
// (fill_tp) and energy density, field fixed.  Per tile and parameter one FP64
// partial: part[jj * ntiles + t].
template <class TH>
__global__ void __launch_bounds__(kThreads)
param_grad_pass(const double* Fv, double* part, PackedPF pf, int j0, int nj, Geom g, Launch L) {
  if constexpr (HasParamGrad<TH>::value) {
    using S = TileShape<double, TH::NF>;
    constexpr int NF = TH::NF, NP = TH::kNumPF;
    __shared__ Tile<double, NF> tile;
    __shared__ double red[kWaves];
    const int t = xcd_tile(blockIdx.x, L.ntiles, L.per_xcd);
    if (t < 0) return;
    const int tx = t / L.ntiles_y, ty = t - tx * L.ntiles_y;
    const int x0 = tx * S::TX, y0 = ty * TY;
    load_tile<double, NF>(tile, g, x0, y0, [&](int a, long long off) -> double { return Fv[off]; });
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int jj = 0; jj < nj; ++jj) {
      const int j = j0 + jj;
      Dual<1> pd[NP];
#pragma unroll
      for (int k = 0; k < NP; ++k) {
        pd[k] = Dual<1>(pf.v[k]);
        pd[k].d[0] = (k == j) ? 1.0 : 0.0;
      }
      typename TH::template TPT<Dual<1>> tpd;
      TH::template fill_tp<Dual<1>>(pd, tpd);
      double acc = 0.0;
      for (int k = 0; k < S::kRowsPerWave; ++k) {
        const int lx = w * S::kRowsPerWave + k, ly = lane;
        const int x = x0 + lx, y = y0 + ly;
        if (x < g.X && y < g.Y) {
          const SiteInfo s = site_info(g, x, y);
          double fd[NF], d1d[NF][2];
          tile_d1<double, double, NF>(tile, lx, ly, g, s, fd, d1d);
          Dual<1> f[NF], d1[NF][2];
#pragma unroll
          for (int a = 0; a < NF; ++a) {
            f[a] = Dual<1>(fd[a]);
            d1[a][0] = Dual<1>(d1d[a][0]);
            d1[a][1] = Dual<1>(d1d[a][1]);
          }
          acc += TH::template energy<Dual<1>>(tpd, g, s, f, d1).d[0] * g.gv;
        }
      }
      const double sum = block_sum<kThreads>(acc, red);
      if (threadIdx.x == 0) part[(long long)jj * L.ntiles + t] = sum;
    }
  }
}

// out[jj] = Kahan sum of the nt partials of parameter jj (one block of NT
// threads each; a template so that every translation unit may instantiate it).
template <int NT>
__global__ void __launch_bounds__(NT) param_grad_fin(const double* part, int nt, double* out) {
  __shared__ double red[NT / 64];
  const double* p = part + (long long)blockIdx.x * nt;
  double s = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < nt; i += NT) {
    const double y = p[i] - c;
    const double tt = s + y;
    c = (tt - s) - y;
    s = tt;
  }
  const double r = block_sum<NT>(s, red);
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}

// ---------------------------------------------------------------------------
// Exact adjoint of the discrete energy E_h = sum_s gv e(f_s, D1 f_s).
//
// dE_h/df_a(t) = gv de/df_a(t) + sum_i sum_k w_k P_ai(t - k e_i), with
// P_ai(s) = gv de/d(D1_i f_a)(s) / h_i where the site's D1_i is the stencil
// (0 in the band, where D1 is the constant 0) and w = (1/12, -2/3, 0, 2/3,
// -1/12) the first-derivative weights at offsets -2..2.  Pass A evaluates the
// local partials with one-direction dual numbers through the theory's own
// energy(); pass B applies the transposed stencil.

template <class TH>
__global__ void __launch_bounds__(kThreads)
adjoint_a(const double* Fv, double* G, double* Px, double* Py, typename TH::TP tp, Geom g, Launch L) {
  using S = TileShape<double, TH::NF>;
  constexpr int NF = TH::NF;
  __shared__ Tile<double, NF> tile;
  const int t = xcd_tile(blockIdx.x, L.ntiles, L.per_xcd);
  if (t < 0) return;
  const int tx = t / L.ntiles_y, ty = t - tx * L.ntiles_y;
  const int x0 = tx * S::TX, y0 = ty * TY;
  load_tile<double, NF>(tile, g, x0, y0, [&](int a, long long off) -> double { return Fv[off]; });
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int k = 0; k < S::kRowsPerWave; ++k) {
    const int lx = w * S::kRowsPerWave + k, ly = lane;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= g.X || y >= g.Y) continue;
    const SiteInfo s = site_info(g, x, y);
    double f[NF], d1[NF][2];
    tile_d1<double, double, NF>(tile, lx, ly, g, s, f, d1);
    const long long off = (long long)x * g.Y + y;
    for (int dir = 0; dir < 3 * NF; ++dir) {
      Dual<1> fd[NF], dd[NF][2];
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        fd[a] = Dual<1>(f[a]);
        dd[a][0] = Dual<1>(d1[a][0]);
        dd[a][1] = Dual<1>(d1[a][1]);
        if (dir == a) fd[a].d[0] = 1.0;
        if (dir == NF + a) dd[a][0].d[0] = 1.0;
        if (dir == 2 * NF + a) dd[a][1].d[0] = 1.0;
      }
      const double de = TH::template energy<Dual<1>>(tp, g, s, fd, dd).d[0] * g.gv;
      const int a = dir % NF;
      const long long o = off + (long long)a * g.N;
      if (dir < NF) G[o] = de;
      else if (dir < 2 * NF) Px[o] = s.inx ? de * g.inv_lsx : 0.0;
      else Py[o] = s.iny ? de * g.inv_lsy : 0.0;
    }
  }
}

// Pass B: G += transposed stencil of (Px, Py).  With ``flow`` the theory's
// flow convention is applied against the stage field Tv (tangent projection,
// saddle-field sign) and l = -dt EG is written to lout.
template <class TH, typename R>
__global__ void __launch_bounds__(kThreads)
adjoint_b(double* G, const double* Px, const double* Py, const void* Tv, void* lout, int flow,
          typename TH::TP tp, Geom g) {
  constexpr int NF = TH::NF;
  const long long n = g.N;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(i / g.Y), y = (int)(i - (long long)x * g.Y);
    double eg[NF];
#pragma unroll
    for (int a = 0; a < NF; ++a) {
      const double* px = Px + (long long)a * n;
      const double* py = Py + (long long)a * n;
      double acc = 0.0;
      if (x + 2 < g.X) acc += SS_C12 * px[i + 2 * (long long)g.Y];
      if (x + 1 < g.X) acc -= SS_C23 * px[i + g.Y];
      if (x - 1 >= 0) acc += SS_C23 * px[i - g.Y];
      if (x - 2 >= 0) acc -= SS_C12 * px[i - 2 * (long long)g.Y];
      if (y + 2 < g.Y) acc += SS_C12 * py[i + 2];
      if (y + 1 < g.Y) acc -= SS_C23 * py[i + 1];
      if (y - 1 >= 0) acc += SS_C23 * py[i - 1];
      if (y - 2 >= 0) acc -= SS_C12 * py[i - 2];
      eg[a] = G[i + (long long)a * n] + acc;
    }
    if (flow) {
      // flow units: EG = (1/gv) dE_h/dphi, the scale of the reference's variation
      const double igv = 1.0 / g.gv;
      const R* T = (const R*)Tv;
      double f[NF];
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        f[a] = (double)T[i + (long long)a * n];
        eg[a] *= igv;
      }
      TH::template flow_fix<double>(tp, f, eg);
      R* l = (R*)lout;
#pragma unroll
      for (int a = 0; a < NF; ++a) l[i + (long long)a * n] = R(-g.dt * eg[a]);
    }
#pragma unroll
    for (int a = 0; a < NF; ++a) G[i + (long long)a * n] = eg[a];
  }
}

// Unfused-flow helpers (used by the exact-adjoint and L2-normalised flows).
// Stage field Temp = F + ca dt (V + cb lprev), V read as 0 when killed.  With
// l2part non-null (256-thread blocks) each block also writes its partial sum
// of Temp^2 gv there, in the order of the separate partial-sum kernel of the
// L2 normalisation launched on the same grid.
template <typename R>
__global__ void stage_field(R* Temp, const R* F, const R* V, const R* lp, double ca, double cb, int use_kill,
                            const DevState* st, Geom g, double* l2part) {
  const long long n = (long long)g.nf * g.N;
  const bool kill = use_kill && st->kill;
  const R dt = R(g.dt), a = R(ca), b = R(cb);
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    R t;
    if (ca == 0.0) {
      t = F[i];
    } else {
      const R v = kill ? R(0) : V[i];
      t = F[i] + a * (dt * (v + (cb == 0.0 ? R(0) : b * lp[i])));
    }
    Temp[i] = t;
    if (l2part) s += (double)t * (double)t * g.gv;
  }
  if (l2part) {  // the same for every thread of the launch
    __shared__ double red[kWaves];
    const double bsum = block_sum<kThreads>(s, red);
    if (threadIdx.x == 0) l2part[blockIdx.x] = bsum;
  }
}

// RK4 combine from stored l1..l3 and the stage-4 gradient EG4 (FP64).
template <class TH, typename R>
__global__ void __launch_bounds__(kThreads)
combine_pass(const R* F, const R* V, const R* l1, const R* l2, const R* l3, const double* EG4, R* Fo, R* Vo,
             double* part0, double* part1, const DevState* st, Geom g, typename TH::TP tp) {
  constexpr int NF = TH::NF;
  __shared__ double red[kWaves];
  const bool kill = st->kill != 0;
  const R dt = R(g.dt), half = R(0.5), sixth = R(1.0 / 6.0);
  double m0 = -1.0e300, m1 = 0.0;
  ErrMax<double> em;
  const long long n = g.N;
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) {
    R Fn[NF], Vn[NF];
#pragma unroll
    for (int a = 0; a < NF; ++a) {
      const long long o = i + (long long)a * n;
      const R Vc = kill ? R(0) : V[o];
      const R a1 = l1[o], a2 = l2[o], a3 = l3[o];
      const double e = EG4[o];
      const R a4 = R(-g.dt * e);
      const R k1 = dt * Vc, k2 = dt * (Vc + half * a1), k3 = dt * (Vc + half * a2), k4 = dt * (Vc + a3);
      Vn[a] = Vc + sixth * (a1 + R(2) * a2 + R(2) * a3 + a4);
      Fn[a] = F[o] + sixth * (k1 + R(2) * k2 + R(2) * k3 + k4);
      em.add(e);
    }
    if (TH::kUnitConstraint && g.unit_mag) {
      TH::template normalize<R>(tp, Fn);
      TH::template project<R>(tp, Vn, Fn);
    }
#pragma unroll
    for (int a = 0; a < NF; ++a) {
      Fo[i + (long long)a * n] = Fn[a];
      Vo[i + (long long)a * n] = Vn[a];
    }
  }
  em.fold(m0, m1);
  const double b0 = block_max<kThreads>(m0, red);
  if (threadIdx.x == 0) part0[blockIdx.x] = b0;
  const double b1 = block_max<kThreads>(m1, red);
  if (threadIdx.x == 0) part1[blockIdx.x] = b1;
}

// ---------------------------------------------------------------------------
// Batched form: many independent lattices (own shape, spacing, time step and
// couplings) relaxed together.  Every pass is ONE launch over the tiles of all
// active instances; a global tile index t maps to (instance, local tile)
// through ``tile_inst``.  Per-instance arrest state, buffer parity and
// activity live on the device, so a captured K-step graph replays with static
// arguments and a converged instance simply stops (its blocks exit at once).

template <class TH>
struct BatchInst {
  Geom g;
  typename TH::TP tp;
  long long off;      // element offset of the instance inside every field buffer
  int tile_begin;     // first global tile index of the instance
  int ntiles;         // tiles of the instance
  int ntiles_y;       // tiles along y
  int pad;
};

struct BatchState {
  double prev_energy, last_energy, err_max, err_absmax;
  int kill, step, active, cur;
  int converged_step, pad[3];
};

struct BatchBufs {
  void* F[2];
  void* V[2];
  void* l[3];
  double* part0;
  double* part1;
  const int* tile_inst;
  const void* insts;   // BatchInst<TH>[ninst]
  BatchState* st;
  int ntiles, per_xcd;
  int sx;              // streaming shapes: rows per strip
};

// Output of pass PASS (1..3, or 23: l2 and l3) inside instance buffers at ``off``.
template <int PASS>
__device__ __forceinline__ void batch_outputs(PassBufs& b, const BatchBufs& B, long long off, size_t es) {
  b.lout = PASS == 23                ? (char*)B.l[1] + off * es
           : PASS == 12             ? (char*)B.l[0] + off * es
           : PASS >= 1 && PASS <= 3 ? (char*)B.l[PASS - 1] + off * es
                                    : nullptr;
  b.lout2 = PASS == 23 ? (char*)B.l[2] + off * es : PASS == 12 ? (char*)B.l[1] + off * es : nullptr;
}

template <class TH, typename R, int PASS, int TXR = 0>
__global__ void __launch_bounds__(kThreads) batch_stage_pass(BatchBufs B) {
  using S = TileShape<R, TH::NF, TXR>;
  __shared__ Tile<R, tile_nf<TH, PASS>(), TXR> tile;
  __shared__ double red[kWaves];
  const int t = xcd_tile(blockIdx.x, B.ntiles, B.per_xcd);
  if (t < 0) return;
  const int i = B.tile_inst[t];
  const BatchState& st = B.st[i];
  if (!st.active) return;  // uniform per block
  const BatchInst<TH>& I = ((const BatchInst<TH>*)B.insts)[i];
  const int lt = t - I.tile_begin;
  const int tx = lt / I.ntiles_y, ty = lt - tx * I.ntiles_y;
  const int cur = st.cur;
  const size_t es = sizeof(R);
  PassBufs b{};
  b.F = (const char*)B.F[cur] + I.off * es;
  b.V = (const char*)B.V[cur] + I.off * es;
  b.l1 = (const char*)B.l[0] + I.off * es;
  b.l2 = (const char*)B.l[1] + I.off * es;
  b.l3 = (const char*)B.l[2] + I.off * es;
  batch_outputs<PASS>(b, B, I.off, es);
  b.Fout = (char*)B.F[cur ^ 1] + I.off * es;
  b.Vout = (char*)B.V[cur ^ 1] + I.off * es;
  b.part0 = B.part0;
  b.part1 = B.part1;
  const bool kill = (PASS >= 2) ? (st.kill != 0) : false;
  // geometry and couplings copied out of the descriptor: as references into
  // global memory they would be re-read after every store of the pass (the
  // compiler cannot rule out that a store aliases them)
  const Geom g = I.g;
  const typename TH::TP tp = I.tp;
  stage_tile<TH, R, PASS, TXR>(tile, red, b, tp, g, tx * S::TX, ty * TY, kill, t);
}

// Streaming form of the batched passes: the table entries are strips of
// B.sx rows by the shape's width, numbered per instance row-band by
// row-band.  ENERGY: pass 1 on every instance, active or not, for its energy
// partials (its l1 output is scratch).
template <class TH, typename R, int PASS, int CW, bool ENERGY = false>
__global__ void __launch_bounds__(kThreads) batch_stream_pass(BatchBufs B) {
  __shared__ typename StreamRing<TH, R, PASS, CW>::type q;
  __shared__ double red[kWaves];
  constexpr int TYW = StreamRing<TH, R, PASS, CW>::shape::TYW;
  const int t = xcd_tile(blockIdx.x, B.ntiles, B.per_xcd);
  if (t < 0) return;
  const int i = B.tile_inst[t];
  const BatchState& st = B.st[i];
  if (!ENERGY && !st.active) return;  // uniform per block
  const BatchInst<TH>& I = ((const BatchInst<TH>*)B.insts)[i];
  const int lt = t - I.tile_begin;
  const int sxi = lt / I.ntiles_y, syi = lt - sxi * I.ntiles_y;
  const int cur = st.cur;
  const size_t es = sizeof(R);
  PassBufs b{};
  b.F = (const char*)B.F[cur] + I.off * es;
  b.V = (const char*)B.V[cur] + I.off * es;
  b.l1 = (const char*)B.l[0] + I.off * es;
  b.l2 = (const char*)B.l[1] + I.off * es;
  b.l3 = (const char*)B.l[2] + I.off * es;
  batch_outputs<PASS>(b, B, I.off, es);
  b.Fout = (char*)B.F[cur ^ 1] + I.off * es;
  b.Vout = (char*)B.V[cur ^ 1] + I.off * es;
  b.part0 = B.part0;
  b.part1 = B.part1;
  const bool kill = (PASS >= 2 && PASS != 12) ? (st.kill != 0) : false;  // 1+2 runs before the decision
  if (PASS == 34 && kill) b.l2 = b.l1;                                      // l2 of an arrested V
  const Geom g = I.g;  // copied out of the descriptor, as in batch_stage_pass
  const typename TH::TP tp = I.tp;
  stream_strip<TH, R, PASS, CW>(q, red, b, tp, g, sxi * B.sx, B.sx, syi * TYW, kill, t);
}

// Energy partials of every instance (active or not) from its current field.
template <class TH, typename R, int TXR = 0>
__global__ void __launch_bounds__(kThreads) batch_energy_pass(BatchBufs B) {
  using S = TileShape<R, TH::NF, TXR>;
  constexpr int NF = TH::NF;
  __shared__ Tile<R, NF, TXR> tile;
  __shared__ double red[kWaves];
  const int t = xcd_tile(blockIdx.x, B.ntiles, B.per_xcd);
  if (t < 0) return;
  const int i = B.tile_inst[t];
  const BatchInst<TH>& I = ((const BatchInst<TH>*)B.insts)[i];
  const Geom g = I.g;
  const typename TH::TP tp = I.tp;
  const int lt = t - I.tile_begin;
  const int tx = lt / I.ntiles_y, ty = lt - tx * I.ntiles_y;
  const int x0 = tx * S::TX, y0 = ty * TY;
  const R* __restrict__ F = (const R*)B.F[B.st[i].cur] + I.off;
  load_tile<R, NF>(tile, g, x0, y0, [&](int a, long long off) -> R { return F[off]; });
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double acc = 0.0;
  for (int k = 0; k < S::kRowsPerWave; ++k) {
    const int lx = w * S::kRowsPerWave + k, x = x0 + lx, y = y0 + lane;
    if (x < g.X && y < g.Y) {
      const SiteInfo s = site_info(g, x, y);
      double fd[NF], d1d[NF][2];
      tile_d1<double, R, NF>(tile, lx, lane, g, s, fd, d1d);
      acc += TH::template energy<double>(tp, g, s, fd, d1d) * g.gv;
    }
  }
  const double sum = block_sum<kThreads>(acc, red);
  if (threadIdx.x == 0) B.part0[t] = sum;
}

// Per-instance finalize of the energy (one block per instance).  ``report``
// 0: arrest decision of the previous step, exactly as the single-lattice
// k_fin_energy (active instances only); 1: only record the energy (every
// instance), leaving the arrest state untouched; 2: start of a relaxation
// (prev_energy = E, step 0, active, no kill).
template <class TH>
__global__ void __launch_bounds__(256) batch_fin_energy(BatchBufs B, int report) {
  __shared__ double red[4];
  const int i = blockIdx.x;
  BatchState& st = B.st[i];
  if (report == 0 && !st.active) return;
  const BatchInst<TH>& I = ((const BatchInst<TH>*)B.insts)[i];
  double s = 0.0, c = 0.0;
  for (int k = threadIdx.x; k < I.ntiles; k += 256) {
    const double y = B.part0[I.tile_begin + k] - c;
    const double tt = s + y;
    c = (tt - s) - y;
    s = tt;
  }
  const double E = block_sum<256>(s, red);
  if (threadIdx.x == 0) {
    st.last_energy = E;
    if (report == 2) {
      st.prev_energy = E;
      st.kill = 0;
      st.step = 0;
      st.active = 1;
      st.converged_step = -1;
    } else if (report == 0) {
      if (st.step > 0) {
        st.kill = (I.g.killkinen && E > st.prev_energy) ? 1 : 0;
        st.prev_energy = E;
      } else {
        st.kill = 0;
      }
    }
  }
}

// Per-instance finalize of the stage-4 error; flips the buffer parity and
// retires instances whose max |EG| fell below ``tol`` (converged_step = the
// step) or is not finite (converged_step = -2: failed, never converged).
template <class TH>
__global__ void __launch_bounds__(256) batch_fin_err(BatchBufs B, double tol) {
  __shared__ double red[4];
  const int i = blockIdx.x;
  BatchState& st = B.st[i];
  if (!st.active) return;
  const BatchInst<TH>& I = ((const BatchInst<TH>*)B.insts)[i];
  double m0 = -1.0e300, m1 = 0.0;
  for (int k = threadIdx.x; k < I.ntiles; k += 256) {
    m0 = nmax(m0, B.part0[I.tile_begin + k]);
    m1 = nmax(m1, B.part1[I.tile_begin + k]);
  }
  const double a = block_max<256>(m0, red);
  const double b = block_max<256>(m1, red);
  if (threadIdx.x == 0) {
    st.err_max = a;
    st.err_absmax = b;
    st.step += 1;
    st.cur ^= 1;
    if (b < tol) {
      st.active = 0;
      st.converged_step = st.step;
    } else if (!(b <= 1.7976931348623157e308)) {  // NaN or +inf: the instance has blown up
      st.active = 0;
      st.converged_step = -2;
    }
  }
}

// Initial condition of every site (theory init_site, FP64) stored in precision R.
template <class TH, typename R>
__global__ void __launch_bounds__(256) init_pass(void* Fv, InitArgs a, Geom g) {
  R* F = (R*)Fv;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < g.N; i += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(i / g.Y), y = (int)(i - (long long)x * g.Y);
    double f[TH::NF];
    TH::init_site(a, g, x, y, f);
#pragma unroll
    for (int c = 0; c < TH::NF; ++c) F[(long long)c * g.N + i] = R(f[c]);
  }
}

// Host-side launchers for one theory.  ``prec`` 0 = FP64 storage, 1 = FP32.
template <class TH>
struct TheoryLaunchers {
  using TP = typename TH::TP;

  // Shape variants of the fused passes: 0..2 tiles with the default rows, halved,
  // quartered (at least 4); 3..5 strips one wavefront (64 columns) wide and
  // 6..8 strips two wavefronts (128 columns) wide, of 64, 128, 256 rows
  // (streaming form).
  static constexpr int kNumShapes = 9;
  static int strip_rows(int v) { return 64 << ((v - 3) % 3); }
  static int strip_cw(int v) { return v >= 6 ? 2 : 1; }
  // Two-wavefront strips are offered when every streaming ring of the theory
  // that fits one wavefront wide also fits two wide.
  template <typename R>
  static constexpr bool wide_fits() {
    constexpr int NF = TH::NF;
    return StreamShape<R, NF, 6>::kFits && (!StreamShape<R, 2 * NF, 2>::kFits || StreamShape<R, 2 * NF, 6>::kFits) &&
           (!StreamShape<R, 2 * NF, 1>::kFits || StreamShape<R, 2 * NF, 5>::kFits);
  }
  // The streaming form addresses each array of nf planes through one buffer
  // resource with 32-bit byte offsets below 2^31; larger fields run the
  // default tile shape.
  template <typename R>
  static int effective_shape(int v, const Geom& g) {
    const bool ok = (long long)TH::NF * g.N * (long long)sizeof(R) < (1LL << 31);
    if (v >= 3 && !ok) return 0;
    if (v >= 6 && !wide_fits<R>()) return v - 3;
    return v;
  }

  // Partials (tiles or strips) of one pass launch in shape variant ``v``.
  template <typename R>
  static int ntiles_r(int v, const Geom& g) {
    constexpr int r0 = variant_rows<R, TH::NF>(0), r1 = variant_rows<R, TH::NF>(1), r2 = variant_rows<R, TH::NF>(2);
    v = effective_shape<R>(v, g);
    if (v >= 3) return make_slaunch<R, TH::NF>(g, strip_rows(v), TY * strip_cw(v)).nblocks;
    return v == 2 ? make_launch<R, TH::NF, r2>(g).ntiles
         : v == 1 ? make_launch<R, TH::NF, r1>(g).ntiles
                  : make_launch<R, TH::NF, r0>(g).ntiles;
  }

  // Rows per tile (variants 0..2) or per strip (3..8).
  static int shape_rows(int prec, int v) {
    if (v >= 3) return strip_rows(v);
    return prec ? variant_rows<float, TH::NF>(v) : variant_rows<double, TH::NF>(v);
  }

  template <typename R, int CW>
  static hipError_t stream_r(int pass, int sx, const PassBufs& b, const TP& tp, const Geom& g, hipStream_t st) {
    const SLaunch L = make_slaunch<R, TH::NF>(g, sx, TY * CW);
    dim3 grid(L.grid), block(kThreads);
    switch (pass) {
      case 0: hipLaunchKernelGGL((stream_pass<TH, R, 0, CW>), grid, block, 0, st, b, tp, g, L); break;
      case 1: hipLaunchKernelGGL((stream_pass<TH, R, 1, CW>), grid, block, 0, st, b, tp, g, L); break;
      case 2: hipLaunchKernelGGL((stream_pass<TH, R, 2, CW>), grid, block, 0, st, b, tp, g, L); break;
      case 3: hipLaunchKernelGGL((stream_pass<TH, R, 3, CW>), grid, block, 0, st, b, tp, g, L); break;
      case 4: hipLaunchKernelGGL((stream_pass<TH, R, 4, CW>), grid, block, 0, st, b, tp, g, L); break;
      case 5:
        if constexpr (StreamRing<TH, R, 23, CW>::shape::kFits) {
          hipLaunchKernelGGL((stream_pass<TH, R, 23, CW>), grid, block, 0, st, b, tp, g, L);
          break;
        } else {
          return hipErrorInvalidValue;
        }
      case 6:
        if constexpr (StreamRing<TH, R, 41, CW>::shape::kFits) {
          hipLaunchKernelGGL((stream_pass<TH, R, 41, CW>), grid, block, 0, st, b, tp, g, L);
          break;
        } else {
          return hipErrorInvalidValue;
        }
      case 7:
        if constexpr (sizeof(R) == 4) {
          hipLaunchKernelGGL((stream_pass_comp<TH, R, 4, CW>), grid, block, 0, st, b, tp, g, L);
          break;
        } else {
          return hipErrorInvalidValue;
        }
      case 8:
        if constexpr (sizeof(R) == 4 && StreamRing<TH, R, 1, CW, true>::shape::kFits) {
          hipLaunchKernelGGL((stream_pass_comp<TH, R, 1, CW>), grid, block, 0, st, b, tp, g, L);
          break;
        } else {
          return hipErrorInvalidValue;
        }
      case 9:
        if constexpr (StreamRing<TH, R, 12, CW>::shape::kFits) {
          hipLaunchKernelGGL((stream_pass<TH, R, 12, CW>), grid, block, 0, st, b, tp, g, L);
          break;
        } else {
          return hipErrorInvalidValue;
        }
      case 10:
        if constexpr (StreamRing<TH, R, 34, CW>::shape::kFits) {
          hipLaunchKernelGGL((stream_pass<TH, R, 34, CW>), grid, block, 0, st, b, tp, g, L);
          break;
        } else {
          return hipErrorInvalidValue;
        }
      case 11:
        if constexpr (sizeof(R) == 4 && StreamRing<TH, R, 34, CW, true>::shape::kFits) {
          hipLaunchKernelGGL((stream_pass_comp<TH, R, 34, CW>), grid, block, 0, st, b, tp, g, L);
          break;
        } else {
          return hipErrorInvalidValue;
        }
      case 12:
        if constexpr (sizeof(R) == 4 && StreamRing<TH, R, 12, CW, true>::shape::kFits) {
          hipLaunchKernelGGL((stream_pass_comp<TH, R, 12, CW>), grid, block, 0, st, b, tp, g, L);
          break;
        } else {
          return hipErrorInvalidValue;
        }
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  template <typename R>
  static hipError_t stream_v(int pass, int v, const PassBufs& b, const TP& tp, const Geom& g, hipStream_t st) {
    if constexpr (wide_fits<R>()) {
      if (strip_cw(v) == 2) return stream_r<R, 2>(pass, strip_rows(v), b, tp, g, st);
    }
    return stream_r<R, 1>(pass, strip_rows(v), b, tp, g, st);
  }

  // Passes 2 and 3 fused into one streaming sweep (pass id 5), and pass 4
  // with pass 1 of the next step (pass id 6), when both halves fit one ring
  // and the field uses 32-bit offsets.  Pass ids 7 and 8: passes 4 (any shape
  // variant) and 1 (streaming variants) of the compensated FP32 form, which
  // needs the streaming pass 1 (can_comp).
  static int can_fuse23(int prec, const Geom& g) {
    const bool fits = prec ? StreamRing<TH, float, 23>::shape::kFits : StreamRing<TH, double, 23>::shape::kFits;
    return fits && (prec ? effective_shape<float>(3, g) : effective_shape<double>(3, g)) == 3;
  }
  static int can_fuse41(int prec, const Geom& g) {
    const bool fits = prec ? StreamRing<TH, float, 41>::shape::kFits : StreamRing<TH, double, 41>::shape::kFits;
    return fits && (prec ? effective_shape<float>(3, g) : effective_shape<double>(3, g)) == 3 && !TH::kL2Normalize;
  }

  // Passes 2+3 as one sweep in shape variant v: strips where can_fuse23, tiles
  // (v < 3) where the two-field tile fits in LDS.
  template <typename R>
  static int fuse23_variant_r(int v, const Geom& g) {
    constexpr int r0 = variant_rows<R, TH::NF>(0), r1 = variant_rows<R, TH::NF>(1), r2 = variant_rows<R, TH::NF>(2);
    if (v >= 3) return effective_shape<R>(v, g) >= 3 && StreamRing<TH, R, 23>::shape::kFits;
    if (v == 2) return fused_tile_fits<TH, R, r2>();
    if (v == 1) return fused_tile_fits<TH, R, r1>();
    return fused_tile_fits<TH, R, r0>();
  }
  static int fuse23_variant(int prec, int v, const Geom& g) {
    return prec ? fuse23_variant_r<float>(v, g) : fuse23_variant_r<double>(v, g);
  }
  // Pass 4 + pass 1 as one sweep in shape variant v: strips where can_fuse41,
  // tiles (v < 3) where the two-field tile fits.
  template <typename R>
  static int fuse41_variant_r(int v, const Geom& g) {
    constexpr int r0 = variant_rows<R, TH::NF>(0), r1 = variant_rows<R, TH::NF>(1), r2 = variant_rows<R, TH::NF>(2);
    if (TH::kL2Normalize) return 0;
    if (v >= 3) return effective_shape<R>(v, g) >= 3 && StreamRing<TH, R, 41>::shape::kFits;
    if (v == 2) return fused_tile_fits<TH, R, r2>();
    if (v == 1) return fused_tile_fits<TH, R, r1>();
    return fused_tile_fits<TH, R, r0>();
  }
  static int fuse41_variant(int prec, int v, const Geom& g) {
    return prec ? fuse41_variant_r<float>(v, g) : fuse41_variant_r<double>(v, g);
  }

  // The paired schedule (pass ids 9 = passes 1+2, 10 = passes 3+4) in strip
  // variant v: both sweeps hold two stage fields in one ring.
  template <typename R, int CW>
  static constexpr bool pairs_fit() {
    return StreamRing<TH, R, 12, CW>::shape::kFits && StreamRing<TH, R, 34, CW>::shape::kFits;
  }
  template <typename R>
  static int pairs_variant_r(int v, const Geom& g) {
    if (TH::kL2Normalize || v < 3) return 0;
    v = effective_shape<R>(v, g);
    if (v < 3) return 0;
    if constexpr (wide_fits<R>()) {
      if (strip_cw(v) == 2) return pairs_fit<R, 2>();
    }
    return pairs_fit<R, 1>();
  }
  // The compensated form (prec 2): bit 0 = its 3+4 sweep (id 11) runs in v,
  // bit 1 = its 1+2 sweep (id 12, a three-part ring: F, Temp2 and the low
  // words) runs there too; the larger ring fitting implies the smaller.
  template <int CW>
  static constexpr int comp_pairs_fit() {
    return (StreamRing<TH, float, 34, CW, true>::shape::kFits ? 1 : 0) |
           (StreamRing<TH, float, 12, CW, true>::shape::kFits && StreamRing<TH, float, 34, CW, true>::shape::kFits
                ? 2
                : 0);
  }
  static int comp_pairs_variant(int v, const Geom& g) {
    if (TH::kL2Normalize || v < 3) return 0;
    v = effective_shape<float>(v, g);
    if (v < 3) return 0;
    if constexpr (wide_fits<float>()) {
      if (strip_cw(v) == 2) return comp_pairs_fit<2>();
    }
    return comp_pairs_fit<1>();
  }
  static int pairs_variant(int prec, int v, const Geom& g) {
    if (prec == 2) return comp_pairs_variant(v, g);
    return prec ? pairs_variant_r<float>(v, g) : pairs_variant_r<double>(v, g);
  }

  // Passes 2+3 as one sweep in the batched form, shape variant v (instances
  // address their buffers from their own offsets, so strips need no size check).
  template <typename R>
  static int batch_fuse23_r(int v) {
    constexpr int r0 = variant_rows<R, TH::NF>(0), r1 = variant_rows<R, TH::NF>(1), r2 = variant_rows<R, TH::NF>(2);
    if (v >= 3) return StreamRing<TH, R, 23>::shape::kFits;
    if (v == 2) return fused_tile_fits<TH, R, r2>();
    if (v == 1) return fused_tile_fits<TH, R, r1>();
    return fused_tile_fits<TH, R, r0>();
  }
  static int batch_fuse23(int prec, int v) { return prec ? batch_fuse23_r<float>(v) : batch_fuse23_r<double>(v); }
  // The paired sweeps (batch pass ids 7 = 1+2, 8 = 3+4) in strip variant v.
  template <typename R>
  static int batch_pairs_r(int v) {
    if (v < 3) return 0;
    if constexpr (wide_fits<R>()) {
      if (strip_cw(v) == 2) return pairs_fit<R, 2>();
    }
    return pairs_fit<R, 1>();
  }
  static int batch_pairs(int prec, int v) { return prec ? batch_pairs_r<float>(v) : batch_pairs_r<double>(v); }

  static int can_comp(int prec, const Geom& g) {
    return prec && StreamRing<TH, float, 1, 1, true>::shape::kFits && effective_shape<float>(3, g) == 3 &&
           !TH::kL2Normalize;
  }

  static int ntiles(int prec, int v, const Geom& g) { return prec ? ntiles_r<float>(v, g) : ntiles_r<double>(v, g); }

  template <typename R, int TXR>
  static hipError_t pass_rv(int pass, const PassBufs& b, const TP& tp, const Geom& g, hipStream_t st) {
    const Launch L = make_launch<R, TH::NF, TXR>(g);
    dim3 grid(L.grid), block(kThreads);
    switch (pass) {
      case 0: hipLaunchKernelGGL((stage_pass<TH, R, 0, TXR>), grid, block, 0, st, b, tp, g, L); break;
      case 1: hipLaunchKernelGGL((stage_pass<TH, R, 1, TXR>), grid, block, 0, st, b, tp, g, L); break;
      case 2: hipLaunchKernelGGL((stage_pass<TH, R, 2, TXR>), grid, block, 0, st, b, tp, g, L); break;
      case 3: hipLaunchKernelGGL((stage_pass<TH, R, 3, TXR>), grid, block, 0, st, b, tp, g, L); break;
      case 4: hipLaunchKernelGGL((stage_pass<TH, R, 4, TXR>), grid, block, 0, st, b, tp, g, L); break;
      case 5:
        if constexpr (fused_tile_fits<TH, R, TXR>()) {
          hipLaunchKernelGGL((stage_pass<TH, R, 23, TXR>), grid, block, 0, st, b, tp, g, L);
          break;
        } else {
          return hipErrorInvalidValue;
        }
      case 6:
        if constexpr (fused_tile_fits<TH, R, TXR>() && !TH::kL2Normalize) {
          hipLaunchKernelGGL((stage_pass<TH, R, 41, TXR>), grid, block, 0, st, b, tp, g, L);
          break;
        } else {
          return hipErrorInvalidValue;
        }
      case 7:
        if constexpr (sizeof(R) == 4) {
          hipLaunchKernelGGL((stage_pass<TH, R, 4, TXR, true>), grid, block, 0, st, b, tp, g, L);
          break;
        } else {
          return hipErrorInvalidValue;
        }
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }

  template <typename R>
  static hipError_t pass_r(int pass, int v, const PassBufs& b, const TP& tp, const Geom& g, hipStream_t st) {
    constexpr int r0 = variant_rows<R, TH::NF>(0), r1 = variant_rows<R, TH::NF>(1), r2 = variant_rows<R, TH::NF>(2);
    static_assert(64 % StreamShape<R, TH::NF>::RB == 0, "strip rows must be a multiple of the ring group");
    v = effective_shape<R>(v, g);
    if (pass >= 8) return v >= 3 ? stream_v<R>(pass, v, b, tp, g, st) : hipErrorInvalidValue;
    if (v >= 3) return stream_v<R>(pass, v, b, tp, g, st);
    if (v == 2 && r2 != r1) return pass_rv<R, r2>(pass, b, tp, g, st);
    if (v >= 1 && r1 != r0) return pass_rv<R, r1>(pass, b, tp, g, st);
    return pass_rv<R, r0>(pass, b, tp, g, st);
  }

  // ``v``: tile-shape variant; passes 0 and the energy/charge kernels use variant 0.
  static hipError_t pass(int prec, int pass, int v, const PassBufs& b, const void* tp, const Geom& g,
                         hipStream_t st) {
    return prec ? pass_r<float>(pass, v, b, *(const TP*)tp, g, st)
                : pass_r<double>(pass, v, b, *(const TP*)tp, g, st);
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

  // dE_h/dp_f[j] for j in [j0, j0 + nj) into out[0 .. nj) (FP64 field; the
  // theory must provide fill_tp).  part holds nj * param_grad_tiles(g) doubles.
  static int param_grad_tiles(const Geom& g) { return make_launch<double, TH::NF>(g).ntiles; }
  static int param_grad_count() {
    if constexpr (HasParamGrad<TH>::value) return TH::kNumPF;
    return 0;
  }
  static hipError_t param_grad(const double* F, double* part, double* out, const double* pf, int j0, int nj,
                               const Geom& g, hipStream_t st) {
    if constexpr (HasParamGrad<TH>::value) {
      if (j0 < 0 || nj < 1 || j0 + nj > TH::kNumPF) return hipErrorInvalidValue;
      const Launch L = make_launch<double, TH::NF>(g);
      PackedPF P{};
      for (int k = 0; k < TH::kNumPF; ++k) P.v[k] = pf[k];
      hipLaunchKernelGGL((param_grad_pass<TH>), dim3(L.grid), dim3(kThreads), 0, st, F, part, P, j0, nj, g, L);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL(param_grad_fin<256>, dim3(nj), dim3(256), 0, st, (const double*)part, L.ntiles, out);
      return hipGetLastError();
    } else {
      return hipErrorNotSupported;
    }
  }

  // Exact adjoint of the energy of Fv (FP64 storage) into G, scratch Px, Py
  // (nf*N doubles each).  With ``flow`` (storage precision ``prec``) also the
  // flow convention against Tv and l = -dt EG into lout.
  static hipError_t adjoint(int prec, const void* Fv, double* G, double* Px, double* Py, const void* Tv, void* lout,
                            int flow, const void* tpv, const Geom& g, hipStream_t st) {
    const TP& tp = *(const TP*)tpv;
    const Launch L = make_launch<double, TH::NF>(g);
    hipLaunchKernelGGL((adjoint_a<TH>), dim3(L.grid), dim3(kThreads), 0, st, (const double*)Fv, G, Px, Py, tp, g, L);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int nb = (int)((g.N + kThreads - 1) / kThreads);
    const dim3 grid(nb < 4096 ? nb : 4096);
    if (prec)
      hipLaunchKernelGGL((adjoint_b<TH, float>), grid, dim3(kThreads), 0, st, G, (const double*)Px,
                         (const double*)Py, Tv, lout, flow, tp, g);
    else
      hipLaunchKernelGGL((adjoint_b<TH, double>), grid, dim3(kThreads), 0, st, G, (const double*)Px,
                         (const double*)Py, Tv, lout, flow, tp, g);
    return hipGetLastError();
  }

  static hipError_t combine(int prec, const PassBufs& b, const double* EG4, const void* tpv, const Geom& g, hipStream_t st,
                            int* nblocks) {
    const int nb = (int)((g.N + kThreads - 1) / kThreads);
    *nblocks = nb;
    if (prec)
      hipLaunchKernelGGL((combine_pass<TH, float>), dim3(nb), dim3(kThreads), 0, st, (const float*)b.F,
                         (const float*)b.V, (const float*)b.l1, (const float*)b.l2, (const float*)b.l3, EG4,
                         (float*)b.Fout, (float*)b.Vout, b.part0, b.part1, b.st, g, *(const TP*)tpv);
    else
      hipLaunchKernelGGL((combine_pass<TH, double>), dim3(nb), dim3(kThreads), 0, st, (const double*)b.F,
                         (const double*)b.V, (const double*)b.l1, (const double*)b.l2, (const double*)b.l3, EG4,
                         (double*)b.Fout, (double*)b.Vout, b.part0, b.part1, b.st, g, *(const TP*)tpv);
    return hipGetLastError();
  }

  static int aux_components(int kind) { return TH::aux_components(kind); }

  static unsigned init_modes() { return TH::kInitModes; }

  static hipError_t init(int prec, const InitArgs& a, void* F, const Geom& g, hipStream_t st) {
    const long long nb = (g.N + 255) / 256;
    const dim3 grid((unsigned)(nb < 8192 ? nb : 8192));
    if (prec) hipLaunchKernelGGL((init_pass<TH, float>), grid, dim3(256), 0, st, F, a, g);
    else hipLaunchKernelGGL((init_pass<TH, double>), grid, dim3(256), 0, st, F, a, g);
    return hipGetLastError();
  }

  static int make_tp(const int32_t* pi, int ni, const double* pf, int nfl, void* tp, char* msg, int cap) {
    static_assert(sizeof(TP) <= 512, "theory parameter block too large");
    return TH::make(pi, ni, pf, nfl, (TP*)tp, msg, cap);
  }

  // -- batched form ----------------------------------------------------------
  static int batch_inst_size() { return (int)sizeof(BatchInst<TH>); }
  static int tile_rows(int prec) { return prec ? TileShape<float, TH::NF>::TX : TileShape<double, TH::NF>::TX; }

  static void batch_fill(void* dst, const Geom& g, const void* tp, long long off, int tile_begin, int ntiles,
                         int ntiles_y) {
    BatchInst<TH> I;
    memset(&I, 0, sizeof(I));
    I.g = g;
    I.tp = *(const TP*)tp;
    I.off = off;
    I.tile_begin = tile_begin;
    I.ntiles = ntiles;
    I.ntiles_y = ntiles_y;
    memcpy(dst, &I, sizeof(I));
  }

  template <typename R, int TXR>
  static hipError_t batch_pass_rv(int pass, const BatchBufs& B, hipStream_t st) {
    const dim3 grid(8 * B.per_xcd), block(kThreads);
    switch (pass) {
      case 1: hipLaunchKernelGGL((batch_stage_pass<TH, R, 1, TXR>), grid, block, 0, st, B); break;
      case 2: hipLaunchKernelGGL((batch_stage_pass<TH, R, 2, TXR>), grid, block, 0, st, B); break;
      case 3: hipLaunchKernelGGL((batch_stage_pass<TH, R, 3, TXR>), grid, block, 0, st, B); break;
      case 4: hipLaunchKernelGGL((batch_stage_pass<TH, R, 4, TXR>), grid, block, 0, st, B); break;
      case 5: hipLaunchKernelGGL((batch_energy_pass<TH, R, TXR>), grid, block, 0, st, B); break;
      case 6:
        if constexpr (fused_tile_fits<TH, R, TXR>()) {
          hipLaunchKernelGGL((batch_stage_pass<TH, R, 23, TXR>), grid, block, 0, st, B);
          break;
        } else {
          return hipErrorInvalidValue;
        }
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  template <typename R>
  static hipError_t batch_pass_r(int pass, int v, const BatchBufs& B, hipStream_t st) {
    constexpr int r0 = variant_rows<R, TH::NF>(0), r1 = variant_rows<R, TH::NF>(1), r2 = variant_rows<R, TH::NF>(2);
    if (v == 2 && r2 != r1) return batch_pass_rv<R, r2>(pass, B, st);
    if (v >= 1 && r1 != r0) return batch_pass_rv<R, r1>(pass, B, st);
    return batch_pass_rv<R, r0>(pass, B, st);
  }

  template <typename R, int CW>
  static hipError_t batch_stream_rv(int pass, const BatchBufs& B, hipStream_t st) {
    const dim3 grid(8 * B.per_xcd), block(kThreads);
    switch (pass) {
      case 1: hipLaunchKernelGGL((batch_stream_pass<TH, R, 1, CW>), grid, block, 0, st, B); break;
      case 2: hipLaunchKernelGGL((batch_stream_pass<TH, R, 2, CW>), grid, block, 0, st, B); break;
      case 3: hipLaunchKernelGGL((batch_stream_pass<TH, R, 3, CW>), grid, block, 0, st, B); break;
      case 4: hipLaunchKernelGGL((batch_stream_pass<TH, R, 4, CW>), grid, block, 0, st, B); break;
      case 5: hipLaunchKernelGGL((batch_stream_pass<TH, R, 1, CW, true>), grid, block, 0, st, B); break;
      case 6:
        if constexpr (StreamRing<TH, R, 23, CW>::shape::kFits) {
          hipLaunchKernelGGL((batch_stream_pass<TH, R, 23, CW>), grid, block, 0, st, B);
          break;
        } else {
          return hipErrorInvalidValue;
        }
      case 7:
        if constexpr (StreamRing<TH, R, 12, CW>::shape::kFits) {
          hipLaunchKernelGGL((batch_stream_pass<TH, R, 12, CW>), grid, block, 0, st, B);
          break;
        } else {
          return hipErrorInvalidValue;
        }
      case 8:
        if constexpr (StreamRing<TH, R, 34, CW>::shape::kFits) {
          hipLaunchKernelGGL((batch_stream_pass<TH, R, 34, CW>), grid, block, 0, st, B);
          break;
        } else {
          return hipErrorInvalidValue;
        }
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }

  // Strip rows and width (columns) of shape variant ``v`` in the batched
  // form, 0 rows for the tile variants; the width is the effective one (a
  // two-wavefront variant runs one wavefront wide where its rings do not fit).
  static void batch_shape(int prec, int v, int* sx, int* tyw) {
    if (v < 3) {
      *sx = 0;
      *tyw = TY;
      return;
    }
    const bool wide = prec ? wide_fits<float>() : wide_fits<double>();
    *sx = strip_rows(v);
    *tyw = (strip_cw(v) == 2 && wide) ? 2 * TY : TY;
  }

  // pass 1..4: fused RK4 passes; 5: energy partials of every instance; 6:
  // passes 2 and 3 in one sweep (where batch_fuse23); 7, 8: the paired sweeps
  // 1+2 and 3+4 (strips, where batch_pairs).  ``v``: shape variant
  // 0..8 (the tile table of the batch must be built for it).
  static hipError_t batch_pass(int prec, int pass, int v, const BatchBufs& B, hipStream_t st) {
    if (v >= 3) return prec ? batch_stream_v<float>(pass, v, B, st) : batch_stream_v<double>(pass, v, B, st);
    return prec ? batch_pass_r<float>(pass, v, B, st) : batch_pass_r<double>(pass, v, B, st);
  }
  template <typename R>
  static hipError_t batch_stream_v(int pass, int v, const BatchBufs& B, hipStream_t st) {
    if constexpr (wide_fits<R>()) {
      if (strip_cw(v) == 2) return batch_stream_rv<R, 2>(pass, B, st);
    }
    return batch_stream_rv<R, 1>(pass, B, st);
  }

  // which 0: energy finalize with ``mode`` (0 arrest, 1 report, 2 begin); 1: error finalize with ``tol``.
  static hipError_t batch_fin(int which, int ninst, const BatchBufs& B, int mode, double tol, hipStream_t st) {
    if (which == 0)
      hipLaunchKernelGGL((batch_fin_energy<TH>), dim3(ninst), dim3(256), 0, st, B, mode);
    else
      hipLaunchKernelGGL((batch_fin_err<TH>), dim3(ninst), dim3(256), 0, st, B, tol);
    return hipGetLastError();
  }
};

// Table of entry points one theory exposes to the context code.
struct TheoryOps {
  const char* name;
  int nf;
  int ncharges;
  int (*make_tp)(const int32_t*, int, const double*, int, void*, char*, int);
  int (*ntiles)(int, int, const Geom&);
  hipError_t (*pass)(int, int, int, const PassBufs&, const void*, const Geom&, hipStream_t);
  hipError_t (*density)(int, int, const void*, double*, double*, const void*, const Geom&, hipStream_t);
  hipError_t (*adjoint)(int, const void*, double*, double*, double*, const void*, void*, int, const void*,
                        const Geom&, hipStream_t);
  hipError_t (*combine)(int, const PassBufs&, const double*, const void*, const Geom&, hipStream_t, int*);
  int (*batch_inst_size)();
  int (*tile_rows)(int);
  int (*shape_rows)(int, int);
  int num_shapes;
  int (*can_fuse23)(int, const Geom&);
  int (*can_fuse41)(int, const Geom&);
  int (*can_comp)(int, const Geom&);
  int (*fuse23_variant)(int, int, const Geom&);
  int (*fuse41_variant)(int, int, const Geom&);
  int (*pairs_variant)(int, int, const Geom&);
  int (*batch_fuse23)(int, int);
  int (*batch_pairs)(int, int);
  int (*aux_components)(int);
  unsigned (*init_modes)();
  hipError_t (*init)(int, const InitArgs&, void*, const Geom&, hipStream_t);
  void (*batch_fill)(void*, const Geom&, const void*, long long, int, int, int);
  hipError_t (*batch_pass)(int, int, int, const BatchBufs&, hipStream_t);
  void (*batch_shape)(int, int, int*, int*);
  hipError_t (*batch_fin)(int, int, const BatchBufs&, int, double, hipStream_t);
  int (*param_grad_tiles)(const Geom&);
  int (*param_grad_count)();
  hipError_t (*param_grad)(const double*, double*, double*, const double*, int, int, const Geom&, hipStream_t);
  int l2_normalize;  // the flow L2-normalises stage fields (unfused normalised path)
};

#define SS_DEFINE_THEORY_OPS(TH, fn)                                                        \
  const ss::TheoryOps* fn() {                                                               \
    static const ss::TheoryOps ops = {TH::kName, TH::NF, TH::kNumCharges,                   \
                                      &ss::TheoryLaunchers<TH>::make_tp,                    \
                                      &ss::TheoryLaunchers<TH>::ntiles,                     \
                                      &ss::TheoryLaunchers<TH>::pass,                       \
                                      &ss::TheoryLaunchers<TH>::density,                    \
                                      &ss::TheoryLaunchers<TH>::adjoint,                    \
                                      &ss::TheoryLaunchers<TH>::combine,                    \
                                      &ss::TheoryLaunchers<TH>::batch_inst_size,            \
                                      &ss::TheoryLaunchers<TH>::tile_rows,                  \
                                      &ss::TheoryLaunchers<TH>::shape_rows,                 \
                                      ss::TheoryLaunchers<TH>::kNumShapes,                  \
                                      &ss::TheoryLaunchers<TH>::can_fuse23,                 \
                                      &ss::TheoryLaunchers<TH>::can_fuse41,                 \
                                      &ss::TheoryLaunchers<TH>::can_comp,                   \
                                      &ss::TheoryLaunchers<TH>::fuse23_variant,             \
                                      &ss::TheoryLaunchers<TH>::fuse41_variant,             \
                                      &ss::TheoryLaunchers<TH>::pairs_variant,              \
                                      &ss::TheoryLaunchers<TH>::batch_fuse23,               \
                                      &ss::TheoryLaunchers<TH>::batch_pairs,                \
                                      &ss::TheoryLaunchers<TH>::aux_components,             \
                                      &ss::TheoryLaunchers<TH>::init_modes,                 \
                                      &ss::TheoryLaunchers<TH>::init,                       \
                                      &ss::TheoryLaunchers<TH>::batch_fill,                 \
                                      &ss::TheoryLaunchers<TH>::batch_pass,                 \
                                      &ss::TheoryLaunchers<TH>::batch_shape,                \
                                      &ss::TheoryLaunchers<TH>::batch_fin,                  \
                                      &ss::TheoryLaunchers<TH>::param_grad_tiles,           \
                                      &ss::TheoryLaunchers<TH>::param_grad_count,           \
                                      &ss::TheoryLaunchers<TH>::param_grad,                 \
                                      TH::kL2Normalize ? 1 : 0};                            \
    return &ops;                                                                            \
  }

}  // namespace ss
