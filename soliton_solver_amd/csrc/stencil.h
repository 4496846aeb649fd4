// LDS-staged halo tiles and the 4th-order derivative stencils.
//
// The stencils are the reference's (core/derivatives.py:13-115) evaluated in
// difference form, sum_k c_k (f_k - f_0), which is algebraically identical
// (the second-difference coefficients sum to zero) and loses less precision
// in FP32.  Masks follow the reference exactly:
//   d/dx, d2/dx2 are 0 unless halo-1 < x < X-halo (same for y);
//   the mixed entry is diag/(2 gv) - xx/(2 lsx^2) - yy/(2 lsy^2), computed
//   only for halo < x < X-halo and halo < y < Y-halo (strict lower bound),
//   with diag the 5-point second difference along the (1,1) diagonal.
#pragma once

#include "common.h"

namespace ss {

// Rows of x per tile.  TXR = 0 picks the largest of 32/16/8/4 rows whose
// tile (with halo) of every component fits in about 48 KiB of LDS; a nonzero
// TXR forces that many rows (a multiple of 4, for the tile-shape variants the
// autotuner times: fewer rows, less LDS, more resident workgroups per CU).
template <typename R, int NF>
struct DefaultRows {
  static constexpr int kBytesRow = NF * (TY + 4) * (int)sizeof(R);
  static constexpr int value = ((32 + 4) * kBytesRow <= 49152) ? 32
                             : ((16 + 4) * kBytesRow <= 49152) ? 16
                             : ((8 + 4) * kBytesRow <= 49152) ? 8
                             : 4;
};

// Rows of tile variant v (0, 1, 2): the default, then halved twice, at least 4.
template <typename R, int NF>
constexpr int variant_rows(int v) {
  return (DefaultRows<R, NF>::value >> v) < 4 ? 4 : (DefaultRows<R, NF>::value >> v);
}

template <typename R, int NF, int TXR = 0>
struct TileShape {
  static constexpr int kBytesRow = NF * (TY + 4) * (int)sizeof(R);
  static constexpr int TX = TXR ? TXR : DefaultRows<R, NF>::value;
  static_assert(TX % kWaves == 0, "tile rows must split evenly over the waves");
  static_assert((TX + 4) * kBytesRow <= 65536, "tile does not fit in 64 KiB of LDS");
  static constexpr int RX = TX + 4;  // rows with halo
  static constexpr int RY = TY + 4;  // columns with halo
  static constexpr int kRowsPerWave = TX / kWaves;
};

template <typename R, int NF, int TXR = 0>
struct Tile {
  using S = TileShape<R, NF, TXR>;
  R v[NF][S::RX][S::RY];
};

// Cooperative load of a (TX+4) x (TY+4) window per component into LDS.
// ``val(a, off)`` returns the stage-field value at flat offset ``off`` of
// component ``a``; sites outside the lattice are stored as 0 and are never
// used by an unmasked stencil.
template <typename R, int NF, int TXR, typename ValFn>
__device__ __forceinline__ void load_tile(Tile<R, NF, TXR>& t, const Geom& g, int x0, int y0, ValFn val) {
  using S = TileShape<R, NF, TXR>;
  constexpr int kElems = S::RX * S::RY;
  for (int idx = threadIdx.x; idx < kElems; idx += kThreads) {
    const int r = idx / S::RY;
    const int c = idx - r * S::RY;
    const int gx = x0 - 2 + r;
    const int gy = y0 - 2 + c;
    const bool in = (gx >= 0) && (gx < g.X) && (gy >= 0) && (gy < g.Y);
    const long long off = (long long)gx * g.Y + gy;
#pragma unroll
    for (int a = 0; a < NF; ++a) {
      t.v[a][r][c] = in ? val(a, off + (long long)a * g.N) : R(0);
    }
  }
}

// Derivatives of every component at interior tile position (lx, ly).
// T is the arithmetic type (R for the variation, double for energies).
template <typename T, typename R, int NF, int TXR>
__device__ __forceinline__ void tile_derivs(const Tile<R, NF, TXR>& t, int lx, int ly, const Geom& g,
                                            const SiteInfo& s, T f[NF], T d1[NF][2], T d2[NF][3]) {
  const int cx = lx + 2, cy = ly + 2;
  const T c12 = T(SS_C12), c23 = T(SS_C23), c43 = T(SS_C43);
  const T ilx = T(g.inv_lsx), ily = T(g.inv_lsy);
  const T ilx2 = T(g.inv_lsx2), ily2 = T(g.inv_lsy2);
  const T i2gv = T(g.inv_2gv), i2lx2 = T(g.inv_2lsx2), i2ly2 = T(g.inv_2lsy2);
#pragma unroll
  for (int a = 0; a < NF; ++a) {
    const T c = T(t.v[a][cx][cy]);
    const T xm2 = T(t.v[a][cx - 2][cy]), xm1 = T(t.v[a][cx - 1][cy]);
    const T xp1 = T(t.v[a][cx + 1][cy]), xp2 = T(t.v[a][cx + 2][cy]);
    const T ym2 = T(t.v[a][cx][cy - 2]), ym1 = T(t.v[a][cx][cy - 1]);
    const T yp1 = T(t.v[a][cx][cy + 1]), yp2 = T(t.v[a][cx][cy + 2]);
    const T sx1 = c12 * (xm2 - xp2) + c23 * (xp1 - xm1);
    const T sy1 = c12 * (ym2 - yp2) + c23 * (yp1 - ym1);
    const T sxx = c43 * ((xm1 - c) + (xp1 - c)) - c12 * ((xm2 - c) + (xp2 - c));
    const T syy = c43 * ((ym1 - c) + (yp1 - c)) - c12 * ((ym2 - c) + (yp2 - c));
    f[a] = c;
    d1[a][0] = s.inx ? sx1 * ilx : T(0);
    d1[a][1] = s.iny ? sy1 * ily : T(0);
    d2[a][0] = s.inx ? sxx * ilx2 : T(0);
    d2[a][1] = s.iny ? syy * ily2 : T(0);
    if (s.inxy) {
      const T dm2 = T(t.v[a][cx - 2][cy - 2]), dm1 = T(t.v[a][cx - 1][cy - 1]);
      const T dp1 = T(t.v[a][cx + 1][cy + 1]), dp2 = T(t.v[a][cx + 2][cy + 2]);
      const T sdd = c43 * ((dm1 - c) + (dp1 - c)) - c12 * ((dm2 - c) + (dp2 - c));
      d2[a][2] = sdd * i2gv - sxx * i2lx2 - syy * i2ly2;
    } else {
      d2[a][2] = T(0);
    }
  }
}

// Derivatives of one component from its x-window (x-2..x+2 at the site's
// column, held in registers by a thread marching along x) plus the y and
// diagonal neighbours read from LDS at tile row cx, column cy.
template <typename T, typename R, int NF, int TXR>
__device__ __forceinline__ void window_derivs(const Tile<R, NF, TXR>& t, int a, const R win[5], int cx, int cy,
                                              const Geom& g, const SiteInfo& s, T& f, T d1[2], T d2[3]) {
  const T c12 = T(SS_C12), c23 = T(SS_C23), c43 = T(SS_C43);
  const T c = T(win[2]);
  const T xm2 = T(win[0]), xm1 = T(win[1]), xp1 = T(win[3]), xp2 = T(win[4]);
  const T ym2 = T(t.v[a][cx][cy - 2]), ym1 = T(t.v[a][cx][cy - 1]);
  const T yp1 = T(t.v[a][cx][cy + 1]), yp2 = T(t.v[a][cx][cy + 2]);
  const T sxx = c43 * ((xm1 - c) + (xp1 - c)) - c12 * ((xm2 - c) + (xp2 - c));
  const T syy = c43 * ((ym1 - c) + (yp1 - c)) - c12 * ((ym2 - c) + (yp2 - c));
  f = c;
  d1[0] = s.inx ? (c12 * (xm2 - xp2) + c23 * (xp1 - xm1)) * T(g.inv_lsx) : T(0);
  d1[1] = s.iny ? (c12 * (ym2 - yp2) + c23 * (yp1 - ym1)) * T(g.inv_lsy) : T(0);
  d2[0] = s.inx ? sxx * T(g.inv_lsx2) : T(0);
  d2[1] = s.iny ? syy * T(g.inv_lsy2) : T(0);
  if (s.inxy) {
    const T dm2 = T(t.v[a][cx - 2][cy - 2]), dm1 = T(t.v[a][cx - 1][cy - 1]);
    const T dp1 = T(t.v[a][cx + 1][cy + 1]), dp2 = T(t.v[a][cx + 2][cy + 2]);
    const T sdd = c43 * ((dm1 - c) + (dp1 - c)) - c12 * ((dm2 - c) + (dp2 - c));
    d2[2] = sdd * T(g.inv_2gv) - sxx * T(g.inv_2lsx2) - syy * T(g.inv_2lsy2);
  } else {
    d2[2] = T(0);
  }
}

// First derivatives of one component from the same window, in type T.
template <typename T, typename R, int NF, int TXR>
__device__ __forceinline__ void window_d1(const Tile<R, NF, TXR>& t, int a, const R win[5], int cx, int cy,
                                          const Geom& g, const SiteInfo& s, T& f, T d1[2]) {
  const T c12 = T(SS_C12), c23 = T(SS_C23);
  const T xm2 = T(win[0]), xm1 = T(win[1]), xp1 = T(win[3]), xp2 = T(win[4]);
  const T ym2 = T(t.v[a][cx][cy - 2]), ym1 = T(t.v[a][cx][cy - 1]);
  const T yp1 = T(t.v[a][cx][cy + 1]), yp2 = T(t.v[a][cx][cy + 2]);
  f = T(win[2]);
  d1[0] = s.inx ? (c12 * (xm2 - xp2) + c23 * (xp1 - xm1)) * T(g.inv_lsx) : T(0);
  d1[1] = s.iny ? (c12 * (ym2 - yp2) + c23 * (yp1 - ym1)) * T(g.inv_lsy) : T(0);
}

// First derivatives only (energies and charges).
template <typename T, typename R, int NF, int TXR>
__device__ __forceinline__ void tile_d1(const Tile<R, NF, TXR>& t, int lx, int ly, const Geom& g,
                                        const SiteInfo& s, T f[NF], T d1[NF][2]) {
  const int cx = lx + 2, cy = ly + 2;
  const T c12 = T(SS_C12), c23 = T(SS_C23);
  const T ilx = T(g.inv_lsx), ily = T(g.inv_lsy);
#pragma unroll
  for (int a = 0; a < NF; ++a) {
    const T xm2 = T(t.v[a][cx - 2][cy]), xm1 = T(t.v[a][cx - 1][cy]);
    const T xp1 = T(t.v[a][cx + 1][cy]), xp2 = T(t.v[a][cx + 2][cy]);
    const T ym2 = T(t.v[a][cx][cy - 2]), ym1 = T(t.v[a][cx][cy - 1]);
    const T yp1 = T(t.v[a][cx][cy + 1]), yp2 = T(t.v[a][cx][cy + 2]);
    f[a] = T(t.v[a][cx][cy]);
    d1[a][0] = s.inx ? (c12 * (xm2 - xp2) + c23 * (xp1 - xm1)) * ilx : T(0);
    d1[a][1] = s.iny ? (c12 * (ym2 - yp2) + c23 * (yp1 - ym1)) * ily : T(0);
  }
}

}  // namespace ss
