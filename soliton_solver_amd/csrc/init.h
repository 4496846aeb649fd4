// Initial conditions evaluated on the device.
//
// The formulas are the reference's initial-condition kernels (e.g.
// chiral_magnet/initial_config.py:48-386, ginzburg_landau_superconductor/
// initial_config.py:114-313): radial profiles, the polar angle about a centre
// shifted to [0, 2 pi), Bloch / Neel / anti textures, and the composition of a
// texture into an existing field by adding stereographic coordinates.  Each
// theory's ``init_site`` picks what its host generator does for a mode, so a
// device-initialised field agrees with the host one to rounding.
#pragma once

#include "common.h"

namespace ss {

enum InitMode { kInitGround = 0, kInitInitial = 1, kInitSkyrmions = 2, kInitVortex = 3 };

struct InitArgs {
  int mode;
  int ansatz;            // 0 Bloch, 1 Neel, 2 anti
  double skN;            // winding of the texture
  double rot;            // rotation of the polar angle
  double scale;          // vortex mode: tanh(scale r)
  double vortex_number;  // winding of the order parameter
  double p[4];           // theory-specific extras (BEC: Thomas-Fermi radius, beta, sign, core size)
  int nsol;              // solitons / vortex seeds in sol
  const double* sol;     // device array, 3 values per soliton (x, y, radius)
  double pf[kMaxPF];     // the packed p_f
};

__device__ __forceinline__ void lattice_centre(const Geom& g, double& xc, double& yc) {
  xc = (0.0 + g.lsx * (double)(g.X - 1)) / 2.0;
  yc = (0.0 + g.lsy * (double)(g.Y - 1)) / 2.0;
}

// r and the angle atan2 in [0, 2 pi) minus ``rot``; at r == 0 the angle is -rot.
__device__ __forceinline__ void polar(const Geom& g, int x, int y, double xc, double yc, double rot, double& r,
                                      double& th) {
  const double gx = g.lsx * (double)x - xc, gy = g.lsy * (double)y - yc;
  r = sqrt(gx * gx + gy * gy);
  double ang = atan2(gy, gx);
  if (ang < 0.0) ang += 2.0 * M_PI;
  th = (r == 0.0) ? -rot : -rot + ang;
}

// pi exp(-(2 r / max_r)^2) for r <= max_r, else 0.
__device__ __forceinline__ double texture_profile(double r, double max_r) {
  if (r > max_r) return 0.0;
  const double t = 2.0 * r / max_r;
  return M_PI * exp(-(t * t));
}

__device__ __forceinline__ void texture(double fm, double th, double skN, int kind, double m[3]) {
  const double s = sin(fm), c = cos(fm);
  if (kind == 0) {
    m[0] = -s * sin(skN * th);
    m[1] = s * cos(skN * th);
  } else if (kind == 1) {
    m[0] = s * cos(skN * th);
    m[1] = s * sin(skN * th);
  } else {
    m[0] = -s * sin(skN * th);
    m[1] = -s * cos(skN * th);
  }
  m[2] = c;
}

// |m| = target (sites with |m| == 0 unchanged).
__device__ __forceinline__ void rescale3(double m[3], double target) {
  const double s = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  if (s != 0.0) {
    m[0] = m[0] * (target / s);
    m[1] = m[1] * (target / s);
    m[2] = m[2] * (target / s);
  }
}

// One centred texture of radius scale max_r (then unit norm).
__device__ __forceinline__ void single_texture(const Geom& g, int x, int y, double max_r, double rot, double skN,
                                               int kind, double m[3]) {
  double xc, yc, r, th;
  lattice_centre(g, xc, yc);
  polar(g, x, y, xc, yc, rot, r, th);
  texture(texture_profile(r, max_r), th, skN, kind, m);
  const double s = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  if (s != 0.0) {
    m[0] /= s;
    m[1] /= s;
    m[2] /= s;
  }
}

// Compose the solitons (x, y, radius) into unit vector m by stereographic
// addition w = (m1 + i m2)/(1 + m3) (denominators clamped at 1e-12), each with
// profile radius 2 radius; sites outside a soliton's disc are unchanged.
__device__ __forceinline__ void compose_textures(const Geom& g, int x, int y, const InitArgs& a, double m[3]) {
  for (int k = 0; k < a.nsol; ++k) {
    const double xc = a.sol[3 * k], yc = a.sol[3 * k + 1], max_r = 2.0 * a.sol[3 * k + 2];
    double r, th;
    polar(g, x, y, xc, yc, a.rot, r, th);
    if (r > max_r) continue;
    double q[3];
    if (a.skN == 0.0) texture(2.0 * texture_profile(r, max_r), th, 1.0, a.ansatz, q);
    else texture(texture_profile(r, max_r), th, a.skN, a.ansatz, q);
    const double den = fmax(1.0 + m[2], 1e-12);
    const double wr = m[0] / den + q[0] / fmax(1.0 + q[2], 1e-12);
    const double wi = m[1] / den + q[1] / fmax(1.0 + q[2], 1e-12);
    const double nw = wr * wr + wi * wi;
    m[0] = 2.0 * wr / (1.0 + nw);
    m[1] = 2.0 * wi / (1.0 + nw);
    m[2] = (1.0 - nw) / (1.0 + nw);
    rescale3(m, 1.0);
  }
}

// Unit-vector theories (chiral magnet, baby Skyrme, liquid crystal): ground
// e3, one centred texture of radius max(xsize, ysize)/10 * skN, or composed
// solitons; components past the third are 0.
template <int NF>
__device__ __forceinline__ bool unit_vector_init(const InitArgs& a, const Geom& g, int x, int y, double f[NF]) {
  for (int c = 0; c < NF; ++c) f[c] = 0.0;
  f[2] = 1.0;
  if (a.mode == kInitGround) return true;
  const double rmax = a.pf[0] > a.pf[1] ? a.pf[0] : a.pf[1];
  double m[3] = {0.0, 0.0, 1.0};
  if (a.mode == kInitInitial) single_texture(g, x, y, rmax / 10.0 * a.skN, a.rot, a.skN, a.ansatz, m);
  else if (a.mode == kInitSkyrmions) compose_textures(g, x, y, a, m);
  else return false;
  f[0] = m[0];
  f[1] = m[1];
  f[2] = m[2];
  return true;
}

// Vortex of winding N about the lattice centre: A = ainf fs (s sin th, -s cos th)/max(r, 1e-12)
// with s = asign, psi = u fs e^{i N th}; the angle has no rotation.
__device__ __forceinline__ void centred_vortex(const Geom& g, int x, int y, double fs_scale, double rcut, double& fs,
                                               double& sth, double& cth, double& rr, double& th) {
  double xc, yc, r;
  lattice_centre(g, xc, yc);
  polar(g, x, y, xc, yc, 0.0, r, th);
  if (r == 0.0) th = 0.0;
  fs = (r > rcut) ? 1.0 : tanh(fs_scale * r);
  sth = sin(th);
  cth = cos(th);
  rr = fmax(r, 1e-12);
}

}  // namespace ss
