// Rotating Bose-Einstein condensate (reference: theories/bose_einstein_condensate/
// kernels.py, observables.py, params.py).  Fields psi_r, psi_i (nf = 2).
// The flow L2-normalises every stage field and the combined field
// (core/integrator.py:295-337), which the context runs through its unfused
// normalised path.
#pragma once

#include "common.h"
#include "init.h"

namespace ss {

struct BoseEinsteinCondensate {
  static constexpr int NF = 2;
  static constexpr int NMAG = 0;
  static constexpr bool kUnitConstraint = false;
  static constexpr bool kL2Normalize = true;
  static constexpr int kNumCharges = 1;  // 0: |psi|^2 (the norm density)
  static constexpr const char* kName = "bose_einstein_condensate";

  struct TP {
    double beta, omega;
  };

  // Packing: bose_einstein_condensate/params.py:198-229.
  static int make(const int32_t* pi, int ni, const double* pf, int nfl, TP* tp, char* msg, int cap) {
    if (ni < 11 || nfl < 10) {
      snprintf(msg, cap, "Bose-Einstein condensate needs p_i[0..10] and p_f[0..9] (got %d, %d)", ni, nfl);
      return -1;
    }
    if (pi[4] != NF) {
      snprintf(msg, cap, "Bose-Einstein condensate kernels are built for 2 fields (got %d)", pi[4]);
      return -1;
    }
    tp->beta = pf[6];
    tp->omega = pf[8];
    return 0;
  }

  // Index-based trap centre ((lsx (X-1))/2, (lsy (Y-1))/2) (kernels.py:74-80).
  template <typename T>
  __device__ __forceinline__ static void position(const Geom& g, const SiteInfo& s, T& xp, T& yp, T& r2) {
    const double xc = (0.0 + g.lsx * (double)(g.X - 1)) / 2.0;
    const double yc = (0.0 + g.lsy * (double)(g.Y - 1)) / 2.0;
    const double xd = g.lsx * (double)s.x - xc;
    const double yd = g.lsy * (double)s.y - yc;
    xp = T(xd);
    yp = T(yd);
    r2 = T(xd * xd + yd * yd);
  }

  // kernels.py:256-328.
  template <typename T>
  __device__ __forceinline__ static void variation(const TP& tp, const Geom& g, const SiteInfo& s,
                                                   const T f[NF], const T d1[NF][2], const T d2[NF][3],
                                                   T eg[NF]) {
    T xp, yp, r2;
    position<T>(g, s, xp, yp, r2);
    const T beta = T(tp.beta), om = T(tp.omega);
    const T psq = f[0] * f[0] + f[1] * f[1];
    eg[0] = f[0] * r2 + T(2) * beta * f[0] * psq - d2[0][0] - d2[0][1];
    eg[1] = f[1] * r2 + T(2) * beta * f[1] * psq - d2[1][0] - d2[1][1];
    eg[0] -= T(2) * om * (xp * d1[1][1] - yp * d1[1][0]);
    eg[1] += T(2) * om * (xp * d1[0][1] - yp * d1[0][0]);
  }

  // kernels.py:45-102, without the gv factor.
  template <typename T>
  __device__ __forceinline__ static T energy(const TP& tp, const Geom& g, const SiteInfo& s, const T f[NF],
                                             const T d1[NF][2]) {
    T xp, yp, r2;
    position<T>(g, s, xp, yp, r2);
    const T beta = T(tp.beta), om = T(tp.omega);
    const T psq = f[0] * f[0] + f[1] * f[1];
    T e = T(0.5) * (d1[0][0] * d1[0][0] + d1[0][1] * d1[0][1] + d1[1][0] * d1[1][0] + d1[1][1] * d1[1][1]);
    e += T(0.5) * r2 * psq;
    e += T(0.5) * beta * psq * psq;
    e -= om * (xp * (f[0] * d1[1][1] - f[1] * d1[0][1]) - yp * (f[0] * d1[1][0] - f[1] * d1[0][0]));
    return e;
  }

  // |psi|^2 (compute_norm_kernel, kernels.py:196-225, without gv).
  template <typename T>
  __device__ __forceinline__ static T charge(int kind, const TP& tp, const Geom& g, const SiteInfo& s,
                                             const T f[NF], const T d1[NF][2], const T d2[NF][3]) {
    return f[0] * f[0] + f[1] * f[1];
  }

  // Bit k set: init mode k (ground 0, initial 1, skyrmions 2, vortex 3) is available on the device.
  static constexpr unsigned kInitModes = 3u;
  // Device initial conditions (bose_einstein_condensate/initial_config.py):
  // Thomas-Fermi amplitude sqrt(max(1/sqrt(beta pi) - r^2/(2 beta), 0)) inside
  // radius p[0], times tanh(r_v/(sqrt 2 xi)) and the phase sign*atan2 of each
  // vortex seed in sol (x, y, -); p[1] = beta, p[2] = sign, p[3] = xi.
  __device__ __forceinline__ static bool init_site(const InitArgs& a, const Geom& g, int x, int y, double f[NF]) {
    if (a.mode != kInitGround && a.mode != kInitInitial) return false;
    double xc, yc;
    lattice_centre(g, xc, yc);
    const double gx = g.lsx * (double)x, gy = g.lsy * (double)y;
    const double dx0 = gx - xc, dy0 = gy - yc, rtf = sqrt(dx0 * dx0 + dy0 * dy0);
    const double beta = a.p[1];
    const double tf = (rtf <= a.p[0]) ? 1.0 / sqrt(beta * M_PI) - 1.0 / (2.0 * beta) * rtf * rtf : 0.0;
    const double amp = sqrt(fmax(tf, 0.0));
    double phase = 0.0, core = 1.0;
    for (int k = 0; k < a.nsol; ++k) {
      const double dx = gx - a.sol[3 * k], dy = gy - a.sol[3 * k + 1];
      phase = phase + a.p[2] * atan2(dy, dx);
      core = core * tanh(sqrt(dx * dx + dy * dy) / (sqrt(2.0) * a.p[3]));
    }
    f[0] = amp * core * cos(phase);
    f[1] = amp * core * sin(phase);
    return true;
  }

  // |psi|^2 (compute_norm_higgs_point, kernels.py).
  __host__ __device__ static constexpr int aux_components(int kind) { return kind == kAuxHiggs ? 1 : 0; }
  template <typename T>
  __device__ __forceinline__ static T aux(int kind, int comp, const TP& tp, const Geom& g, const SiteInfo& s,
                                          const T f[NF], const T d1[NF][2], const T d2[NF][3]) {
    return f[0] * f[0] + f[1] * f[1];
  }

  template <typename T>
  __device__ __forceinline__ static void flow_fix(const TP& tp, const T f[NF], T eg[NF]) {}
  template <typename T>
  __device__ __forceinline__ static void normalize(const TP& tp, T f[NF]) {}
  template <typename T>
  __device__ __forceinline__ static void project(const TP& tp, T v[NF], const T f[NF]) {}
};

}  // namespace ss
