// Rotating Bose-Einstein condensate (reference: theories/bose_einstein_condensate/
// kernels.py, observables.py, params.py).  Fields psi_r, psi_i (nf = 2).
// The flow L2-normalises every stage field and the combined field
// (core/integrator.py:295-337), which the context runs through its unfused
// normalised path.
#pragma once

#include "common.h"

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
