// Maxwell-Chern-Simons-Higgs model, the anyon superconductor (reference:
// theories/maxwell_chern_simons_higgs/kernels.py, params.py).
// Fields A1, A2, psi1, psi2, A0 (nf = 5); A0 is a saddle field.
#pragma once

#include "common.h"
#include "init.h"

namespace ss {

struct MaxwellChernSimonsHiggs {
  static constexpr int NF = 5;
  static constexpr int NMAG = 0;
  static constexpr bool kUnitConstraint = false;
  static constexpr bool kL2Normalize = false;
  static constexpr int kNumCharges = 3;  // 0 vortex number, 1 electric charge, 2 Noether charge
  static constexpr const char* kName = "maxwell_chern_simons_higgs";

  struct TP {
    double q, lam, kappa, u1;
  };

  // Packing: maxwell_chern_simons_higgs/params.py:173-207.
  static int make(const int32_t* pi, int ni, const double* pf, int nfl, TP* tp, char* msg, int cap) {
    if (ni < 12 || nfl < 12) {
      snprintf(msg, cap, "Maxwell-Chern-Simons-Higgs needs p_i[0..11] and p_f[0..11] (got %d, %d)", ni, nfl);
      return -1;
    }
    if (pi[4] != NF) {
      snprintf(msg, cap, "Maxwell-Chern-Simons-Higgs kernels are built for 5 fields (got %d)", pi[4]);
      return -1;
    }
    tp->q = pf[6];
    tp->lam = pf[7];
    tp->kappa = pf[8];
    tp->u1 = pf[9];
    return 0;
  }

  // kernels.py:480-567.
  template <typename T>
  __device__ __forceinline__ static void variation(const TP& tp, const Geom& g, const SiteInfo& s,
                                                   const T f[NF], const T d1[NF][2], const T d2[NF][3],
                                                   T eg[NF]) {
    const T q = T(tp.q), lam = T(tp.lam), k = T(tp.kappa), u1 = T(tp.u1);
    const T A1 = f[0], A2 = f[1], p1 = f[2], p2 = f[3], V = f[4];
    const T psq = p1 * p1 + p2 * p2;
    const T pot = lam / T(2) * (u1 * u1 - psq);
    const T mass = q * q * (A1 * A1 + A2 * A2 - V * V);
    const T divA = d1[0][0] + d1[1][1];
    eg[0] = -d2[0][1] + d2[1][2] + q * (p1 * d1[3][0] - p2 * d1[2][0]) + q * q * A1 * psq - k * d1[4][1];
    eg[1] = -d2[1][0] + d2[0][2] + q * (p1 * d1[3][1] - p2 * d1[2][1]) + q * q * A2 * psq + k * d1[4][0];
    eg[2] = -d2[2][0] - d2[2][1] + T(2) * q * (A1 * d1[3][0] + A2 * d1[3][1]) + q * p2 * divA + mass * p1 - pot * p1;
    eg[3] = -d2[3][0] - d2[3][1] - T(2) * q * (A1 * d1[2][0] + A2 * d1[2][1]) - q * p1 * divA + mass * p2 - pot * p2;
    eg[4] = -d2[4][0] - d2[4][1] + q * q * V * psq + k * (d1[1][0] - d1[0][1]);
  }

  // kernels.py:49-116, without the gv factor.
  template <typename T>
  __device__ __forceinline__ static T energy(const TP& tp, const Geom& g, const SiteInfo& s, const T f[NF],
                                             const T d1[NF][2]) {
    const T q = T(tp.q), lam = T(tp.lam), u1 = T(tp.u1);
    const T A1 = f[0], A2 = f[1], p1 = f[2], p2 = f[3], A0 = f[4];
    const T psq = p1 * p1 + p2 * p2;
    const T Asq = A1 * A1 + A2 * A2 + A0 * A0;
    T e = T(0.5) * (d1[2][0] * d1[2][0] + d1[2][1] * d1[2][1] + d1[3][0] * d1[3][0] + d1[3][1] * d1[3][1]);
    e += q * A1 * (p1 * d1[3][0] - p2 * d1[2][0]);
    e += q * A2 * (p1 * d1[3][1] - p2 * d1[2][1]);
    e += T(0.5) * q * q * Asq * psq;
    const T w = u1 * u1 - psq;
    e += lam / T(8) * w * w;
    const T curl = d1[1][0] - d1[0][1];
    e += T(0.5) * curl * curl;
    e += T(0.5) * (d1[4][0] * d1[4][0] + d1[4][1] * d1[4][1]);
    return e;
  }

  // 0: (d_x A2 - d_y A1)/(2 pi q); 1: -(lap A0); 2: q^2 A0 |psi|^2 (without gv).
  template <typename T>
  __device__ __forceinline__ static T charge(int kind, const TP& tp, const Geom& g, const SiteInfo& s,
                                             const T f[NF], const T d1[NF][2], const T d2[NF][3]) {
    if (kind == 0) return (d1[1][0] - d1[0][1]) / (T(2.0 * M_PI) * T(tp.q));
    if (kind == 1) return -(d2[4][0] + d2[4][1]);
    return T(tp.q) * T(tp.q) * f[4] * (f[2] * f[2] + f[3] * f[3]);
  }

  // Bit k set: init mode k (ground 0, initial 1, skyrmions 2, vortex 3) is available on the device.
  static constexpr unsigned kInitModes = 3u;
  // Device initial conditions (maxwell_chern_simons_higgs/initial_config.py):
  // ground psi = u1; initial vortex as the Ginzburg-Landau one, A0 = 0.
  __device__ __forceinline__ static bool init_site(const InitArgs& a, const Geom& g, int x, int y, double f[NF]) {
    const double u1 = a.pf[9], ainf = a.pf[11];
    for (int c = 0; c < NF; ++c) f[c] = 0.0;
    f[2] = u1;
    if (a.mode == kInitGround) return true;
    if (a.mode != kInitInitial) return false;
    const double rmax = a.pf[0] > a.pf[1] ? a.pf[0] : a.pf[1];
    double fs, sth, cth, rr, th;
    centred_vortex(g, x, y, 0.2, rmax, fs, sth, cth, rr, th);
    f[0] = ainf * fs * sth / rr;
    f[1] = -ainf * fs * cth / rr;
    f[2] = u1 * fs * cos(a.vortex_number * th);
    f[3] = u1 * fs * sin(a.vortex_number * th);
    return true;
  }

  // |psi|^2 (compute_norm_higgs_point, psi at fields 2, 3) and the 2-component
  // supercurrent (compute_supercurrent_point): (d_yx A2 - d_yy A1, d_xy A1 - d_xx A2).
  __host__ __device__ static constexpr int aux_components(int kind) {
    return kind == kAuxHiggs ? 1 : kind == kAuxSupercurrent ? 2 : 0;
  }
  template <typename T>
  __device__ __forceinline__ static T aux(int kind, int comp, const TP& tp, const Geom& g, const SiteInfo& s,
                                          const T f[NF], const T d1[NF][2], const T d2[NF][3]) {
    if (kind == kAuxHiggs) return f[2] * f[2] + f[3] * f[3];
    return comp == 0 ? d2[1][2] - d2[0][1] : d2[0][2] - d2[1][0];
  }

  // The energy as written (A0^2 |psi|^2 and |grad A0|^2 with + signs) is
  // convex in A0, so the exact-adjoint flow minimises over every field.
  template <typename T>
  __device__ __forceinline__ static void flow_fix(const TP& tp, const T f[NF], T eg[NF]) {}
  template <typename T>
  __device__ __forceinline__ static void normalize(const TP& tp, T f[NF]) {}
  template <typename T>
  __device__ __forceinline__ static void project(const TP& tp, T v[NF], const T f[NF]) {}
};

}  // namespace ss
