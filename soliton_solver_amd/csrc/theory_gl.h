// Ginzburg-Landau superconductor / Abelian Higgs model
// (reference: theories/ginzburg_landau_superconductor/kernels.py, params.py).
// Fields A1, A2, psi1, psi2 (nf = 4).
#pragma once

#include "common.h"
#include "init.h"

namespace ss {

struct GinzburgLandau {
  static constexpr int NF = 4;
  static constexpr int NMAG = 0;
  static constexpr bool kUnitConstraint = false;  // no do_rk4_kernel: unconstrained combine
  static constexpr bool kL2Normalize = false;
  static constexpr int kNumCharges = 1;           // 0 vortex number
  static constexpr const char* kName = "ginzburg_landau";

  struct TP {
    double q, lam;
    double u_var;   // p_f[8] (u1), read by the variation (kernels.py:379)
    double u_en;    // p_f[9] (vortex_number), read by the energy (kernels.py:76)
  };

  // Packing: ginzburg_landau_superconductor/params.py:161-194.
  static int make(const int32_t* pi, int ni, const double* pf, int nfl, TP* tp, char* msg, int cap) {
    if (ni < 12 || nfl < 11) {
      snprintf(msg, cap, "Ginzburg-Landau needs p_i[0..11] and p_f[0..10] (got %d, %d)", ni, nfl);
      return -1;
    }
    if (pi[4] != NF) {
      snprintf(msg, cap, "Ginzburg-Landau kernels are built for number_total_fields=4 (got %d)", pi[4]);
      return -1;
    }
    tp->q = pf[6];
    tp->lam = pf[7];
    tp->u_var = pf[8];
    tp->u_en = pf[9];
    return 0;
  }

  // kernels.py:342-419 (g0..g3 at :407-410); no band mask, no projection.
  template <typename T>
  __device__ __forceinline__ static void variation(const TP& tp, const Geom& g, const SiteInfo& s,
                                                   const T f[NF], const T d1[NF][2], const T d2[NF][3],
                                                   T eg[NF]) {
    const T q = T(tp.q), lam = T(tp.lam), u1 = T(tp.u_var);
    const T A1 = f[0], A2 = f[1], p1 = f[2], p2 = f[3];
    const T psq = p1 * p1 + p2 * p2;
    const T Asq = A1 * A1 + A2 * A2;
    const T pot = lam / T(2) * (u1 * u1 - psq);
    const T divA = d1[0][0] + d1[1][1];
    eg[0] = -d2[0][1] + d2[1][2] + q * (p1 * d1[3][0] - p2 * d1[2][0]) + q * q * A1 * psq;
    eg[1] = -d2[1][0] + d2[0][2] + q * (p1 * d1[3][1] - p2 * d1[2][1]) + q * q * A2 * psq;
    eg[2] = -d2[2][0] - d2[2][1] + T(2) * q * (A1 * d1[3][0] + A2 * d1[3][1]) + q * p2 * divA
            + q * q * Asq * p1 - pot * p1;
    eg[3] = -d2[3][0] - d2[3][1] - T(2) * q * (A1 * d1[2][0] + A2 * d1[2][1]) - q * p1 * divA
            + q * q * Asq * p2 - pot * p2;
  }

  // kernels.py:45-108, without the gv factor.
  template <typename T>
  __device__ __forceinline__ static T energy(const TP& tp, const Geom& g, const SiteInfo& s,
                                             const T f[NF], const T d1[NF][2]) {
    const T q = T(tp.q), lam = T(tp.lam), u1 = T(tp.u_en);
    const T A1 = f[0], A2 = f[1], p1 = f[2], p2 = f[3];
    const T psq = p1 * p1 + p2 * p2;
    const T Asq = A1 * A1 + A2 * A2;
    T e = T(0.5) * (d1[2][0] * d1[2][0] + d1[2][1] * d1[2][1] + d1[3][0] * d1[3][0] + d1[3][1] * d1[3][1]);
    e += q * A1 * (p1 * d1[3][0] - p2 * d1[2][0]);
    e += q * A2 * (p1 * d1[3][1] - p2 * d1[2][1]);
    e += T(0.5) * q * q * Asq * psq;
    const T w = u1 * u1 - psq;
    e += lam / T(8) * w * w;
    const T curl = d1[1][0] - d1[0][1];
    e += T(0.5) * curl * curl;
    return e;
  }

  // Vortex number density (d_x A2 - d_y A1)/(2 pi q) (kernels.py:145-178).
  template <typename T>
  __device__ __forceinline__ static T charge(int kind, const TP& tp, const Geom& g, const SiteInfo& s,
                                             const T f[NF], const T d1[NF][2], const T d2[NF][3]) {
    return (d1[1][0] - d1[0][1]) / (T(2.0 * M_PI) * T(tp.q));
  }

  // Bit k set: init mode k (ground 0, initial 1, skyrmions 2, vortex 3) is available on the device.
  static constexpr unsigned kInitModes = 11u;
  // Device initial conditions (ginzburg_landau_superconductor/initial_config.py):
  // ground (A = 0, psi = u1); vortex psi = u1 tanh(scale r) e^{i N th}, A = 0;
  // initial: profile tanh(0.2 r) (1 beyond max(xsize, ysize)), A = ainf fs (sin, -cos)/r.
  __device__ __forceinline__ static bool init_site(const InitArgs& a, const Geom& g, int x, int y, double f[NF]) {
    const double u1 = a.pf[8], ainf = a.pf[10];
    f[0] = f[1] = f[3] = 0.0;
    f[2] = u1;
    if (a.mode == kInitGround) return true;
    if (a.mode != kInitVortex && a.mode != kInitInitial) return false;
    const double rmax = a.pf[0] > a.pf[1] ? a.pf[0] : a.pf[1];
    double fs, sth, cth, rr, th;
    if (a.mode == kInitVortex) {
      centred_vortex(g, x, y, a.scale, 1e300, fs, sth, cth, rr, th);
    } else {
      centred_vortex(g, x, y, 0.2, rmax, fs, sth, cth, rr, th);
      f[0] = ainf * fs * sth / rr;
      f[1] = -ainf * fs * cth / rr;
    }
    f[2] = u1 * fs * cos(a.vortex_number * th);
    f[3] = u1 * fs * sin(a.vortex_number * th);
    return true;
  }

  // |psi|^2 (compute_norm_higgs_point) and the 2-component supercurrent
  // (compute_supercurrent_point): (d_yx A2 - d_yy A1, d_xy A1 - d_xx A2).
  __host__ __device__ static constexpr int aux_components(int kind) {
    return kind == kAuxHiggs ? 1 : kind == kAuxSupercurrent ? 2 : 0;
  }
  template <typename T>
  __device__ __forceinline__ static T aux(int kind, int comp, const TP& tp, const Geom& g, const SiteInfo& s,
                                          const T f[NF], const T d1[NF][2], const T d2[NF][3]) {
    if (kind == kAuxHiggs) return f[2] * f[2] + f[3] * f[3];
    return comp == 0 ? d2[1][2] - d2[0][1] : d2[0][2] - d2[1][0];
  }

  template <typename T>
  __device__ __forceinline__ static void flow_fix(const TP& tp, const T f[NF], T eg[NF]) {}
  template <typename T>
  __device__ __forceinline__ static void normalize(const TP& tp, T f[NF]) {}
  template <typename T>
  __device__ __forceinline__ static void project(const TP& tp, T v[NF], const T f[NF]) {}
};

}  // namespace ss
