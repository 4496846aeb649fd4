// Spin-triplet superconducting ferromagnet (reference: theories/
// spin_triplet_superconducting_magnet/kernels.py, params.py).  Fields m1, m2, m3,
// A1, A2, A3, psi1 (re, im), psi2 (re, im) (nf = 10); |m| = M0.
#pragma once

#include "common.h"
#include "init.h"

namespace ss {

struct SpinTripletSFM {
  static constexpr int NF = 10;
  static constexpr int NMAG = 3;
  static constexpr bool kUnitConstraint = true;
  static constexpr bool kL2Normalize = false;
  static constexpr int kNumCharges = 4;  // 0 skyrmion, 1 vortex (which=2), 2 which=0, 3 which=1
  static constexpr const char* kName = "spin_triplet_superconducting_magnet";

  struct TP {
    double q, ha, hb1, hb2, hc, alpha, beta, gamma, M0, vac;
  };

  // Packing: spin_triplet_superconducting_magnet/params.py, pack_device_params.
  static int make(const int32_t* pi, int ni, const double* pf, int nfl, TP* tp, char* msg, int cap) {
    if (ni < 13 || nfl < 27) {
      snprintf(msg, cap, "spin-triplet ferromagnet needs p_i[0..12] and p_f[0..26] (got %d, %d)", ni, nfl);
      return -1;
    }
    if (pi[4] != NF || pi[10] != NMAG) {
      snprintf(msg, cap, "spin-triplet ferromagnet kernels are built for 10 fields, 3 magnetisation (got %d, %d)",
               pi[4], pi[10]);
      return -1;
    }
    tp->q = pf[6];
    tp->ha = pf[7];
    tp->hb1 = pf[8];
    tp->hb2 = pf[9];
    tp->hc = pf[10];
    tp->alpha = pf[17];
    tp->beta = pf[18];
    tp->gamma = pf[19];
    tp->M0 = pf[20];
    if (tp->M0 == 0.0) {
      snprintf(msg, cap, "spin-triplet ferromagnet needs M0 != 0");
      return -1;
    }
    const double u1 = pf[11], u2 = pf[12], M0 = tp->M0;
    // subtracted vacuum constant (kernels.py:189)
    tp->vac = tp->ha / 2.0 * (u1 * u1 + u2 * u2) + tp->hb1 / 4.0 * (u1 * u1 * u1 * u1 + u2 * u2 * u2 * u2) +
              tp->hb2 * u1 * u1 * u2 * u2 + 2.0 * tp->hc * u1 * u2 + tp->alpha / 2.0 * M0 * M0 +
              tp->beta / 4.0 * M0 * M0 * M0 * M0;
    return 0;
  }

  // kernels.py:601-690.
  template <typename T>
  __device__ __forceinline__ static void variation(const TP& tp, const Geom& g, const SiteInfo& s,
                                                   const T f[NF], const T d1[NF][2], const T d2[NF][3],
                                                   T eg[NF]) {
    const T q = T(tp.q), ha = T(tp.ha), hb1 = T(tp.hb1), hb2 = T(tp.hb2), hc = T(tp.hc);
    const T al = T(tp.alpha), be = T(tp.beta), ga2 = T(tp.gamma) * T(tp.gamma);
    const T A1 = f[3], A2 = f[4], A3 = f[5], p1 = f[6], p2 = f[7], p3 = f[8], p4 = f[9];
    const T msq = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
    const T Asq = A1 * A1 + A2 * A2 + A3 * A3;
    const T p12 = p1 * p1 + p2 * p2, p34 = p3 * p3 + p4 * p4, psq = p12 + p34;
    const T mterm = al + be * msq;
#pragma unroll
    for (int a = 0; a < 3; ++a) eg[a] = mterm * f[a] - ga2 * d2[a][0] - ga2 * d2[a][1];
    eg[0] -= d1[5][1];
    eg[1] += d1[5][0];
    eg[2] += -d1[4][0] + d1[3][1];
    eg[3] = q * q * A1 * psq + q * (p1 * d1[7][0] - p2 * d1[6][0] + p3 * d1[9][0] - p4 * d1[8][0]) + d2[4][2] -
            d2[3][1] - d1[2][1];
    eg[4] = q * q * A2 * psq + q * (p1 * d1[7][1] - p2 * d1[6][1] + p3 * d1[9][1] - p4 * d1[8][1]) + d2[3][2] -
            d2[4][0] + d1[2][0];
    eg[5] = q * q * A3 * psq - d2[5][0] - d2[5][1] - d1[1][0] + d1[0][1];
    const T divA = d1[3][0] + d1[4][1];
    const T qqA = q * q * Asq;
    eg[6] = (ha + hb1 * p12 + qqA + T(2) * hb2 * p34) * p1 - d2[6][0] - d2[6][1] + T(2) * q * A1 * d1[7][0] +
            T(2) * q * A2 * d1[7][1] + q * p2 * divA + T(2) * hc * p3;
    eg[7] = (ha + hb1 * p12 + qqA + T(2) * hb2 * p34) * p2 - d2[7][0] - d2[7][1] - T(2) * q * A1 * d1[6][0] -
            T(2) * q * A2 * d1[6][1] - q * p1 * divA + T(2) * hc * p4;
    eg[8] = (ha + hb1 * p34 + qqA + T(2) * hb2 * p12) * p3 - d2[8][0] - d2[8][1] + T(2) * q * A1 * d1[9][0] +
            T(2) * q * A2 * d1[9][1] + q * p4 * divA + T(2) * hc * p1;
    eg[9] = (ha + hb1 * p34 + qqA + T(2) * hb2 * p12) * p4 - d2[9][0] - d2[9][1] - T(2) * q * A1 * d1[8][0] -
            T(2) * q * A2 * d1[8][1] - q * p3 * divA + T(2) * hc * p2;
    project(tp, eg, f);
  }

  // kernels.py:125-191, without the gv factor.
  template <typename T>
  __device__ __forceinline__ static T energy(const TP& tp, const Geom& g, const SiteInfo& s, const T f[NF],
                                             const T d1[NF][2]) {
    const T q = T(tp.q), ha = T(tp.ha), hb1 = T(tp.hb1), hb2 = T(tp.hb2), hc = T(tp.hc);
    const T al = T(tp.alpha), be = T(tp.beta), ga = T(tp.gamma);
    const T m1 = f[0], m2 = f[1], m3 = f[2], A1 = f[3], A2 = f[4], A3 = f[5];
    const T p1 = f[6], p2 = f[7], p3 = f[8], p4 = f[9];
    const T p12 = p1 * p1 + p2 * p2, p34 = p3 * p3 + p4 * p4, psq = p12 + p34;
    const T Asq = A1 * A1 + A2 * A2 + A3 * A3;
    const T msq = m1 * m1 + m2 * m2 + m3 * m3;
    T e = ha / T(2) * psq + hb1 / T(4) * p12 * p12 + hb1 / T(4) * p34 * p34 + hb2 * p12 * p34 +
          T(2) * hc * (p1 * p3 + p2 * p4);
    T kin = T(0), gm = T(0);
#pragma unroll
    for (int a = 6; a < 10; ++a) kin += d1[a][0] * d1[a][0] + d1[a][1] * d1[a][1];
#pragma unroll
    for (int a = 0; a < 3; ++a) gm += d1[a][0] * d1[a][0] + d1[a][1] * d1[a][1];
    e += T(0.5) * kin;
    e += T(0.5) * q * q * Asq * psq;
    e += q * A1 * (p1 * d1[7][0] - p2 * d1[6][0] + p3 * d1[9][0] - p4 * d1[8][0]) +
         q * A2 * (p1 * d1[7][1] - p2 * d1[6][1] + p3 * d1[9][1] - p4 * d1[8][1]);
    const T dA1dy = d1[3][1], dA2dx = d1[4][0], dA3dx = d1[5][0], dA3dy = d1[5][1];
    e += T(0.5) * dA2dx * dA2dx + T(0.5) * dA1dy * dA1dy - dA2dx * dA1dy + T(0.5) * dA3dx * dA3dx +
         T(0.5) * dA3dy * dA3dy;
    e += al / T(2) * msq + be / T(4) * msq * msq;
    e += T(0.5) * ga * ga * gm;
    e += m2 * dA3dx - m1 * dA3dy - m3 * dA2dx + m3 * dA1dy;
    return e - T(tp.vac);
  }

  // 0: m.(d_x m x d_y m)/(4 pi M0^3); 1..3: flux with which = 2, 0, 1 over 2 pi q.
  template <typename T>
  __device__ __forceinline__ static T charge(int kind, const TP& tp, const Geom& g, const SiteInfo& s,
                                             const T f[NF], const T d1[NF][2], const T d2[NF][3]) {
    if (kind == 0) {
      const T cx0 = d1[1][0] * d1[2][1] - d1[2][0] * d1[1][1];
      const T cx1 = d1[2][0] * d1[0][1] - d1[0][0] * d1[2][1];
      const T cx2 = d1[0][0] * d1[1][1] - d1[1][0] * d1[0][1];
      return (f[0] * cx0 + f[1] * cx1 + f[2] * cx2) / (T(4.0 * M_PI) * T(tp.M0) * T(tp.M0) * T(tp.M0));
    }
    T b;
    if (kind == 2) b = d1[5][1];
    else if (kind == 3) b = -d1[5][0];
    else b = d1[4][0] - d1[3][1];
    return b / (T(2.0 * M_PI) * T(tp.q));
  }

  // Bit k set: init mode k (ground 0, initial 1, skyrmions 2, vortex 3) is available on the device.
  static constexpr unsigned kInitModes = 3u;
  // Device initial conditions (spin_triplet_superconducting_magnet/initial_config.py): ground m = (0, 0, M0), psi = u;
  // initial: texture of radius max(xsize, ysize)/10 rescaled to |m| = M0, and a
  // vortex A = ainf fs (-sin th, cos th)/r, psi = u fs e^{i N th} with the
  // rotated angle and fs = tanh(0.2 r) (1 beyond max(xsize, ysize)).
  __device__ __forceinline__ static bool init_site(const InitArgs& a, const Geom& g, int x, int y, double f[NF]) {
    const double M0 = a.pf[20], u1 = a.pf[11], u2 = a.pf[12], ainf = a.pf[16], skN = a.pf[21];
    for (int c = 0; c < NF; ++c) f[c] = 0.0;
    f[2] = M0;
    f[6] = u1;
    f[8] = u2;
    if (a.mode == kInitGround) return true;
    if (a.mode != kInitInitial) return false;
    const double rmax = a.pf[0] > a.pf[1] ? a.pf[0] : a.pf[1];
    double xc, yc, r, th;
    lattice_centre(g, xc, yc);
    polar(g, x, y, xc, yc, a.rot, r, th);
    double m[3];
    texture(texture_profile(r, rmax / 10.0), th, skN, a.ansatz, m);
    rescale3(m, M0);
    const double fs = (r > rmax) ? 1.0 : tanh(0.2 * r);
    const double rr = fmax(r, 1e-12);
    f[0] = m[0];
    f[1] = m[1];
    f[2] = m[2];
    f[3] = -ainf * fs * sin(th) / rr;
    f[4] = ainf * fs * cos(th) / rr;
    f[6] = u1 * fs * cos(a.pf[13] * th);
    f[7] = u1 * fs * sin(a.pf[13] * th);
    f[8] = u2 * fs * cos(a.pf[14] * th);
    f[9] = u2 * fs * sin(a.pf[14] * th);
    return true;
  }

  // |psi1|^2, |psi2|^2 (compute_norm_higgs1/2_point), the 3-component supercurrent
  // (d_yx A2 - d_yy A1, d_xy A1 - d_xx A2, -(d_xx + d_yy) A3) and the magnetic
  // field: component a is the flux density ``which = a`` times gv/(2 pi q)
  // (compute_supercurrent_point, compute_magnetic_field_kernel).
  __host__ __device__ static constexpr int aux_components(int kind) {
    return (kind == kAuxHiggs1 || kind == kAuxHiggs2) ? 1 : (kind == kAuxSupercurrent || kind == kAuxMagneticField) ? 3 : 0;
  }
  template <typename T>
  __device__ __forceinline__ static T aux(int kind, int comp, const TP& tp, const Geom& g, const SiteInfo& s,
                                          const T f[NF], const T d1[NF][2], const T d2[NF][3]) {
    if (kind == kAuxHiggs1) return f[6] * f[6] + f[7] * f[7];
    if (kind == kAuxHiggs2) return f[8] * f[8] + f[9] * f[9];
    if (kind == kAuxMagneticField) {
      // flux kinds: 1 which=2, 2 which=0, 3 which=1
      const int fk = comp == 2 ? 1 : comp + 2;
      return charge<T>(fk, tp, g, s, f, d1, d2) * T(g.gv);
    }
    if (comp == 0) return d2[4][2] - d2[3][1];
    if (comp == 1) return d2[3][2] - d2[4][0];
    return -(d2[5][0] + d2[5][1]);
  }

  template <typename T>
  __device__ __forceinline__ static void flow_fix(const TP& tp, const T f[NF], T eg[NF]) {
    project(tp, eg, f);
  }

  // kernels.py:47-84.
  template <typename T>
  __device__ __forceinline__ static void normalize(const TP& tp, T f[NF]) {
    const T n = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (n != T(0)) {
      const T sc = T(tp.M0) / n;
      f[0] = f[0] * sc;
      f[1] = f[1] * sc;
      f[2] = f[2] * sc;
    }
  }

  // kernels.py:86-120.
  template <typename T>
  __device__ __forceinline__ static void project(const TP& tp, T v[NF], const T f[NF]) {
    const T M0 = T(tp.M0);
    const T lm = v[0] * f[0] / M0 + v[1] * f[1] / M0 + v[2] * f[2] / M0;
    v[0] -= lm * f[0] / M0;
    v[1] -= lm * f[1] / M0;
    v[2] -= lm * f[2] / M0;
  }
};

}  // namespace ss
