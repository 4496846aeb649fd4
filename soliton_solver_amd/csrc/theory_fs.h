// Ferromagnetic superconductor (reference: theories/ferromagnetic_superconductor/
// kernels.py, params.py).  Fields m1, m2, m3, A1, A2, A3, psi1, psi2 (nf = 8);
// the magnetisation is constrained to |m| = M0.
#pragma once

#include "common.h"
#include "init.h"

namespace ss {

struct FerromagneticSC {
  static constexpr int NF = 8;
  static constexpr int NMAG = 3;
  static constexpr bool kUnitConstraint = true;
  static constexpr bool kL2Normalize = false;
  static constexpr int kNumCharges = 4;  // 0 skyrmion, 1 vortex (which=2), 2 which=0, 3 which=1
  static constexpr const char* kName = "ferromagnetic_superconductor";

  struct TP {
    double q, ha, hb, eta1, eta2, alpha, beta, gamma, M0, vac;
  };

  // Packing: ferromagnetic_superconductor/params.py:266-312.
  static int make(const int32_t* pi, int ni, const double* pf, int nfl, TP* tp, char* msg, int cap) {
    if (ni < 13 || nfl < 24) {
      snprintf(msg, cap, "ferromagnetic superconductor needs p_i[0..12] and p_f[0..23] (got %d, %d)", ni, nfl);
      return -1;
    }
    if (pi[4] != NF || pi[10] != NMAG) {
      snprintf(msg, cap, "ferromagnetic superconductor kernels are built for 8 fields, 3 magnetisation (got %d, %d)",
               pi[4], pi[10]);
      return -1;
    }
    tp->q = pf[6];
    tp->ha = pf[7];
    tp->hb = pf[8];
    tp->eta1 = pf[9];
    tp->eta2 = pf[10];
    tp->alpha = pf[14];
    tp->beta = pf[15];
    tp->gamma = pf[16];
    tp->M0 = pf[17];
    if (tp->M0 == 0.0) {
      snprintf(msg, cap, "ferromagnetic superconductor needs M0 != 0");
      return -1;
    }
    // subtracted constant (kernels.py:178-179)
    const double den = 16.0 * tp->eta1 * tp->eta1 - 4.0 * tp->hb * tp->beta;
    tp->vac = (tp->alpha * tp->alpha * tp->hb + tp->ha * tp->ha * tp->beta - 4.0 * tp->ha * tp->alpha * tp->eta1) / den;
    return 0;
  }

  // kernels.py:533-619.
  template <typename T>
  __device__ __forceinline__ static void variation(const TP& tp, const Geom& g, const SiteInfo& s,
                                                   const T f[NF], const T d1[NF][2], const T d2[NF][3],
                                                   T eg[NF]) {
    const T q = T(tp.q), ha = T(tp.ha), hb = T(tp.hb), e1 = T(tp.eta1), e2 = T(tp.eta2);
    const T al = T(tp.alpha), be = T(tp.beta), ga2 = T(tp.gamma) * T(tp.gamma);
    const T m1 = f[0], m2 = f[1], m3 = f[2], A1 = f[3], A2 = f[4], A3 = f[5], p1 = f[6], p2 = f[7];
    const T psq = p1 * p1 + p2 * p2;
    const T msq = m1 * m1 + m2 * m2 + m3 * m3;
    T dmsq = T(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) dmsq += d1[a][0] * d1[a][0] + d1[a][1] * d1[a][1];
    const T Asq = A1 * A1 + A2 * A2 + A3 * A3;
    const T mterm = al + be * msq + T(2) * e1 * psq;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const T lap = d2[a][0] + d2[a][1];
      eg[a] = mterm * f[a] - ga2 * d2[a][0] - ga2 * d2[a][1] - T(2) * e2 * lap * psq;
    }
    eg[0] -= d1[5][1];
    eg[1] += d1[5][0];
    eg[2] += -d1[4][0] + d1[3][1];
    eg[3] = q * q * A1 * psq + q * (p1 * d1[7][0] - p2 * d1[6][0]) + d2[4][2] - d2[3][1] - d1[2][1];
    eg[4] = q * q * A2 * psq + q * (p1 * d1[7][1] - p2 * d1[6][1]) + d2[3][2] - d2[4][0] + d1[2][0];
    eg[5] = q * q * A3 * psq - d2[5][0] - d2[5][1] - d1[1][0] + d1[0][1];
    const T divA = d1[3][0] + d1[4][1];
    const T common = ha + hb * psq + q * q * Asq + T(2) * e1 * msq;
    eg[6] = common * p1 - d2[6][0] - d2[6][1] + T(2) * q * A1 * d1[7][0] + T(2) * q * A2 * d1[7][1] + q * p2 * divA +
            T(2) * e2 * p1 * p1 * dmsq;
    eg[7] = common * p2 - d2[7][0] - d2[7][1] - T(2) * q * A1 * d1[6][0] - T(2) * q * A2 * d1[6][1] - q * p1 * divA +
            T(2) * e2 * p2 * p2 * dmsq;
    project(tp, eg, f);
  }

  // kernels.py:124-181, without the gv factor.
  template <typename T>
  __device__ __forceinline__ static T energy(const TP& tp, const Geom& g, const SiteInfo& s, const T f[NF],
                                             const T d1[NF][2]) {
    const T q = T(tp.q), ha = T(tp.ha), hb = T(tp.hb), e1 = T(tp.eta1), e2 = T(tp.eta2);
    const T al = T(tp.alpha), be = T(tp.beta), ga = T(tp.gamma);
    const T m1 = f[0], m2 = f[1], m3 = f[2], A1 = f[3], A2 = f[4], A3 = f[5], p1 = f[6], p2 = f[7];
    const T psq = p1 * p1 + p2 * p2;
    const T msq = m1 * m1 + m2 * m2 + m3 * m3;
    const T Asq = A1 * A1 + A2 * A2 + A3 * A3;
    T gradm = T(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) gradm += d1[a][0] * d1[a][0] + d1[a][1] * d1[a][1];
    const T dA1dy = d1[3][1], dA2dx = d1[4][0], dA3dx = d1[5][0], dA3dy = d1[5][1];
    T e = ha / T(2) * psq + hb / T(4) * psq * psq;
    e += T(0.5) * (d1[6][0] * d1[6][0] + d1[6][1] * d1[6][1] + d1[7][0] * d1[7][0] + d1[7][1] * d1[7][1]);
    e += T(0.5) * q * q * Asq * psq;
    e += q * A1 * (p1 * d1[7][0] - p2 * d1[6][0]) + q * A2 * (p1 * d1[7][1] - p2 * d1[6][1]);
    e += T(0.5) * dA2dx * dA2dx + T(0.5) * dA1dy * dA1dy - dA2dx * dA1dy + T(0.5) * dA3dx * dA3dx +
         T(0.5) * dA3dy * dA3dy;
    e += al / T(2) * msq + be / T(4) * msq * msq;
    e += ga * ga / T(2) * gradm;
    e += m2 * dA3dx - m1 * dA3dy - m3 * dA2dx + m3 * dA1dy;
    e += e1 * msq * psq;
    e += e2 * gradm * psq;
    return e - T(tp.vac);
  }

  // 0: m.(d_x m x d_y m)/(4 pi M0^3); 1..3: vortex density with which = 2, 0, 1 over 2 pi q.
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
  // Device initial conditions (ferromagnetic_superconductor/initial_config.py): ground m = (0, 0, M0), psi = u;
  // initial: texture of radius max(xsize, ysize)/10 rescaled to |m| = M0, and a
  // vortex A = ainf fs (-sin th, cos th)/r, psi = u fs e^{i N th} with the
  // rotated angle and fs = tanh(0.2 r) (1 beyond max(xsize, ysize)).
  __device__ __forceinline__ static bool init_site(const InitArgs& a, const Geom& g, int x, int y, double f[NF]) {
    const double M0 = a.pf[17], u1 = a.pf[11], ainf = a.pf[13], skN = a.pf[18];
    for (int c = 0; c < NF; ++c) f[c] = 0.0;
    f[2] = M0;
    f[6] = u1;
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
    f[6] = u1 * fs * cos(a.vortex_number * th);
    f[7] = u1 * fs * sin(a.vortex_number * th);
    return true;
  }

  // |psi|^2 (compute_norm_higgs_point), the 3-component supercurrent
  // (d_yx A2 - d_yy A1, d_xy A1 - d_xx A2, -(d_xx + d_yy) A3) and the magnetic
  // field: component a is the flux density ``which = a`` times gv/(2 pi q)
  // (compute_supercurrent_point, compute_magnetic_field_kernel).
  __host__ __device__ static constexpr int aux_components(int kind) {
    return kind == kAuxHiggs ? 1 : (kind == kAuxSupercurrent || kind == kAuxMagneticField) ? 3 : 0;
  }
  template <typename T>
  __device__ __forceinline__ static T aux(int kind, int comp, const TP& tp, const Geom& g, const SiteInfo& s,
                                          const T f[NF], const T d1[NF][2], const T d2[NF][3]) {
    if (kind == kAuxHiggs) return f[6] * f[6] + f[7] * f[7];
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

  // |m| = M0 (kernels.py:46-83).
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

  // v -= (v.m/M0) m/M0 (kernels.py:85-120).
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
