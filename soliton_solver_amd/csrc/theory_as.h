// Anisotropic s+id superconductor (reference: theories/anisotropic_superconductor/
// kernels.py, params.py).  Fields A1, A2, psi1 (re, im), psi2 (re, im): nf = 6.
//
// gamma_entry(i, j, a, b) (kernels.py:45-94) is non-zero only for i == j, so
// every sum over coordinate pairs here runs over i == j; the reference's
// i != j terms are products with an exact zero.
#pragma once

#include "common.h"
#include "init.h"

namespace ss {

struct AnisotropicSC {
  static constexpr int NF = 6;
  static constexpr int NMAG = 0;
  static constexpr bool kUnitConstraint = false;
  static constexpr bool kL2Normalize = false;
  static constexpr int kNumCharges = 1;  // 0 vortex number
  static constexpr const char* kName = "anisotropic_superconductor";

  // The coupling constants in scalar type P: double for the flow, dual
  // numbers for the derivatives of the energy with respect to the packed
  // parameters (energy_param_grad).
  template <typename P>
  struct TPT {
    P kappa, Q, a1, a2, b1, b2, b3, b4;
    P G[2][2][2];  // G[i][a][b] = gamma_entry(i, i, a, b)
    P vac;         // vacuum constant subtracted per site (kernels.py:168)
  };
  using TP = TPT<double>;
  static constexpr int kNumPF = 25;  // packed p_f entries (anisotropic_superconductor/params.py:287-308)
  static constexpr bool kParamGrad = true;

  // Packed parameters -> coupling constants (anisotropic_superconductor/
  // params.py:266-313; gamma_entry kernels.py:45-94; vacuum constant
  // kernels.py:140-168).
  template <typename P>
  __host__ __device__ static void fill_tp(const P* pf, TPT<P>& tp) {
    tp.kappa = pf[6];
    tp.Q = pf[7];
    tp.a1 = pf[8];
    tp.a2 = pf[9];
    tp.b1 = pf[10];
    tp.b2 = pf[11];
    tp.b3 = pf[12];
    tp.b4 = pf[13];
    const P g1 = pf[14], g2 = pf[15], g12 = pf[16];
    tp.G[0][0][0] = P(2.0) * g1;
    tp.G[0][0][1] = -(P(2.0) * g12);
    tp.G[0][1][0] = -(P(2.0) * g12);
    tp.G[0][1][1] = P(2.0) * g2;
    tp.G[1][0][0] = P(2.0) * g1;
    tp.G[1][0][1] = P(2.0) * g12;
    tp.G[1][1][0] = P(2.0) * g12;
    tp.G[1][1][1] = P(2.0) * g2;
    const P u1 = pf[17], u2 = pf[18], th = pf[19];
    tp.vac = tp.a1 * u1 * u1 + tp.a2 * u2 * u2 + tp.b1 * u1 * u1 * u1 * u1 + tp.b2 * u2 * u2 * u2 * u2 +
             tp.b3 * u1 * u1 * u2 * u2 + P(2.0) * tp.b4 * u1 * u1 * u2 * u2 * cos(P(2.0) * th);
  }

  static int make(const int32_t* pi, int ni, const double* pf, int nfl, TP* tp, char* msg, int cap) {
    if (ni < 14 || nfl < 25) {
      snprintf(msg, cap, "anisotropic superconductor needs p_i[0..13] and p_f[0..24] (got %d, %d)", ni, nfl);
      return -1;
    }
    if (pi[4] != NF) {
      snprintf(msg, cap, "anisotropic superconductor kernels are built for 6 fields (got %d)", pi[4]);
      return -1;
    }
    fill_tp<double>(pf, *tp);
    return 0;
  }

  // Second derivative with the reference's idx_d2(i, j, a) meaning.
  template <typename T>
  __device__ __forceinline__ static T D2(const T d2[NF][3], int i, int j, int a) {
    return i == j ? d2[a][i] : d2[a][2];
  }

  // kernels.py:555-658.
  template <typename T>
  __device__ __forceinline__ static void variation(const TP& tp, const Geom& g, const SiteInfo& s,
                                                   const T f[NF], const T d1[NF][2], const T d2[NF][3],
                                                   T eg[NF]) {
    const T k2 = T(tp.kappa) * T(tp.kappa), Q = T(tp.Q);
    const T p1r = f[2], p1i = f[3], p2r = f[4], p2i = f[5];
    const T s1 = p1r * p1r + p1i * p1i, s2 = p2r * p2r + p2i * p2i;
    const T c12 = p1r * p2r + p1i * p2i;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const T Aj = f[j];
      T v = T(0);
#pragma unroll
      for (int k = 0; k < 2; ++k) v += k2 * (D2(d2, j, k, k) - D2(d2, k, k, j));
      const T g00 = T(tp.G[j][0][0]), g01 = T(tp.G[j][0][1]), g10 = T(tp.G[j][1][0]), g11 = T(tp.G[j][1][1]);
      v -= Q * g00 * (p1i * d1[2][j] - p1r * d1[3][j] - Q * Aj * s1);
      v -= Q * g01 * (p2i * d1[2][j] - p2r * d1[3][j] - Q * Aj * c12);
      v -= Q * g10 * (p1i * d1[4][j] - p1r * d1[5][j] - Q * Aj * c12);
      v -= Q * g11 * (p2i * d1[4][j] - p2r * d1[5][j] - Q * Aj * s2);
      eg[j] = v;
    }
    const T a1 = T(tp.a1), a2 = T(tp.a2), b1 = T(tp.b1), b2 = T(tp.b2), b3 = T(tp.b3), b4 = T(tp.b4);
    const T dd2 = p2r * p2r - p2i * p2i, dd1 = p1r * p1r - p1i * p1i;
    // Potential part: twice dV/dpsi (the reference's factor, kernels.py:623-626).
    eg[2] = T(2) * (T(2) * a1 * p1r + T(4) * b1 * p1r * s1 + T(2) * b3 * p1r * s2 +
                    T(2) * b4 * (T(2) * p1r * dd2 + T(4) * p1i * p2r * p2i));
    eg[3] = T(2) * (T(2) * a1 * p1i + T(4) * b1 * p1i * s1 + T(2) * b3 * p1i * s2 +
                    T(2) * b4 * (-T(2) * p1i * dd2 + T(4) * p1r * p2r * p2i));
    eg[4] = T(2) * (T(2) * a2 * p2r + T(4) * b2 * p2r * s2 + T(2) * b3 * p2r * s1 +
                    T(2) * b4 * (T(2) * p2r * dd1 + T(4) * p1r * p1i * p2i));
    eg[5] = T(2) * (T(2) * a2 * p2i + T(4) * b2 * p2i * s2 + T(2) * b3 * p2i * s1 +
                    T(2) * b4 * (-T(2) * p2i * dd1 + T(4) * p1r * p1i * p2r));
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const T Ai = f[i];
      const T g00 = T(tp.G[i][0][0]), g01 = T(tp.G[i][0][1]), g10 = T(tp.G[i][1][0]), g11 = T(tp.G[i][1][1]);
      const T dA = d1[i][i];
      const T QQAA = Q * Q * Ai * Ai;
      const T t2 = d2[2][i] - QQAA * p1r - Q * (p1i * dA + T(2) * Ai * d1[3][i]);
      const T t4 = d2[4][i] - QQAA * p2r - Q * (p2i * dA + T(2) * Ai * d1[5][i]);
      const T t3 = d2[3][i] - QQAA * p1i + Q * (p1r * dA + T(2) * Ai * d1[2][i]);
      const T t5 = d2[5][i] - QQAA * p2i + Q * (p2r * dA + T(2) * Ai * d1[4][i]);
      eg[2] -= g00 * t2 + g01 * t4;
      eg[3] -= g00 * t3 + g01 * t5;
      eg[4] -= g10 * t2 + g11 * t4;
      eg[5] -= g10 * t3 + g11 * t5;
    }
  }

  // kernels.py:96-170, without the gv factor.  TPX: TP, or TPT<T> for
  // derivatives with respect to the coupling constants.
  template <typename T, class TPX = TP>
  __device__ __forceinline__ static T energy(const TPX& tp, const Geom& g, const SiteInfo& s,
                                             const T f[NF], const T d1[NF][2]) {
    const T k = T(tp.kappa), Q = T(tp.Q);
    const T p1r = f[2], p1i = f[3], p2r = f[4], p2i = f[5];
    const T s1 = p1r * p1r + p1i * p1i, s2 = p2r * p2r + p2i * p2i;
    const T curl = d1[1][0] - d1[0][1];
    const T Fm = T(0.5) * k * k * curl * curl;
    T Fp = T(tp.a1) * s1 + T(tp.a2) * s2;
    Fp += T(tp.b1) * s1 * s1 + T(tp.b2) * s2 * s2 + T(tp.b3) * s1 * s2;
    Fp += T(2) * T(tp.b4) * (p1r * p1r * p2r * p2r + p1i * p1i * p2i * p2i - p1i * p1i * p2r * p2r -
                             p1r * p1r * p2i * p2i + T(4) * p1r * p1i * p2r * p2i);
    T Fk = T(0);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const T A = f[i];
      const T g00 = T(tp.G[i][0][0]), g01 = T(tp.G[i][0][1]), g10 = T(tp.G[i][1][0]), g11 = T(tp.G[i][1][1]);
      const T r1 = d1[2][i], i1 = d1[3][i], r2 = d1[4][i], i2 = d1[5][i];
      Fk += T(0.5) * g00 * (r1 * r1 + i1 * i1 + T(2) * Q * A * (p1r * i1 - p1i * r1) + Q * Q * A * A * s1);
      Fk += T(0.5) * g01 * (r1 * r2 + i1 * i2 + Q * A * (p1r * i2 - p1i * r2) + Q * A * (p2r * i1 - p2i * r1) +
                            Q * Q * A * A * (p1r * p2r + p1i * p2i));
      Fk += T(0.5) * g10 * (r2 * r1 + i2 * i1 + Q * A * (p2r * i1 - p2i * r1) + Q * A * (p1r * i2 - p1i * r2) +
                            Q * Q * A * A * (p2r * p1r + p2i * p1i));
      Fk += T(0.5) * g11 * (r2 * r2 + i2 * i2 + T(2) * Q * A * (p2r * i2 - p2i * r2) + Q * Q * A * A * s2);
    }
    return Fk + Fm + Fp - T(tp.vac);
  }

  // Vortex number density (d_x A2 - d_y A1)/(2 pi Q) (kernels.py:207-240).
  template <typename T>
  __device__ __forceinline__ static T charge(int kind, const TP& tp, const Geom& g, const SiteInfo& s,
                                             const T f[NF], const T d1[NF][2], const T d2[NF][3]) {
    return (d1[1][0] - d1[0][1]) / (T(2.0 * M_PI) * T(tp.Q));
  }

  // Bit k set: init mode k (ground 0, initial 1, skyrmions 2, vortex 3) is available on the device.
  static constexpr unsigned kInitModes = 3u;
  // Device initial conditions (anisotropic_superconductor/initial_config.py):
  // ground A = 0, Delta_s = u1, Delta_d = u2 e^{i theta_diff}; initial vortex
  // with profile tanh(r), windings N1, N2 of the two order parameters.
  __device__ __forceinline__ static bool init_site(const InitArgs& a, const Geom& g, int x, int y, double f[NF]) {
    const double u1 = a.pf[17], u2 = a.pf[18], thd = a.pf[19];
    f[0] = f[1] = f[3] = 0.0;
    f[2] = u1;
    f[4] = u2 * cos(thd);
    f[5] = u2 * sin(thd);
    if (a.mode == kInitGround) return true;
    if (a.mode != kInitInitial) return false;
    const double rmax = a.pf[0] > a.pf[1] ? a.pf[0] : a.pf[1];
    const double ainf = a.pf[22], N1 = a.pf[23], N2 = a.pf[24];
    double fs, sth, cth, rr, th;
    centred_vortex(g, x, y, 1.0, rmax, fs, sth, cth, rr, th);
    f[0] = ainf * fs * sth / rr;
    f[1] = -ainf * fs * cth / rr;
    f[2] = u1 * fs * cos(N1 * th);
    f[3] = u1 * fs * sin(N1 * th);
    f[4] = u2 * fs * cos(N2 * th);
    f[5] = u2 * fs * sin(N2 * th);
    return true;
  }

  // |Delta_s|^2, |Delta_d|^2, cos(arg Delta_s - arg Delta_d), the 2-component
  // supercurrent, and the magnetic field B3 times gv/(2 pi q)
  // (compute_norm_higgs1/2_point, compute_phase_difference_point,
  // compute_supercurrent_point, compute_magnetic_field_kernel).
  __host__ __device__ static constexpr int aux_components(int kind) {
    return (kind == kAuxHiggs1 || kind == kAuxHiggs2 || kind == kAuxPhaseDifference || kind == kAuxMagneticField) ? 1
         : kind == kAuxSupercurrent ? 2 : 0;
  }
  template <typename T>
  __device__ __forceinline__ static T aux(int kind, int comp, const TP& tp, const Geom& g, const SiteInfo& s,
                                          const T f[NF], const T d1[NF][2], const T d2[NF][3]) {
    if (kind == kAuxHiggs1) return f[2] * f[2] + f[3] * f[3];
    if (kind == kAuxHiggs2) return f[4] * f[4] + f[5] * f[5];
    if (kind == kAuxPhaseDifference) return cos(atan2(f[3], f[2]) - atan2(f[5], f[4]));
    if (kind == kAuxMagneticField) return charge<T>(0, tp, g, s, f, d1, d2) * T(g.gv);
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
