// Baby Skyrme model (reference: theories/baby_skyrme/kernels.py, params.py).
// Fields m1, m2, m3 (nf = 3), unit-norm constrained.
#pragma once

#include "common.h"
#include "init.h"

namespace ss {

struct BabySkyrme {
  static constexpr int NF = 3;
  static constexpr int NMAG = 3;
  static constexpr bool kUnitConstraint = true;
  static constexpr bool kL2Normalize = false;
  static constexpr int kNumCharges = 1;  // 0 skyrmion number
  static constexpr const char* kName = "baby_skyrme";

  struct TP {
    double mpi2, kappa2;
    int standard, holomorphic, easyplane, dihedral2, aloof, dihedral3, broken, doublevacua;
    int N;
  };

  // Packing: baby_skyrme/params.py:331-378.
  static int make(const int32_t* pi, int ni, const double* pf, int nfl, TP* tp, char* msg, int cap) {
    if (ni < 20 || nfl < 8) {
      snprintf(msg, cap, "baby Skyrme needs p_i[0..19] and p_f[0..7] (got %d, %d)", ni, nfl);
      return -1;
    }
    if (pi[4] != NF || pi[10] != NMAG) {
      snprintf(msg, cap, "baby Skyrme kernels are built for 3 magnetisation fields (got nf=%d, nmag=%d)", pi[4], pi[10]);
      return -1;
    }
    tp->mpi2 = pf[6] * pf[6];
    tp->kappa2 = pf[7] * pf[7];
    tp->standard = pi[11];
    tp->holomorphic = pi[12];
    tp->easyplane = pi[13];
    tp->dihedral2 = pi[14];
    tp->aloof = pi[15];
    tp->dihedral3 = pi[16];
    tp->broken = pi[17];
    tp->doublevacua = pi[18];
    tp->N = pi[19];
    if (tp->broken && tp->N < 1) {
      snprintf(msg, cap, "broken potential needs N >= 1 (got %d)", tp->N);
      return -1;
    }
    return 0;
  }

  // kernels.py:307-458.
  template <typename T>
  __device__ __forceinline__ static void variation(const TP& tp, const Geom& g, const SiteInfo& s,
                                                   const T f[NF], const T d1[NF][2], const T d2[NF][3],
                                                   T eg[NF]) {
    const T k2 = T(tp.kappa2), mpi2 = T(tp.mpi2);
    const T s0 = d1[0][0] * d1[0][0] + d1[1][0] * d1[1][0] + d1[2][0] * d1[2][0];
    const T s1 = d1[0][1] * d1[0][1] + d1[1][1] * d1[1][1] + d1[2][1] * d1[2][1];
    const T dot01 = d1[0][0] * d1[0][1] + d1[1][0] * d1[1][1] + d1[2][0] * d1[2][1];
    T sumA = T(0), sumB = T(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      sumA += d2[a][2] * d1[a][1] - d2[a][1] * d1[a][0];
      sumB += d2[a][2] * d1[a][0] - d2[a][0] * d1[a][1];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      // -lap - k^2 [ Sxx s1 + Syy s0 + F0 sumA + F1 sumB - 2 Sxy dot01 ]
      T v = -(d2[a][0] + d2[a][1]);
      v -= k2 * d2[a][0] * s1;
      v -= k2 * (d1[a][0] * sumA - d2[a][2] * dot01);
      v -= k2 * d2[a][1] * s0;
      v -= k2 * (d1[a][1] * sumB - d2[a][2] * dot01);
      eg[a] = v;
    }
    const T m1 = f[0], m2 = f[1], m3 = f[2];
    if (tp.standard) eg[2] -= mpi2;
    if (tp.holomorphic) {
      const T t = T(1) - m3;
      eg[2] -= T(4) * mpi2 * t * t * t;
    }
    if (tp.easyplane) eg[0] += mpi2 * m1;
    if (tp.dihedral2) {
      eg[0] -= mpi2 * (T(1) - m3 * m3) * m1;
      eg[2] -= mpi2 * (T(1) - m1 * m1) * m3;
    }
    if (tp.aloof) {
      const T t = T(1) - m3;
      const T t3 = t * t * t;
      eg[2] -= mpi2 * (T(0.5) * (T(1) + t3) + T(1.5) * t3);
    }
    if (tp.dihedral3) {
      eg[0] += T(48) * mpi2 * (m2 * m2 - m1 * m1) * (T(1) - m3);
      eg[1] += T(96) * mpi2 * m1 * m2 * (T(1) - m3);
      const T u = T(1) - T(3) * m3;
      eg[2] -= T(16) * mpi2 * (-(m1 * m1 * m1) + T(3) * m1 * (m2 * m2) + u * u);
    }
    if (tp.broken) {
      const T N = T(tp.N);
      const T r2 = m1 * m1 + m2 * m2;
      const T theta = atan2(m2, m1);
      const T rN2 = pow(r2, T(0.5) * N);
      const T rN = pow(r2, N);
      const T c = cos(N * theta), sn = sin(N * theta);
      const T nrm = T(1) - T(2) * c * rN2 + rN;
      const T den = sqrt(nrm);
      if (den != T(0) && r2 != T(0)) {
        const T pref = T(2) * mpi2 * N * sqrt(nrm) * pow(r2, T(-1) + T(0.5) * N) * (T(-1) + m3) / den;
        eg[0] += pref * (sn * m2 + m1 * (c - rN2));
        eg[1] += (-pref) * (sn * m1 + m2 * (-c + rN2));
      }
      eg[2] -= mpi2 * nrm;
    }
    if (tp.doublevacua) eg[2] -= T(2) * mpi2 * m3;
    project(tp, eg, f);
  }

  // kernels.py:126-199, without the gv factor.
  template <typename T>
  __device__ __forceinline__ static T energy(const TP& tp, const Geom& g, const SiteInfo& s,
                                             const T f[NF], const T d1[NF][2]) {
    const T k2 = T(tp.kappa2), mpi2 = T(tp.mpi2);
    const T gx = d1[0][0] * d1[0][0] + d1[1][0] * d1[1][0] + d1[2][0] * d1[2][0];
    const T gy = d1[0][1] * d1[0][1] + d1[1][1] * d1[1][1] + d1[2][1] * d1[2][1];
    T e = T(0.5) * (gx + gy);
    const T cx = d1[1][0] * d1[2][1] - d1[2][0] * d1[1][1];
    const T cy = d1[2][0] * d1[0][1] - d1[0][0] * d1[2][1];
    const T cz = d1[0][0] * d1[1][1] - d1[1][0] * d1[0][1];
    e += T(0.5) * k2 * (cx * cx + cy * cy + cz * cz);
    const T m1 = f[0], m2 = f[1], m3 = f[2];
    if (tp.standard) e += mpi2 * (T(1) - m3);
    if (tp.holomorphic) {
      const T t = T(1) - m3;
      e += mpi2 * t * t * t * t;
    }
    if (tp.easyplane) e += T(0.5) * mpi2 * m1 * m1;
    if (tp.dihedral2) e += T(0.5) * mpi2 * (T(1) - m1 * m1) * (T(1) - m3 * m3);
    if (tp.aloof) {
      const T t = T(1) - m3;
      e += T(0.5) * mpi2 * t * (T(1) + t * t * t);
    }
    if (tp.dihedral3) e += T(16) * mpi2 * (T(1) - m3) * (T(1) + T(3) * m3 * m3 + T(3) * m1 * m2 * m2 - m1 * m1 * m1);
    if (tp.broken) {
      T zr = m1, zi = m2;
      for (int k = 0; k < tp.N - 1; ++k) {
        const T nr = zr * m1 - zi * m2;
        zi = zr * m2 + zi * m1;
        zr = nr;
      }
      const T wr = T(1) - zr, wi = -zi;
      e += mpi2 * (T(1) - m3) * (wr * wr + wi * wi);
    }
    if (tp.doublevacua) e += mpi2 * (T(1) - m3 * m3);
    return e;
  }

  // Skyrmion number density m.(d_x m x d_y m)/(4 pi) (kernels.py:236-270).
  template <typename T>
  __device__ __forceinline__ static T charge(int kind, const TP& tp, const Geom& g, const SiteInfo& s,
                                             const T f[NF], const T d1[NF][2], const T d2[NF][3]) {
    const T cx0 = d1[1][0] * d1[2][1] - d1[2][0] * d1[1][1];
    const T cx1 = d1[2][0] * d1[0][1] - d1[0][0] * d1[2][1];
    const T cx2 = d1[0][0] * d1[1][1] - d1[1][0] * d1[0][1];
    return (f[0] * cx0 + f[1] * cx1 + f[2] * cx2) * T(1.0 / (4.0 * M_PI));
  }

  // Bit k set: init mode k (ground 0, initial 1, skyrmions 2, vortex 3) is available on the device.
  static constexpr unsigned kInitModes = 7u;
  // Device initial conditions (baby_skyrme/initial_config.py).
  __device__ __forceinline__ static bool init_site(const InitArgs& a, const Geom& g, int x, int y, double f[NF]) {
    return unit_vector_init<NF>(a, g, x, y, f);
  }

  // No auxiliary densities.
  __host__ __device__ static constexpr int aux_components(int kind) { return 0; }
  template <typename T>
  __device__ __forceinline__ static T aux(int kind, int comp, const TP& tp, const Geom& g, const SiteInfo& s,
                                          const T f[NF], const T d1[NF][2], const T d2[NF][3]) {
    return T(0);
  }

  template <typename T>
  __device__ __forceinline__ static void flow_fix(const TP& tp, const T f[NF], T eg[NF]) {
    project(tp, eg, f);
  }

  template <typename T>
  __device__ __forceinline__ static void normalize(const TP& tp, T f[NF]) {
    const T n = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (n != T(0)) {
      const T inv = T(1) / n;
      f[0] = f[0] * inv;
      f[1] = f[1] * inv;
      f[2] = f[2] * inv;
    }
  }

  template <typename T>
  __device__ __forceinline__ static void project(const TP& tp, T v[NF], const T f[NF]) {
    const T lm = v[0] * f[0] + v[1] * f[1] + v[2] * f[2];
    v[0] -= lm * f[0];
    v[1] -= lm * f[1];
    v[2] -= lm * f[2];
  }
};

}  // namespace ss
