// Chiral liquid crystal (reference: theories/liquid_crystal/kernels.py,
// params.py).  Fields n1, n2, n3 (director, |n| = 1) and Phi, the potential of
// the flexoelectric depolarisation (nf = 4).
#pragma once

#include "common.h"
#include "init.h"

namespace ss {

struct LiquidCrystal {
  static constexpr int NF = 4;
  static constexpr int NMAG = 3;
  static constexpr bool kUnitConstraint = true;
  static constexpr bool kL2Normalize = false;
  static constexpr int kNumCharges = 2;  // 0 skyrmion number, 1 bound electric charge
  static constexpr const char* kName = "liquid_crystal";

  struct TP {
    double potE, potw0, ceps, r;  // coup_PotE, coup_Potw0, coup_eps, e3/e1
    double D00, D01, D10, D11;     // twist (Dresselhaus) or splay-bend (Rashba)
    int depol;
  };

  // Packing: liquid_crystal/params.py, pack_device_params.
  static int make(const int32_t* pi, int ni, const double* pf, int nfl, TP* tp, char* msg, int cap) {
    if (ni < 14 || nfl < 11) {
      snprintf(msg, cap, "liquid crystal needs p_i[0..13] and p_f[0..10] (got %d, %d)", ni, nfl);
      return -1;
    }
    if (pi[4] != NF || pi[10] != NMAG) {
      snprintf(msg, cap, "liquid crystal kernels are built for 4 fields, 3 director components (got %d, %d)",
               pi[4], pi[10]);
      return -1;
    }
    tp->potE = pf[6];
    tp->potw0 = pf[7];
    tp->ceps = pf[8];
    tp->depol = pi[13];
    if (tp->depol && (pf[9] == 0.0 || pf[8] == 0.0)) {
      snprintf(msg, cap, "liquid crystal depolarisation needs e1 != 0 and coup_eps != 0");
      return -1;
    }
    tp->r = pf[9] != 0.0 ? pf[10] / pf[9] : 0.0;
    tp->D00 = tp->D01 = tp->D10 = tp->D11 = 0.0;
    if (pi[11]) { tp->D00 = -1.0; tp->D11 = -1.0; }
    else if (pi[12]) { tp->D01 = -1.0; tp->D10 = 1.0; }
    return 0;
  }

  __device__ __forceinline__ static constexpr double levi(int i, int j, int k) {
    return ((i == 0 && j == 1 && k == 2) || (i == 1 && j == 2 && k == 0) || (i == 2 && j == 0 && k == 1)) ? 1.0
         : ((i == 0 && j == 2 && k == 1) || (i == 2 && j == 1 && k == 0) || (i == 1 && j == 0 && k == 2)) ? -1.0
         : 0.0;
  }

  // d_i d_j of component a; (0, 1) and (1, 0) share the mixed entry.
  template <typename T>
  __device__ __forceinline__ static T dd(const T d2[NF][3], int a, int i, int j) {
    return i == j ? d2[a][i] : d2[a][2];
  }

  // kernels.py:403-509: exchange, DMI, anchoring/field potential on n3, the
  // depolarisation terms (:492-500), boundary-band mask, tangent projection.
  template <typename T>
  __device__ __forceinline__ static void variation(const TP& tp, const Geom& g, const SiteInfo& s,
                                                   const T f[NF], const T d1[NF][2], const T d2[NF][3],
                                                   T eg[NF]) {
    const T D00 = T(tp.D00), D01 = T(tp.D01), D10 = T(tp.D10), D11 = T(tp.D11);
#pragma unroll
    for (int i = 0; i < NMAG; ++i) eg[i] = -(d2[i][0] + d2[i][1]);
    eg[3] = T(0);
#pragma unroll
    for (int i = 0; i < NMAG; ++i) {
#pragma unroll
      for (int j = 0; j < NMAG; ++j) {
        if (levi(0, i, j) != 0.0) {
          const T e0 = T(levi(0, i, j));
          eg[i] += T(2) * D00 * e0 * d1[j][0];
          eg[i] += T(2) * D01 * e0 * d1[j][1];
        }
        if (levi(1, i, j) != 0.0) {
          const T e1 = T(levi(1, i, j));
          eg[i] += T(2) * D10 * e1 * d1[j][0];
          eg[i] += T(2) * D11 * e1 * d1[j][1];
        }
      }
    }
    eg[2] -= T(tp.potE) * f[2];
    eg[2] -= T(tp.potw0) * f[2];
    if (tp.depol) {
      const T r = T(tp.r), ic = T(1.0 / tp.ceps);
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        eg[3] -= d2[3][j];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          eg[3] += ic * (d1[i][i] * d1[j][j] + f[i] * dd(d2, j, i, j) -
                         r * (d1[j][i] * d1[i][j] + f[j] * dd(d2, i, i, j)));
          eg[i] += d1[3][i] * d1[j][j] - d1[3][j] * d1[j][i] - f[j] * dd(d2, 3, i, j);
          eg[i] += r * (d1[3][i] * d1[j][j] - d1[3][j] * d1[j][i] + f[j] * dd(d2, 3, i, j));
        }
#pragma unroll
        for (int i = 0; i < NMAG; ++i) eg[i] -= r * f[i] * d2[3][j];
      }
    }
    if (s.band) {
#pragma unroll
      for (int a = 0; a < NF; ++a) eg[a] = T(0);
    }
    project(tp, eg, f);
  }

  // kernels.py:124-216, without the gv factor.
  template <typename T>
  __device__ __forceinline__ static T energy(const TP& tp, const Geom& g, const SiteInfo& s, const T f[NF],
                                             const T d1[NF][2]) {
    T e = T(0);
#pragma unroll
    for (int a = 0; a < NMAG; ++a) e += T(0.5) * (d1[a][0] * d1[a][0] + d1[a][1] * d1[a][1]);
    const T Dm[2][2] = {{T(tp.D00), T(tp.D10)}, {T(tp.D01), T(tp.D11)}};
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
#pragma unroll
        for (int k = 0; k < NMAG; ++k) {
#pragma unroll
          for (int l = 0; l < NMAG; ++l) {
            if (levi(j, k, l) != 0.0) e += Dm[i][j] * T(levi(j, k, l)) * f[k] * d1[l][i];
          }
        }
      }
    }
    if (tp.depol) {
      const T r = T(tp.r);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < NMAG; ++j) e += r * f[j] * d1[j][i] * d1[3][i] / T(2);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          e += d1[i][i] * f[j] * d1[3][j] / T(2);
          e -= r * f[j] * d1[i][j] * d1[3][i] / T(2);
        }
      }
    }
    const T mz = f[2];
    e += T(0.5) * T(tp.potE) * (T(1) - mz * mz);
    e += T(0.5) * T(tp.potw0) * (T(1) - mz * mz);
    return e;
  }

  // 0: n.(d_x n x d_y n)/(4 pi) (kernels.py:254-287);
  // 1: -(d_i n_i d_j n_j + n_i d_i d_j n_j - (e3/e1)(d_i n_j d_j n_i + n_j d_i d_j n_i)) (kernels.py:325-362).
  template <typename T>
  __device__ __forceinline__ static T charge(int kind, const TP& tp, const Geom& g, const SiteInfo& s,
                                             const T f[NF], const T d1[NF][2], const T d2[NF][3]) {
    if (kind == 0) {
      const T cx0 = d1[1][0] * d1[2][1] - d1[2][0] * d1[1][1];
      const T cx1 = d1[2][0] * d1[0][1] - d1[0][0] * d1[2][1];
      const T cx2 = d1[0][0] * d1[1][1] - d1[1][0] * d1[0][1];
      return (f[0] * cx0 + f[1] * cx1 + f[2] * cx2) * T(1.0 / (4.0 * M_PI));
    }
    const T r = T(tp.r);
    T c = T(0);
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        c -= d1[i][i] * d1[j][j] + f[i] * dd(d2, j, i, j) - r * (d1[j][i] * d1[i][j] + f[j] * dd(d2, i, i, j));
      }
    }
    return c;
  }

  // Bit k set: init mode k (ground 0, initial 1, skyrmions 2, vortex 3) is available on the device.
  static constexpr unsigned kInitModes = 3u;
  // Device initial conditions (liquid_crystal/initial_config.py): ground and
  // the centred texture; Phi = 0.
  __device__ __forceinline__ static bool init_site(const InitArgs& a, const Geom& g, int x, int y, double f[NF]) {
    return a.mode != kInitSkyrmions && unit_vector_init<NF>(a, g, x, y, f);
  }

  // No auxiliary densities.
  __host__ __device__ static constexpr int aux_components(int kind) { return 0; }
  template <typename T>
  __device__ __forceinline__ static T aux(int kind, int comp, const TP& tp, const Geom& g, const SiteInfo& s,
                                          const T f[NF], const T d1[NF][2], const T d2[NF][3]) {
    return T(0);
  }

  // The energy is linear in Phi, so the exact-adjoint flow has no stationary
  // point in Phi when depol is on; the director gradient is projected.
  template <typename T>
  __device__ __forceinline__ static void flow_fix(const TP& tp, const T f[NF], T eg[NF]) {
    project(tp, eg, f);
  }

  // kernels.py:48-84.
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

  // kernels.py:86-120.
  template <typename T>
  __device__ __forceinline__ static void project(const TP& tp, T v[NF], const T f[NF]) {
    const T lm = v[0] * f[0] + v[1] * f[1] + v[2] * f[2];
    v[0] -= lm * f[0];
    v[1] -= lm * f[1];
    v[2] -= lm * f[2];
  }
};

}  // namespace ss
