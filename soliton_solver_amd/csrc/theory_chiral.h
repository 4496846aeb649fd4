// Chiral magnet (reference: theories/chiral_magnet/kernels.py, params.py).
// Fields m1, m2, m3 and the demagnetising potential psi (nf = 4).
#pragma once

#include "common.h"
#include "init.h"

namespace ss {

struct ChiralMagnet {
  static constexpr int NF = 4;
  static constexpr int NMAG = 3;
  static constexpr bool kUnitConstraint = true;   // do_rk4_kernel with norm + projection (kernels.py:122)
  static constexpr bool kL2Normalize = false;
  static constexpr int kNumCharges = 2;           // 0 skyrmion number, 1 magnetic charge
  static constexpr const char* kName = "chiral_magnet";

  struct TP {
    double coupK, couph, coupmu;   // p_f[6..8]
    double inv_2mu;                // 1/(2 coupmu)
    double D00, D01, D10, D11;     // DMI matrix by flag (kernels.py:171-183)
    int demag;                     // p_i[15]
  };

  // Packing: chiral_magnet/params.py:333-378.
  static int make(const int32_t* pi, int ni, const double* pf, int nfl, TP* tp, char* msg, int cap) {
    if (ni < 16 || nfl < 9) {
      snprintf(msg, cap, "chiral magnet needs p_i[0..15] and p_f[0..8] (got %d, %d)", ni, nfl);
      return -1;
    }
    if (pi[4] != NF || pi[10] != NMAG) {
      snprintf(msg, cap, "chiral magnet kernels are built for number_total_fields=4, "
               "number_magnetization_fields=3 (got %d, %d)", pi[4], pi[10]);
      return -1;
    }
    tp->coupK = pf[6];
    tp->couph = pf[7];
    tp->coupmu = pf[8];
    tp->inv_2mu = 1.0 / (2.0 * pf[8]);
    tp->D00 = tp->D01 = tp->D10 = tp->D11 = 0.0;
    if (pi[11]) { tp->D00 = -1.0; tp->D11 = -1.0; }
    else if (pi[12]) { tp->D01 = -1.0; tp->D10 = 1.0; }
    else if (pi[13]) { tp->D00 = 1.0; tp->D11 = -1.0; }
    else if (pi[14]) {
      const double s = 0.7071067811865476;
      tp->D00 = -s; tp->D01 = -s; tp->D10 = s; tp->D11 = -s;
    }
    tp->demag = pi[15];
    return 0;
  }

  // Levi-Civita symbol, folded at compile time inside unrolled loops.
  __device__ __forceinline__ static constexpr double levi(int i, int j, int k) {
    return ((i == 0 && j == 1 && k == 2) || (i == 1 && j == 2 && k == 0) || (i == 2 && j == 0 && k == 1)) ? 1.0
         : ((i == 0 && j == 2 && k == 1) || (i == 2 && j == 1 && k == 0) || (i == 1 && j == 0 && k == 2)) ? -1.0
         : 0.0;
  }

  // Energy gradient EG at one site (kernels.py:397-511): exchange Laplacian,
  // DMI, Zeeman + anisotropy on m3, demag coupling, boundary-band mask, then
  // the tangent projection against the (unrenormalised) stage field.
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
        const T e0 = T(levi(0, i, j));
        const T e1 = T(levi(1, i, j));
        if (levi(0, i, j) != 0.0) {
          eg[i] += T(2) * D00 * e0 * d1[j][0];
          eg[i] += T(2) * D01 * e0 * d1[j][1];
        }
        if (levi(1, i, j) != 0.0) {
          eg[i] += T(2) * D10 * e1 * d1[j][0];
          eg[i] += T(2) * D11 * e1 * d1[j][1];
        }
      }
    }
    eg[2] -= T(tp.couph);
    eg[2] -= T(2) * T(tp.coupK) * f[2];
    if (tp.demag) {
      const T mu = T(tp.coupmu);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        eg[i] += d1[3][i];
        eg[3] += mu * d1[i][i] - d2[3][i];
      }
    }
    if (s.band) {
#pragma unroll
      for (int a = 0; a < NF; ++a) eg[a] = T(0);
    }
    project(tp, eg, f);
  }

  // Energy density without the gv factor (kernels.py:124-216).
  template <typename T>
  __device__ __forceinline__ static T energy(const TP& tp, const Geom& g, const SiteInfo& s,
                                             const T f[NF], const T d1[NF][2]) {
    T e = T(0);
#pragma unroll
    for (int a = 0; a < NMAG; ++a) e += T(0.5) * (d1[a][0] * d1[a][0] + d1[a][1] * d1[a][1]);
    const T Dm[2][2] = {{T(tp.D00), T(tp.D10)}, {T(tp.D01), T(tp.D11)}};  // Dm[i][j] = D_{ij}
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const T Dij = Dm[i][j];
#pragma unroll
        for (int k = 0; k < NMAG; ++k) {
#pragma unroll
          for (int l = 0; l < NMAG; ++l) {
            if (levi(j, k, l) != 0.0) e += Dij * T(levi(j, k, l)) * f[k] * d1[l][i];
          }
        }
      }
      if (tp.demag) e -= f[3] * d1[i][i] + d1[3][i] * d1[3][i] * T(tp.inv_2mu);
    }
    const T mz = f[2];
    e += T(tp.couph) * (T(1) - mz);
    e += T(tp.coupK) * (T(1) - mz * mz);
    return e;
  }

  // Charge densities without the gv factor.
  // 0: m.(d_x m x d_y m)/(4 pi) (kernels.py:253-291); 1: -(d_x m1 + d_y m2) (kernels.py:328-360).
  template <typename T>
  __device__ __forceinline__ static T charge(int kind, const TP& tp, const Geom& g, const SiteInfo& s,
                                             const T f[NF], const T d1[NF][2], const T d2[NF][3]) {
    if (kind == 0) {
      const T cx0 = d1[1][0] * d1[2][1] - d1[2][0] * d1[1][1];
      const T cx1 = d1[2][0] * d1[0][1] - d1[0][0] * d1[2][1];
      const T cx2 = d1[0][0] * d1[1][1] - d1[1][0] * d1[0][1];
      return (f[0] * cx0 + f[1] * cx1 + f[2] * cx2) * T(1.0 / (4.0 * M_PI));
    }
    return -(d1[0][0] + d1[1][1]);
  }

  // Bit k set: init mode k (ground 0, initial 1, skyrmions 2, vortex 3) is available on the device.
  static constexpr unsigned kInitModes = 7u;
  // Device initial conditions (chiral_magnet/initial_config.py): ground,
  // centred texture, composed solitons; psi = 0.
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

  // Exact-adjoint flow convention: psi is a saddle field, flowed along
  // -coup_mu dE/dpsi as the reference's variation does (kernels.py:503); the
  // magnetisation gradient is projected onto the tangent space.
  template <typename T>
  __device__ __forceinline__ static void flow_fix(const TP& tp, const T f[NF], T eg[NF]) {
    if (tp.demag) eg[3] = -T(tp.coupmu) * eg[3];
    project(tp, eg, f);
  }

  // Unit-norm constraint (kernels.py:47-83): sites with |m| == 0 are left as is.
  template <typename T>
  __device__ __forceinline__ static void normalize(const TP& tp, T f[NF]) {
    T s2 = T(0);
#pragma unroll
    for (int a = 0; a < NMAG; ++a) s2 += f[a] * f[a];
    const T n = sqrt(s2);
    if (n != T(0)) {
      const T inv = T(1) / n;
#pragma unroll
      for (int a = 0; a < NMAG; ++a) f[a] = f[a] * inv;
    }
  }

  // v -= (v.m) m (kernels.py:85-119).
  template <typename T>
  __device__ __forceinline__ static void project(const TP& tp, T v[NF], const T f[NF]) {
    T lm = T(0);
#pragma unroll
    for (int a = 0; a < NMAG; ++a) lm += v[a] * f[a];
#pragma unroll
    for (int a = 0; a < NMAG; ++a) v[a] -= lm * f[a];
  }
};

}  // namespace ss
