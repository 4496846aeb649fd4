// FP64 C++/OpenMP restatement of the reference's chiral-magnet relaxation path
// -- TEST INFRASTRUCTURE ONLY (the CPU baseline of bench.py and a second,
// independent check of the NumPy oracle).  It is deliberately unfused, with
// the reference's buffer structure: per stage, first and second derivatives
// of every component into scratch arrays (core/derivatives.py:13-115), the
// theory's gradient point (chiral_magnet/kernels.py:397-511) writing EG and
// l = -dt EG, the RK4 stage arrays k/Temp (core/integrator.py:50-96), the
// constrained combine (:151-189), energy (chiral_magnet/kernels.py:124-251)
// summed in 1024-wide trees (core/utils.py:187-228,316-341), the velocity
// kill (integrator.py:339-343) and the signed max of the stage-4 EG.
//
// Layout: F[a*X*Y + x*Y + y]; scratch d1[(a*2+i)*N + site], d2[(a*3+k)*N + site].
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <string.h>

#include <vector>

namespace {

struct P {
  int X, Y, halo, nf;
  long long N;
  double lsx, lsy, gv, dt;
  double coupK, couph, coupmu;
  double D00, D01, D10, D11;
  int demag, unit_mag, killkinen;
};

const double C12 = 1.0 / 12.0, C23 = 2.0 / 3.0, C43 = 4.0 / 3.0, C52 = 5.0 / 2.0;

void derivs(const P& p, const double* F, double* d1, double* d2) {
  const int X = p.X, Y = p.Y, h = p.halo;
  const long long N = p.N;
#pragma omp parallel for schedule(static)
  for (int x = 0; x < X; ++x) {
    for (int a = 0; a < p.nf; ++a) {
      const double* f = F + a * N;
      for (int y = 0; y < Y; ++y) {
        const long long s = (long long)x * Y + y;
        const bool inx = x > h - 1 && x < X - h, iny = y > h - 1 && y < Y - h;
        double dx = 0, dy = 0, xx = 0, yy = 0, xy = 0;
        if (inx) {
          dx = (C12 * f[s - 2 * Y] - C23 * f[s - Y] + C23 * f[s + Y] - C12 * f[s + 2 * Y]) / p.lsx;
          xx = (-C12 * f[s - 2 * Y] + C43 * f[s - Y] - C52 * f[s] + C43 * f[s + Y] - C12 * f[s + 2 * Y]) / (p.lsx * p.lsx);
        }
        if (iny) {
          dy = (C12 * f[s - 2] - C23 * f[s - 1] + C23 * f[s + 1] - C12 * f[s + 2]) / p.lsy;
          yy = (-C12 * f[s - 2] + C43 * f[s - 1] - C52 * f[s] + C43 * f[s + 1] - C12 * f[s + 2]) / (p.lsy * p.lsy);
        }
        if (x > h && y > h && x < X - h && y < Y - h) {
          const double diag = (-C12 * f[s - 2 * Y - 2] + C43 * f[s - Y - 1] - C52 * f[s] + C43 * f[s + Y + 1] -
                               C12 * f[s + 2 * Y + 2]) / (2.0 * p.gv);
          const double xh = (-C12 * f[s - 2 * Y] + C43 * f[s - Y] - C52 * f[s] + C43 * f[s + Y] - C12 * f[s + 2 * Y]) /
                            (2.0 * p.lsx * p.lsx);
          const double yh = (-C12 * f[s - 2] + C43 * f[s - 1] - C52 * f[s] + C43 * f[s + 1] - C12 * f[s + 2]) /
                            (2.0 * p.lsy * p.lsy);
          xy = diag - xh - yh;
        }
        d1[(a * 2 + 0) * N + s] = dx;
        d1[(a * 2 + 1) * N + s] = dy;
        d2[(a * 3 + 0) * N + s] = xx;
        d2[(a * 3 + 1) * N + s] = yy;
        d2[(a * 3 + 2) * N + s] = xy;
      }
    }
  }
}

double levi(int i, int j, int k) {
  if ((i == 0 && j == 1 && k == 2) || (i == 1 && j == 2 && k == 0) || (i == 2 && j == 0 && k == 1)) return 1.0;
  if ((i == 0 && j == 2 && k == 1) || (i == 2 && j == 1 && k == 0) || (i == 1 && j == 0 && k == 2)) return -1.0;
  return 0.0;
}

void gradient(const P& p, const double* T, const double* d1, const double* d2, double* EG, double* l) {
  const long long N = p.N;
  const int X = p.X, Y = p.Y, h = p.halo;
#pragma omp parallel for schedule(static)
  for (int x = 0; x < X; ++x) {
    for (int y = 0; y < Y; ++y) {
      const long long s = (long long)x * Y + y;
      double eg[4] = {0, 0, 0, 0};
      auto D1 = [&](int a, int i) { return d1[(a * 2 + i) * N + s]; };
      auto D2 = [&](int a, int k) { return d2[(a * 3 + k) * N + s]; };
      for (int i = 0; i < 3; ++i) {
        double lap = 0.0;
        lap += D2(i, 0);
        lap += D2(i, 1);
        eg[i] -= lap;
      }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          double e = levi(0, i, j);
          if (e != 0.0) eg[i] += 2.0 * p.D00 * e * D1(j, 0);
          if (e != 0.0 && p.D01 != 0.0) eg[i] += 2.0 * p.D01 * e * D1(j, 1);
          e = levi(1, i, j);
          if (e != 0.0 && p.D10 != 0.0) eg[i] += 2.0 * p.D10 * e * D1(j, 0);
          if (e != 0.0 && p.D11 != 0.0) eg[i] += 2.0 * p.D11 * e * D1(j, 1);
        }
      const double mz = T[2 * N + s];
      eg[2] -= p.couph;
      eg[2] -= 2.0 * p.coupK * mz;
      if (p.demag)
        for (int i = 0; i < 2; ++i) {
          eg[i] += D1(3, i);
          eg[3] += p.coupmu * D1(i, i) - D2(3, i);
        }
      if (x < h || x > X - h - 1 || y < h || y > Y - h - 1) eg[0] = eg[1] = eg[2] = eg[3] = 0.0;
      double lm = 0.0;
      for (int a = 0; a < 3; ++a) lm += eg[a] * T[a * N + s];
      for (int a = 0; a < 3; ++a) eg[a] -= lm * T[a * N + s];
      for (int a = 0; a < 4; ++a) {
        EG[a * N + s] = eg[a];
        l[a * N + s] = -p.dt * eg[a];
      }
    }
  }
}

double energy(const P& p, const double* F, double* d1, double* d2, std::vector<double>& en) {
  derivs(p, F, d1, d2);
  const long long N = p.N;
  const double Dm[2][2] = {{p.D00, p.D10}, {p.D01, p.D11}};
#pragma omp parallel for schedule(static)
  for (long long s = 0; s < N; ++s) {
    auto D1 = [&](int a, int i) { return d1[(a * 2 + i) * N + s]; };
    double e = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < 2; ++i) e += 0.5 * D1(a, i) * D1(a, i);
    for (int i = 0; i < 2; ++i) {
      for (int j = 0; j < 2; ++j) {
        const double Dij = Dm[i][j];
        if (Dij == 0.0) continue;
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) {
            const double eps = levi(j, k, l);
            if (eps != 0.0) e += Dij * eps * F[k * N + s] * D1(l, i);
          }
      }
      if (p.demag) e -= F[3 * N + s] * D1(i, i) + D1(3, i) * D1(3, i) / (2.0 * p.coupmu);
    }
    const double mz = F[2 * N + s];
    e += p.couph * (1.0 - mz);
    e += p.coupK * (1.0 - mz * mz);
    en[s] = e * p.gv;
  }
  // 1024-wide trees with halving stride, then the partials in order.
  const long long nb = (N + 1023) / 1024;
  std::vector<double> part(nb);
#pragma omp parallel for schedule(static)
  for (long long b = 0; b < nb; ++b) {
    double buf[1024];
    for (int t = 0; t < 1024; ++t) {
      const long long i = b * 1024 + t;
      buf[t] = i < N ? en[i] : 0.0;
    }
    for (int st = 512; st > 0; st >>= 1)
      for (int t = 0; t < st; ++t) buf[t] += buf[t + st];
    part[b] = buf[0];
  }
  double sum = 0.0;
  for (long long b = 0; b < nb; ++b) sum += part[b];
  return sum;
}

}  // namespace

// nsteps reference steps (or fewer: with tol > 0 the run stops after the
// first step whose max |stage-4 EG| is below tol); energies, errs (signed
// max, the reference's err) and errs_abs (max |EG|) get one entry per step
// taken (any may be null).  Returns the number of steps taken, or -1.
static int chiral_relax(const int32_t* pi, int ni, const double* pf, int nfl, double* F, double* V,
                        double prev_energy, int nsteps, double tol, double* energies, double* errs, double* errs_abs,
                        int nthreads) {
  if (ni < 16 || nfl < 9 || pi[4] != 4 || pi[10] != 3 || pi[2] < 2) return -1;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  P p;
  p.X = pi[0]; p.Y = pi[1]; p.halo = pi[2]; p.nf = 4;
  p.N = (long long)p.X * p.Y;
  p.lsx = pf[2]; p.lsy = pf[3]; p.gv = pf[4]; p.dt = pf[5];
  p.coupK = pf[6]; p.couph = pf[7]; p.coupmu = pf[8];
  p.D00 = p.D01 = p.D10 = p.D11 = 0.0;
  if (pi[11]) { p.D00 = -1; p.D11 = -1; }
  else if (pi[12]) { p.D01 = -1; p.D10 = 1; }
  else if (pi[13]) { p.D00 = 1; p.D11 = -1; }
  else if (pi[14]) { const double s = 0.7071067811865476; p.D00 = -s; p.D01 = -s; p.D10 = s; p.D11 = -s; }
  p.demag = pi[15]; p.unit_mag = pi[9]; p.killkinen = pi[7];
  const long long M = 4 * p.N;
  std::vector<double> d1(2 * M), d2(3 * M), EG(M), T(M), en(p.N);
  std::vector<double> k[4], l[4];
  for (int i = 0; i < 4; ++i) { k[i].assign(M, 0.0); l[i].assign(M, 0.0); }
  double E = prev_energy;
  const double dt = p.dt;
  for (int step = 0; step < nsteps; ++step) {
    const double fac[4] = {0.0, 0.5, 0.5, 1.0};
    for (int st = 0; st < 4; ++st) {
      const double* lin = st == 0 ? V : l[st - 1].data();
      const double* kp = st == 0 ? F : k[st - 1].data();
      double* ko = k[st].data();
      const double c = fac[st];
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < M; ++i) {
        ko[i] = dt * (V[i] + c * lin[i]);
        T[i] = F[i] + c * kp[i];
      }
      derivs(p, T.data(), d1.data(), d2.data());
      gradient(p, T.data(), d1.data(), d2.data(), EG.data(), l[st].data());
    }
#pragma omp parallel for schedule(static)
    for (long long s = 0; s < p.N; ++s) {
      for (int a = 0; a < 4; ++a) {
        const long long i = a * p.N + s;
        V[i] += (1.0 / 6.0) * (l[0][i] + 2.0 * l[1][i] + 2.0 * l[2][i] + l[3][i]);
        F[i] += (1.0 / 6.0) * (k[0][i] + 2.0 * k[1][i] + 2.0 * k[2][i] + k[3][i]);
      }
      if (p.unit_mag) {
        double n2 = 0.0;
        for (int a = 0; a < 3; ++a) n2 += F[a * p.N + s] * F[a * p.N + s];
        const double n = sqrt(n2);
        if (n != 0.0)
          for (int a = 0; a < 3; ++a) F[a * p.N + s] /= n;
        double lm = 0.0;
        for (int a = 0; a < 3; ++a) lm += V[a * p.N + s] * F[a * p.N + s];
        for (int a = 0; a < 3; ++a) V[a * p.N + s] -= lm * F[a * p.N + s];
      }
    }
    const double En = energy(p, F, d1.data(), d2.data(), en);
    if (p.killkinen && En > E) memset(V, 0, sizeof(double) * M);
    double mx = -1.0e300, ma = 0.0;
    for (long long i = 0; i < M; ++i) {
      if (mx < EG[i]) mx = EG[i];
      if (!(fabs(EG[i]) <= ma)) ma = fabs(EG[i]);  // NaN stays
    }
    E = En;
    if (energies) energies[step] = En;
    if (errs) errs[step] = mx;
    if (errs_abs) errs_abs[step] = ma;
    if (tol > 0.0 && ma < tol) return step + 1;
  }
  return nsteps;
}

extern "C" int cref_chiral_relax(const int32_t* pi, int ni, const double* pf, int nfl, double* F, double* V,
                                 double prev_energy, int nsteps, double* energies, double* errs, int nthreads) {
  return chiral_relax(pi, ni, pf, nfl, F, V, prev_energy, nsteps, 0.0, energies, errs, nullptr, nthreads) < 0 ? -1
                                                                                                             : 0;
}

// As cref_chiral_relax, stopping after the first step whose max |stage-4 EG|
// is below tol; returns the steps taken (or -1).
extern "C" int cref_chiral_relax_tol(const int32_t* pi, int ni, const double* pf, int nfl, double* F, double* V,
                                     double prev_energy, int nsteps, double tol, double* energies, double* errs,
                                     double* errs_abs, int nthreads) {
  return chiral_relax(pi, ni, pf, nfl, F, V, prev_energy, nsteps, tol, energies, errs, errs_abs, nthreads);
}

extern "C" int cref_max_threads(void) { return omp_get_max_threads(); }
