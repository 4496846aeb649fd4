// Streaming form of the fused RK4 passes.
//
// A workgroup owns a strip of 64 or 128 columns (y; one or two wavefronts
// across) by ``sx`` rows (x) and marches down it RB rows at a time.  The stage
// field of the rows it needs (RB computed rows plus the 2-row halo either
// side) lives in an LDS ring of NR >= RB + 4 rows (row r in slot
// (r - x0 + 2) mod NR); after each group of RB rows is computed, the next RB
// rows are written into slots no longer read.  The global loads for those
// rows are issued one iteration ahead into registers, so they are in flight
// while the current rows are computed.  Compared with the tile form, rows are
// fetched once per strip instead of (TX + 4)/TX times, and the ring is small
// enough for several resident workgroups per CU.
//
// Per site the arithmetic is the tile form's (same stencils, same order, same
// theory calls), so the two forms agree to rounding (contraction of the same
// expressions may differ between the two inlined contexts); the FP64 partial
// sums differ in order.
#pragma once

#include "common.h"
#include "stencil.h"

namespace ss {

// Buffers of one fused pass (tile and streaming forms).
struct PassBufs {
  const void* F;
  const void* V;
  const void* l1;
  const void* l2;
  const void* l3;
  void* lout;      // pass 0: EG; passes 1-3: l_pass; fused passes 2+3: l2; fused 4+1: l1 of the next step
  void* lout2;     // fused passes 2+3: l3
  void* Fout;      // pass 4
  void* Vout;      // pass 4
  double* part0;   // pass 1: energy partial per tile; pass 4: signed max partial
  double* part1;   // pass 4: max |EG| partial
  double* part2;   // fused 4+1: energy partial of the new field
  const DevState* st;
  float* Flo;      // compensated FP32 form, pass 4: low words of the field (read and written in place)
  // Passes 2+3 of the 4+1 schedule on small lattices (fin_n > 0): the
  // previous step's finalize folded in (see fold_finalize).
  const double* fin_p0;
  const double* fin_p1;
  const double* fin_p2;
  DevState* fin_st;
  double* histE;
  double* histErr;
  int fin_n, hist_cap, killkinen, parity;
};

// The finalize of the 4+1 schedule (err of step n, energy of F_{n+1} and the
// arrest decision) folded into the passes 2+3 of step n+1: every workgroup
// reduces the same partials in the same order, so all reach the same kill
// decision with no atomics; the workgroup of tile 0 (``writer``) writes the
// arrest state.  The energy it compares with is eslot[parity ^ 1], the
// previous field's (parity = that of F_{n+1}), which no workgroup of this
// launch writes.  Same arithmetic as k_fin_err_energy.  All threads of the
// block call it.
template <int NT>
__device__ __forceinline__ bool fold_finalize(const PassBufs& b, bool writer, double* red) {
  double m0 = -1.0e300, m1 = 0.0, s = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < b.fin_n; i += NT) {
    m0 = nmax(m0, b.fin_p0[i]);
    m1 = nmax(m1, b.fin_p1[i]);
    const double y = b.fin_p2[i] - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
  }
  const double a = block_max<NT>(m0, red);
  const double e = block_max<NT>(m1, red);
  const double E = block_sum<NT>(s, red);
  __shared__ int s_kill;  // the reductions are valid in thread 0 only: it broadcasts the decision
  DevState* st = b.fin_st;
  if (threadIdx.x == 0) s_kill = (b.killkinen && E > st->eslot[b.parity ^ 1]) ? 1 : 0;
  if (writer && threadIdx.x == 0) {
    const bool kill = s_kill != 0;
    st->err_max = a;
    st->err_absmax = e;
    const int step = st->step;
    if (step < b.hist_cap) b.histErr[step] = a;
    st->step = step + 1;
    st->last_energy = E;
    st->kill = kill ? 1 : 0;
    st->prev_energy = E;
    st->eslot[b.parity] = E;
    if (step < b.hist_cap) b.histE[step] = E;
  }
  __syncthreads();
  return s_kill != 0;
}

// Ring of NFR components (NF for one stage field, 2 NF when two stage fields
// share a pass).
// Shape code SC: bits 0-1 = most rows per wave and iteration (1 where the
// register prefetch of two rows would not fit the register file); bit 2 = the
// strip is two wavefronts (128 columns) wide, the workgroup's 4 waves split as
// 2 across the columns x 2 down the rows (the y-halo columns are then shared
// by twice as many sites), else one wavefront (64 columns) with all 4 waves
// down the rows.
template <typename R, int NFR, int SC = 2>
struct StreamShape {
  static constexpr int RPM = SC & 3;
  static constexpr int CW = (SC & 4) ? 2 : 1;  // wavefronts across the columns
  static constexpr int TYW = TY * CW;          // strip width in columns
  static constexpr int RW = kWaves / CW;       // wavefronts down the rows
  static constexpr int RY = TYW + 4;
  static constexpr int kBytesRow = NFR * RY * (int)sizeof(R);
  static constexpr int nr_for(int rb) { return (rb + 4 <= 8) ? 8 : 16; }
  // Two rows per wave and iteration when that ring fits 40 KiB, else one.
  static constexpr int RPW = (RPM >= 2 && nr_for(2 * RW) * kBytesRow <= 40960) ? 2 : 1;
  static constexpr int RB = RW * RPW;  // rows computed per iteration
  static constexpr int NR = nr_for(RB);  // ring rows: a power of two >= RB + 4
  static constexpr bool kFits = NR * kBytesRow <= 65536;
  static_assert(RB + 4 <= NR, "ring too small");
};

// A copy of v defined here, in program order.  The held prefetch values are
// kept this way: a plain assignment is a rename, whose copy the compiler
// places at the loop latch, after the next group's loads, so those loads
// cannot reuse the registers of the values they replace and go to
// temporaries copied over, which waits for them.
template <typename T>
__device__ __forceinline__ T pinned(T v) {
  asm volatile("" : "+v"(v));
  return v;
}

// Column wave (0 .. CW-1) and row wave (0 .. RW-1) of the calling thread.
template <class S>
__device__ __forceinline__ int col_wave() {
  return (threadIdx.x >> 6) % S::CW;
}
template <class S>
__device__ __forceinline__ int row_wave() {
  return (threadIdx.x >> 6) / S::CW;
}

template <typename R, int NFR, int SC = 2>
struct alignas(16) Ring {
  using S = StreamShape<R, NFR, SC>;
  static_assert(S::kFits, "ring does not fit in 64 KiB of LDS");
  R v[NFR][S::NR][S::RY];
};

// The compensated 1+2 sweep's ring, where four parts fit, holds the field
// joined in FP64, (double)hi + (double)lo, formed once per site when a row
// enters the ring: NF planes of doubles over float components 2NF .. 4NF-1
// (after F and Temp2), plane a at [a][slot][column].
template <typename R, int NFR, int SC>
__device__ __forceinline__ double* joined_plane(Ring<R, NFR, SC>& q, int a) {
  using S = StreamShape<R, NFR, SC>;
  static_assert(sizeof(R) == 4 && NFR % 4 == 0, "joined planes: FP32 ring of four parts");
  return reinterpret_cast<double*>(&q.v[NFR / 2][0][0]) + a * S::NR * S::RY;
}
template <typename R, int NFR, int SC>
__device__ __forceinline__ const double* joined_plane(const Ring<R, NFR, SC>& q, int a) {
  using S = StreamShape<R, NFR, SC>;
  static_assert(sizeof(R) == 4 && NFR % 4 == 0, "joined planes: FP32 ring of four parts");
  return reinterpret_cast<const double*>(&q.v[NFR / 2][0][0]) + a * S::NR * S::RY;
}

struct SLaunch {
  int sx;       // rows per strip (a multiple of RB)
  int nsy;      // strips along y
  int nblocks;  // strips
  int per_xcd, grid;
};

template <typename R, int NF>
inline SLaunch make_slaunch(const Geom& g, int sx, int tyw) {
  SLaunch L;
  L.sx = sx;
  L.nsy = (g.Y + tyw - 1) / tyw;
  L.nblocks = ((g.X + sx - 1) / sx) * L.nsy;
  L.per_xcd = (L.nblocks + 7) / 8;
  L.grid = 8 * L.per_xcd;
  return L;
}

// Stencil constants in the arithmetic type, converted once per thread.
template <typename T>
struct StencilK {
  T ilx, ily, ilx2, ily2, i2gv, i2lx2, i2ly2;
  __device__ __forceinline__ explicit StencilK(const Geom& g)
      : ilx(T(g.inv_lsx)), ily(T(g.inv_lsy)), ilx2(T(g.inv_lsx2)), ily2(T(g.inv_lsy2)), i2gv(T(g.inv_2gv)),
        i2lx2(T(g.inv_2lsx2)), i2ly2(T(g.inv_2lsy2)) {}
};

// Element ``a*N + off`` of a field array as (plane base, 32-bit byte offset):
// the base is uniform (scalar registers) and the offset a 32-bit VGPR, the
// form a global load with a scalar base takes.  The host keeps every plane
// below 4 GiB and nf*N below 2^31 for this form.
template <typename R>
__device__ __forceinline__ const char* plane(const void* base, int a, unsigned N) {
  return (const char*)base + (size_t)a * N * sizeof(R);
}
template <typename R>
__device__ __forceinline__ R ld(const char* plane_base, unsigned boff) {
  return *(const R*)(plane_base + boff);
}
// A value read once, at the reading thread's own site (no neighbour reads it
// from L2 later): a non-temporal load leaves the L2 lines to the halo rows.
template <typename R>
__device__ __forceinline__ R ldnt(const char* plane_base, unsigned boff) {
  return __builtin_nontemporal_load((const R*)(plane_base + boff));
}
// The prefetch loads of the streaming form go through buffer resources with
// hardware range checking: a site outside the lattice gets the offset kOutside,
// beyond the array's size, and the load returns 0.  The value is then final as
// loaded (no select after the load, which would make the wave wait for it at
// once), and an array read as all zeros (V after an arrest) is a resource of
// size 0.  A resource spans one array of nf planes, so the host keeps
// nf*X*Y*sizeof(R) below 2^31 bytes for this form; plane a is at the scalar
// offset a*X*Y*sizeof(R).
constexpr unsigned kOutside = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t array_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
template <typename R>
__device__ __forceinline__ R bld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(R) == 4)
    return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  else
    return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// Outputs of a streaming pass are read by the next launch only: non-temporal
// stores keep them from displacing the input rows (halo lines still to be read
// by neighbouring strips) in L2.  Measured at 4096^2: pass 4 380 -> 351 us,
// the 2+3 sweep 298 -> 280 us (FP32).
template <typename R>
__device__ __forceinline__ void st(void* base, int a, unsigned N, unsigned boff, R v) {
  __builtin_nontemporal_store(v, (R*)((char*)base + (size_t)a * N * sizeof(R) + boff));
}

// Stencils of component a at ring column cy; sl[0..4] are the ring slots of
// rows x-2 .. x+2.  Same formulas and masks as window_derivs; INTERIOR (every
// mask true) drops the selects.
template <bool INTERIOR, typename T, typename R, int NFR, int SC>
__device__ __forceinline__ void ring_derivs(const Ring<R, NFR, SC>& q, int a, const int sl[5], int cy, const StencilK<T>& K,
                                            const SiteInfo& s, T& f, T d1[2], T d2[3]) {
  const T c12 = T(SS_C12), c23 = T(SS_C23), c43 = T(SS_C43);
  const T c = T(q.v[a][sl[2]][cy]);
  const T xm2 = T(q.v[a][sl[0]][cy]), xm1 = T(q.v[a][sl[1]][cy]);
  const T xp1 = T(q.v[a][sl[3]][cy]), xp2 = T(q.v[a][sl[4]][cy]);
  const T ym2 = T(q.v[a][sl[2]][cy - 2]), ym1 = T(q.v[a][sl[2]][cy - 1]);
  const T yp1 = T(q.v[a][sl[2]][cy + 1]), yp2 = T(q.v[a][sl[2]][cy + 2]);
  const T sxx = c43 * ((xm1 - c) + (xp1 - c)) - c12 * ((xm2 - c) + (xp2 - c));
  const T syy = c43 * ((ym1 - c) + (yp1 - c)) - c12 * ((ym2 - c) + (yp2 - c));
  const T sx1 = (c12 * (xm2 - xp2) + c23 * (xp1 - xm1)) * K.ilx;
  const T sy1 = (c12 * (ym2 - yp2) + c23 * (yp1 - ym1)) * K.ily;
  f = c;
  d1[0] = (INTERIOR || s.inx) ? sx1 : T(0);
  d1[1] = (INTERIOR || s.iny) ? sy1 : T(0);
  d2[0] = (INTERIOR || s.inx) ? sxx * K.ilx2 : T(0);
  d2[1] = (INTERIOR || s.iny) ? syy * K.ily2 : T(0);
  if (INTERIOR || s.inxy) {
    const T dm2 = T(q.v[a][sl[0]][cy - 2]), dm1 = T(q.v[a][sl[1]][cy - 1]);
    const T dp1 = T(q.v[a][sl[3]][cy + 1]), dp2 = T(q.v[a][sl[4]][cy + 2]);
    const T sdd = c43 * ((dm1 - c) + (dp1 - c)) - c12 * ((dm2 - c) + (dp2 - c));
    d2[2] = sdd * K.i2gv - sxx * K.i2lx2 - syy * K.i2ly2;
  } else {
    d2[2] = T(0);
  }
}

// ring_derivs in R and, from the same ring values (one set of LDS reads), the
// value and first derivatives in FP64: the inputs of the variation and of the
// energy density of pass 1 (the FP32 form converts the FP32 ring values).
template <bool INTERIOR, typename R, int NFR, int SC>
__device__ __forceinline__ void ring_derivs_d1(const Ring<R, NFR, SC>& q, int a, const int sl[5], int cy,
                                               const StencilK<R>& K, const StencilK<double>& Kd, const SiteInfo& s,
                                               R& f, R d1[2], R d2[3], double& fd, double d1d[2]) {
  const R c12 = R(SS_C12), c23 = R(SS_C23), c43 = R(SS_C43);
  const R c = q.v[a][sl[2]][cy];
  const R xm2 = q.v[a][sl[0]][cy], xm1 = q.v[a][sl[1]][cy];
  const R xp1 = q.v[a][sl[3]][cy], xp2 = q.v[a][sl[4]][cy];
  const R ym2 = q.v[a][sl[2]][cy - 2], ym1 = q.v[a][sl[2]][cy - 1];
  const R yp1 = q.v[a][sl[2]][cy + 1], yp2 = q.v[a][sl[2]][cy + 2];
  const R sxx = c43 * ((xm1 - c) + (xp1 - c)) - c12 * ((xm2 - c) + (xp2 - c));
  const R syy = c43 * ((ym1 - c) + (yp1 - c)) - c12 * ((ym2 - c) + (yp2 - c));
  const R sx1 = (c12 * (xm2 - xp2) + c23 * (xp1 - xm1)) * K.ilx;
  const R sy1 = (c12 * (ym2 - yp2) + c23 * (yp1 - ym1)) * K.ily;
  f = c;
  d1[0] = (INTERIOR || s.inx) ? sx1 : R(0);
  d1[1] = (INTERIOR || s.iny) ? sy1 : R(0);
  d2[0] = (INTERIOR || s.inx) ? sxx * K.ilx2 : R(0);
  d2[1] = (INTERIOR || s.iny) ? syy * K.ily2 : R(0);
  if (INTERIOR || s.inxy) {
    const R dm2 = q.v[a][sl[0]][cy - 2], dm1 = q.v[a][sl[1]][cy - 1];
    const R dp1 = q.v[a][sl[3]][cy + 1], dp2 = q.v[a][sl[4]][cy + 2];
    const R sdd = c43 * ((dm1 - c) + (dp1 - c)) - c12 * ((dm2 - c) + (dp2 - c));
    d2[2] = sdd * K.i2gv - sxx * K.i2lx2 - syy * K.i2ly2;
  } else {
    d2[2] = R(0);
  }
  const double e12 = SS_C12, e23 = SS_C23;
  fd = (double)c;
  const double tx = (e12 * ((double)xm2 - (double)xp2) + e23 * ((double)xp1 - (double)xm1)) * Kd.ilx;
  const double ty = (e12 * ((double)ym2 - (double)yp2) + e23 * ((double)yp1 - (double)ym1)) * Kd.ily;
  d1d[0] = (INTERIOR || s.inx) ? tx : 0.0;
  d1d[1] = (INTERIOR || s.iny) ? ty : 0.0;
}

// FP64 value and first derivatives from v(slot, column), the compensated
// form's field hi + lo in FP64.
template <bool INTERIOR, class VF>
__device__ __forceinline__ void d1_fp64(VF v, const int sl[5], int cy, const StencilK<double>& Kd,
                                        const SiteInfo& s, double& f, double d1[2]) {
  const double c12 = SS_C12, c23 = SS_C23;
  const double xm2 = v(sl[0], cy), xm1 = v(sl[1], cy), xp1 = v(sl[3], cy), xp2 = v(sl[4], cy);
  const double ym2 = v(sl[2], cy - 2), ym1 = v(sl[2], cy - 1), yp1 = v(sl[2], cy + 1), yp2 = v(sl[2], cy + 2);
  f = v(sl[2], cy);
  const double sx1 = (c12 * (xm2 - xp2) + c23 * (xp1 - xm1)) * Kd.ilx;
  const double sy1 = (c12 * (ym2 - yp2) + c23 * (yp1 - ym1)) * Kd.ily;
  d1[0] = (INTERIOR || s.inx) ? sx1 : 0.0;
  d1[1] = (INTERIOR || s.iny) ? sy1 : 0.0;
}

// hi in ring component ah and lo in al, joined at every read.
template <bool INTERIOR, typename R, int NFR, int SC>
__device__ __forceinline__ void ring_d1_hilo(const Ring<R, NFR, SC>& q, int ah, int al, const int sl[5], int cy,
                                             const StencilK<double>& Kd, const SiteInfo& s, double& f,
                                             double d1[2]) {
  d1_fp64<INTERIOR>([&](int r, int c) { return (double)q.v[ah][r][c] + (double)q.v[al][r][c]; }, sl, cy, Kd, s, f,
                    d1);
}

// hi + lo from the joined FP64 plane a (the same values ring_d1_hilo forms).
template <bool INTERIOR, typename R, int NFR, int SC>
__device__ __forceinline__ void ring_d1_joined(const Ring<R, NFR, SC>& q, int a, const int sl[5], int cy,
                                               const StencilK<double>& Kd, const SiteInfo& s, double& f,
                                               double d1[2]) {
  using S = StreamShape<R, NFR, SC>;
  const double* pl = joined_plane(q, a);
  d1_fp64<INTERIOR>([&](int r, int c) { return pl[r * S::RY + c]; }, sl, cy, Kd, s, f, d1);
}

// Raw inputs of the stage field for the RB rows of one group, per thread: RPW
// interior sites (its column) and HP (site, component) pairs of the group's
// 4*RB halo-column sites (ring columns 0, 1, TYW+2, TYW+3), spread one pair
// per thread so that each thread holds one component of a halo site, not all
// NF.  A group starts 2 rows after the first row computed in the next
// iteration, so the interior row of (row wave w, k) is (w*RPW + k - 2) mod RB:
// each thread loads the sites it computes next iteration, except the first
// two compute rows (w*RPW + k < 2), whose loaders hold them one iteration
// longer.  Sites outside the lattice read 0 (range-checked loads), so the
// values are used as loaded and stay in flight across the compute of the
// current rows.
template <typename R, int NF, int RPW, int HP>
struct StreamPrefetch {
  R f[RPW][NF], v[RPW][NF], l[RPW][NF];
  R la[RPW][NF], lc[RPW][NF];  // fused 4+1: l1 and l3; 3+4: l2 in la
  R hf[HP], hv[HP], hl[HP], hla[HP];  // halo pairs: F, V, l, la of one component
};

// Halo-column (site, component) pairs of a group: 4*RB sites x NF components,
// pair p = threadIdx.x + j*kThreads (site p / NF, component p mod NF).
template <class S, int NF>
struct HaloPairs {
  static constexpr int kPairs = 4 * S::RB * NF;
  static constexpr int HP = (kPairs + kThreads - 1) / kThreads;
};
// Row r (in the group), ring column c and component a of pair j of this
// thread; false when the thread has no such pair.
template <class S, int NF>
__device__ __forceinline__ bool halo_pair(int j, int& r, int& c, int& a) {
  const int p = threadIdx.x + j * kThreads;
  const int site = p / NF;
  a = p - site * NF;
  r = site >> 2;
  const int qd = site & 3;
  c = qd < 2 ? qd : S::TYW + qd;
  return p < HaloPairs<S, NF>::kPairs;
}

template <typename R, int NF, int NFR, int SC>
using PrefetchFor = StreamPrefetch<R, NF, StreamShape<R, NFR, SC>::RPW, HaloPairs<StreamShape<R, NFR, SC>, NF>::HP>;

// Ring slot row r and column c of the RPW interior sites (k < RPW) and, for
// the 4+1 sweep, of its halo-column site (k == RPW, threads < 4*RB; one site
// with all components, loaded when stored).
template <typename R, int NFR, int SC>
__device__ __forceinline__ void group_site(int k, int& r, int& c) {
  using S = StreamShape<R, NFR, SC>;
  const int lane = threadIdx.x & 63, rw = row_wave<S>();
  if (k < S::RPW) {
    r = (rw * S::RPW + k + S::RB - 2) & (S::RB - 1);
    c = col_wave<S>() * TY + lane + 2;
  } else {
    r = threadIdx.x >> 2;
    const int qd = threadIdx.x & 3;
    c = qd < 2 ? qd : S::TYW + qd;
  }
}

// Issue the loads of the group of rows starting at rb (range-checked: a site
// outside the lattice, or V after an arrest, reads 0; no load sits behind a
// branch).  PASS
// selects the arrays: 0/1 F; 2 and 12 F, V; 3 and 23 F, V, l1; 4 F, V, l2; 34
// F, V, l1, l2 (into la); 41 F, V, l2, l1, l3; pass 1 of the compensated form (PASS 1 with a two-half ring) F
// and its low words (into l); passes 1+2 of the compensated form (PASS 12 with a three-part ring) F, V and the
// low words (into l).
template <int PASS, typename R, int NF, int NFR, int SC>
__device__ __forceinline__ void group_load(PrefetchFor<R, NF, NFR, SC>& p, const PassBufs& b, const Geom& g, int rb,
                                           int y0, bool kill, bool live = true) {
  using S = StreamShape<R, NFR, SC>;
  const void* L = (PASS == 4 || PASS == 41) ? b.l2 : b.l1;
  // the low words of the compensated passes 1 and 1+2 travel in l
  constexpr bool kLo = (PASS == 1 && NFR == 2 * NF) || (PASS == 12 && NFR >= 3 * NF);
  const unsigned NB = (unsigned)g.N * (unsigned)sizeof(R), AB = NB * (unsigned)NF;
  const auto rF = array_rsrc(b.F, AB), rV = array_rsrc(b.V, kill ? 0u : AB);
  const auto rL = array_rsrc(kLo ? (const void*)b.Flo : L, AB);
  const auto r1 = array_rsrc(b.l1, AB), r2 = array_rsrc(b.l2, AB), r3 = array_rsrc(b.l3, AB);
#pragma unroll
  for (int k = 0; k < S::RPW; ++k) {
    int r, c;
    group_site<R, NFR, SC>(k, r, c);
    const int x = rb + r, y = y0 - 2 + c;
    const bool in = live && x >= 0 && x < g.X && y >= 0 && y < g.Y;
    const unsigned boff = in ? (unsigned)(x * g.Y + y) * (unsigned)sizeof(R) : kOutside;
#pragma unroll
    for (int a = 0; a < NF; ++a) {
      const unsigned so = (unsigned)a * NB;
      p.f[k][a] = bld<R>(rF, boff, so);
      p.v[k][a] = (PASS >= 2) ? bld<R>(rV, boff, so) : R(0);
      p.l[k][a] = ((PASS >= 3 && PASS != 12) || kLo) ? bld<R>(rL, boff, so) : R(0);
      if (PASS == 41) {
        p.la[k][a] = bld<R>(r1, boff, so);
        p.lc[k][a] = bld<R>(r3, boff, so);
      }
      if (PASS == 34)  // l2 (l1 after an arrest: the caller points b.l2 there)
        p.la[k][a] = bld<R>(r2, boff, so);
    }
  }
  if constexpr (PASS == 41) return;  // the 4+1 sweep loads its halo-column sites when it stores them
#pragma unroll
  for (int j = 0; j < HaloPairs<S, NF>::HP; ++j) {
    int r, c, a;
    const bool mine = halo_pair<S, NF>(j, r, c, a);
    const int x = rb + r, y = y0 - 2 + c;
    const bool in = live && mine && x >= 0 && x < g.X && y >= 0 && y < g.Y;
    // component a is a runtime plane: its offset joins the site's
    const unsigned boff = in ? (unsigned)a * NB + (unsigned)(x * g.Y + y) * (unsigned)sizeof(R) : kOutside;
    p.hf[j] = bld<R>(rF, boff, 0u);
    p.hv[j] = (PASS >= 2) ? bld<R>(rV, boff, 0u) : R(0);
    p.hl[j] = ((PASS >= 3 && PASS != 12) || kLo) ? bld<R>(rL, boff, 0u) : R(0);
    p.hla[j] = (PASS == 34) ? bld<R>(r2, boff, 0u) : R(0);
  }
}

// The field after a step at one site, from the inputs of pass 4 (pointwise:
// F' = Temp4 + dt (l1 - 2 l2 + l3)/6, then the unit constraint), the same
// arithmetic as the combine of stream_rows.  ``in``: the site is in the
// lattice (outside it the value stays 0).
template <class TH, typename R>
__device__ __forceinline__ void new_field(const typename TH::TP& tp, const Geom& g, bool in, const R* f, const R* v,
                                          const R* l1, const R* l2, const R* l3, R dt, R* T4, R* Fn) {
  constexpr int NF = TH::NF;
  const R dt6 = dt * R(1.0 / 6.0);
#pragma unroll
  for (int a = 0; a < NF; ++a) {
    T4[a] = f[a] + dt * (v[a] + R(0.5) * l2[a]);
    Fn[a] = T4[a] + dt6 * (l1[a] - R(2) * l2[a] + l3[a]);
  }
  if (TH::kUnitConstraint && g.unit_mag && in) TH::template normalize<R>(tp, Fn);
}

// Compensated FP32 form: the field is hi + lo, hi the stored FP32 field (what
// every stencil reads) and lo a second FP32 word per component holding the
// part of the field below hi's rounding.  The step's increment (FP32, relative
// error ~1e-7 of the increment itself) is added to hi + lo in FP64
// (compensated_sum), the unit constraint applied in FP64, and the sum split
// again: hi' = fl32(F'), lo' = fl32(F' - hi') (compensated_split).
// Increments smaller than an FP32 ulp of the field, which the plain form
// rounds away, accumulate in lo until they move hi.
__device__ __forceinline__ double compensated_sum(float hi, float lo, float inc) {
  return ((double)hi + (double)lo) + (double)inc;
}
template <class TH>
__device__ __forceinline__ void compensated_split(const typename TH::TP& tp, const Geom& g, double* Fd, float* Fn,
                                                  float* lon) {
  constexpr int NF = TH::NF;
  if (TH::kUnitConstraint && g.unit_mag) TH::template normalize<double>(tp, Fd);
#pragma unroll
  for (int a = 0; a < NF; ++a) {
    Fn[a] = (float)Fd[a];
    lon[a] = (float)(Fd[a] - (double)Fn[a]);
  }
}

template <int PASS, typename R>
__device__ __forceinline__ R stage_value(R fo, R v, R l, R dt) {
  const R half = R(0.5);
  if (PASS <= 1) return fo;
  if (PASS == 2) return fo + half * (dt * v);
  if (PASS == 3) return fo + half * (dt * (v + half * l));
  return fo + dt * (v + half * l);
}

// Write a group into the ring; slot0 is the ring slot of its first row.  PASS
// 23 writes both stage fields, Temp2 to components 0..NF-1 and Temp3 to
// NF..2NF-1; PASS 12 F and Temp2; PASS 34 Temp3 and Temp4.
// One component a of one site into ring slot ``slot``, column c.
template <int PASS, typename R, int NF, int NFR, int SC>
__device__ __forceinline__ void ring_put(Ring<R, NFR, SC>& q, int slot, int c, int a, R fo, R vv, R lv, R la, R dt) {
  using S = StreamShape<R, NFR, SC>;
  if (PASS == 23) {
    q.v[a][slot][c] = stage_value<2, R>(fo, vv, lv, dt);
    q.v[NF + a][slot][c] = stage_value<3, R>(fo, vv, lv, dt);
  } else if (PASS == 12) {
    q.v[a][slot][c] = fo;
    q.v[NF + a][slot][c] = stage_value<2, R>(fo, vv, R(0), dt);
    if constexpr (NFR == 3 * NF) q.v[2 * NF + a][slot][c] = lv;  // compensated: low words
    if constexpr (NFR == 4 * NF)  // compensated: hi + lo joined in FP64
      joined_plane(q, a)[slot * S::RY + c] = (double)fo + (double)lv;
  } else if (PASS == 34) {
    q.v[a][slot][c] = stage_value<3, R>(fo, vv, lv, dt);
    q.v[NF + a][slot][c] = stage_value<4, R>(fo, vv, la, dt);
  } else {
    q.v[a][slot][c] = stage_value<PASS, R>(fo, vv, lv, dt);
    if constexpr (PASS == 1 && NFR == 2 * NF) q.v[NF + a][slot][c] = lv;  // low words
  }
}

template <int PASS, typename R, int NF, int NFR, int SC>
__device__ __forceinline__ void group_store(Ring<R, NFR, SC>& q, const PrefetchFor<R, NF, NFR, SC>& p, int slot0,
                                            R dt) {
  using S = StreamShape<R, NFR, SC>;
#pragma unroll
  for (int k = 0; k < S::RPW; ++k) {
    int r, c;
    group_site<R, NFR, SC>(k, r, c);
    const int slot = (slot0 + r) & (S::NR - 1);
#pragma unroll
    for (int a = 0; a < NF; ++a)
      ring_put<PASS, R, NF, NFR, SC>(q, slot, c, a, p.f[k][a], p.v[k][a], p.l[k][a], p.la[k][a], dt);
  }
#pragma unroll
  for (int j = 0; j < HaloPairs<S, NF>::HP; ++j) {
    int r, c, a;
    if (!halo_pair<S, NF>(j, r, c, a)) break;
    const int slot = (slot0 + r) & (S::NR - 1);
    ring_put<PASS, R, NF, NFR, SC>(q, slot, c, a, p.hf[j], p.hv[j], p.hl[j], p.hla[j], dt);
  }
}

// Fused 4+1: Temp4 into ring components 0..NF-1 and the new field F' into
// NF..2NF-1, both from the group's loads (rows from rb, columns from y0 - 2);
// the halo-column sites (threads < 4 RB) are loaded here, not prefetched,
// which keeps five arrays of one site out of the prefetch registers.
template <class TH, typename R, int NFR, int SC>
__device__ __forceinline__ void group_store41(Ring<R, NFR, SC>& q, const PrefetchFor<R, TH::NF, NFR, SC>& p,
                                              int slot0, R dt, const typename TH::TP& tp, const PassBufs& b,
                                              const Geom& g, int rb, int y0, bool kill) {
  using S = StreamShape<R, NFR, SC>;
  constexpr int NF = TH::NF;
  const unsigned NP = (unsigned)g.N;
#pragma unroll
  for (int k = 0; k <= S::RPW; ++k) {
    if (k == S::RPW && threadIdx.x >= 4 * S::RB) break;
    int r, c;
    group_site<R, NFR, SC>(k, r, c);
    const int slot = (slot0 + r) & (S::NR - 1);
    const int x = rb + r, y = y0 - 2 + c;
    const bool in = x >= 0 && x < g.X && y >= 0 && y < g.Y;
    R T4[NF], Fn[NF];
    if (k < S::RPW) {
      new_field<TH, R>(tp, g, in, p.f[k], p.v[k], p.la[k], p.l[k], p.lc[k], dt, T4, Fn);
    } else {
      const unsigned boff = in ? (unsigned)(x * g.Y + y) * (unsigned)sizeof(R) : kOutside;
      const unsigned NB = NP * (unsigned)sizeof(R), AB = NB * (unsigned)NF;
      const auto rF = array_rsrc(b.F, AB), rV = array_rsrc(b.V, kill ? 0u : AB);
      const auto r1 = array_rsrc(b.l1, AB), r2 = array_rsrc(b.l2, AB), r3 = array_rsrc(b.l3, AB);
      R f[NF], v[NF], l1[NF], l2[NF], l3[NF];
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        const unsigned so = (unsigned)a * NB;
        f[a] = bld<R>(rF, boff, so);
        v[a] = bld<R>(rV, boff, so);
        l1[a] = bld<R>(r1, boff, so);
        l2[a] = bld<R>(r2, boff, so);
        l3[a] = bld<R>(r3, boff, so);
      }
      new_field<TH, R>(tp, g, in, f, v, l1, l2, l3, dt, T4, Fn);
    }
#pragma unroll
    for (int a = 0; a < NF; ++a) {
      q.v[a][slot][c] = T4[a];
      q.v[NF + a][slot][c] = Fn[a];
    }
  }
}

// The RPW sites of this thread in one iteration (rows c0 + rw*RPW + k, rw its
// row wave; column y), from
// the stage field in ring components AOFF .. AOFF+NF-1.
// Pass 4 takes V and l2 of its sites from kV / kL ([k][a], kill applied), and
// in the 4+1 and 3+4 sweeps (FOFF >= 0, FOFF == -2) also l1 and l3 from kA /
// kC; with FOFF >= 0 it takes the new field F' from ring components FOFF ..
// FOFF+NF-1 (the 4+1 sweep, which builds F' there) instead of forming it.
// Passes 0-3 with FOFF == -3 leave l in kOut ([k][a]) instead of storing it.
// COMP (FP32): pass 4 forms the new field with compensated_sum / _split from
// F and its low words; pass 1 takes the energy of F + low words (ring
// components NF..2NF-1).
template <class TH, typename R, int PASS, bool INTERIOR, int NFR, int AOFF, int FOFF, int SC, bool COMP = false>
__device__ __forceinline__ void stream_rows(const Ring<R, NFR, SC>& q, const PassBufs& b, const typename TH::TP& tp,
                                            const Geom& g, const StencilK<R>& K, const StencilK<double>& Kd, int xs,
                                            int c0, int y, int cy, const R* kV, const R* kL, double& acc0,
                                            double& acc1, const R* kA = nullptr, const R* kC = nullptr,
                                            R* kOut = nullptr, const float* kHL = nullptr) {
  using S = StreamShape<R, NFR, SC>;
  constexpr int NF = TH::NF, NR = S::NR, RPW = S::RPW;
  const R dt = R(g.dt);
  const unsigned NP = (unsigned)g.N;
  const int w = row_wave<S>();
  // Pass 4 in FP32: l1 and l3 of this thread's sites, issued before the
  // stencils so their latency overlaps the variation (the stores of the loop
  // below would otherwise keep the compiler from hoisting them).  In FP64 the
  // registers are not there to spare, and they are loaded at the combine.
  constexpr bool kPre = PASS == 4 && sizeof(R) == 4 && FOFF == -1 && !COMP;
  R pl1[kPre ? RPW : 1][NF], pl3[kPre ? RPW : 1][NF];
  // Compensated pass 4 of the 3+4 sweep: hi and lo of this thread's sites come
  // in kHL (hi [k][a], then lo) from the caller: hi kept from its prefetch, lo
  // loaded before pass 3.  (The separate compensated pass 4 has no registers
  // to spare for them and loads both at the combine.)
  constexpr bool kPreC = PASS == 4 && COMP && FOFF == -2;
  if (kPre) {
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
      const int x = c0 + w * RPW + k;
      const bool in = INTERIOR || (x < g.X && y < g.Y);
      const unsigned boff = in ? (unsigned)(x * g.Y + y) * (unsigned)sizeof(R) : 0u;
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        pl1[k][a] = ldnt<R>(plane<R>(b.l1, a, NP), boff);
        pl3[k][a] = ldnt<R>(plane<R>(b.l3, a, NP), boff);
      }
    }
  }
  ErrMax<R> em;  // pass 4: the stage-4 error of this call's sites
#pragma unroll
  for (int k = 0; k < RPW; ++k) {
    const int x = c0 + w * RPW + k;
    if (INTERIOR || (x < g.X && y < g.Y)) {
      int sl[5];
      const int base = x - xs;  // slot of row x-2
#pragma unroll
      for (int d = 0; d < 5; ++d) sl[d] = (base + d) & (NR - 1);
      SiteInfo s;
      if (INTERIOR) {
        s.x = x;
        s.y = y;
        s.inx = s.iny = s.inxy = true;
        s.band = false;
      } else {
        s = site_info(g, x, y);
      }
      R f[NF], d1[NF][2], d2[NF][3], eg[NF];
      double fd[NF], d1d[NF][2];  // pass 1: FP64 inputs of the energy density, from the same ring reads
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        if constexpr (PASS == 1 && !COMP)
          ring_derivs_d1<INTERIOR, R, NFR>(q, AOFF + a, sl, cy, K, Kd, s, f[a], d1[a], d2[a], fd[a], d1d[a]);
        else
          ring_derivs<INTERIOR, R, R, NFR>(q, AOFF + a, sl, cy, K, s, f[a], d1[a], d2[a]);
      }
      TH::template variation<R>(tp, g, s, f, d1, d2, eg);
      const unsigned boff = (unsigned)(x * g.Y + y) * (unsigned)sizeof(R);
      if (PASS == 0) {
#pragma unroll
        for (int a = 0; a < NF; ++a) st<R>(b.lout, a, NP, boff, eg[a]);
      } else if (PASS <= 3 && FOFF == -3) {  // kept in registers for the pass 4 that follows
#pragma unroll
        for (int a = 0; a < NF; ++a) kOut[k * NF + a] = -dt * eg[a];
      } else if (PASS <= 3) {
#pragma unroll
        for (int a = 0; a < NF; ++a) st<R>(b.lout, a, NP, boff, -dt * eg[a]);
      }
      if (PASS == 1) {
        if constexpr (COMP) {
          // compensated form: the energy of hi + lo (low words in the last
          // NF ring components), read after the variation so that the FP64
          // values are not held across it
#pragma unroll
          for (int a = 0; a < NF; ++a) {
            if constexpr (NFR == 4 * NF)
              ring_d1_joined<INTERIOR, R, NFR>(q, a, sl, cy, Kd, s, fd[a], d1d[a]);
            else
              ring_d1_hilo<INTERIOR, R, NFR>(q, AOFF + a, NFR - NF + a, sl, cy, Kd, s, fd[a], d1d[a]);
          }
        }
        acc0 += TH::template energy<double>(tp, g, s, fd, d1d) * g.gv;
      }
      if (PASS == 4) {
        // F' = Temp4 + dt (l1 - 2 l2 + l3)/6 with Temp4 the ring centre.
        R Fn[NF], Vn[NF];
        double Fd[COMP ? NF : 1];  // COMP: F + increment in FP64
        const R sixth = R(1.0 / 6.0);
        const R dt6 = dt * sixth;
#pragma unroll
        for (int a = 0; a < NF; ++a) {
          const R Vc = kV[k * NF + a];
          const R a2 = kL[k * NF + a];
          R a1, a3;
          if constexpr (FOFF >= 0 || FOFF == -2) {  // 4+1 sweep: kept from its front loads; 3+4: l3 just formed
            a1 = kA[k * NF + a];
            a3 = kC[k * NF + a];
          } else if constexpr (kPre) {  // issued before the stencils
            a1 = pl1[k][a];
            a3 = pl3[k][a];
          } else {  // FP64: issued at the combine, where a non-temporal load measured slower
            a1 = ld<R>(plane<R>(b.l1, a, NP), boff);
            a3 = ld<R>(plane<R>(b.l3, a, NP), boff);
          }
          const R a4 = -dt * eg[a];
          Vn[a] = Vc + sixth * (a1 + R(2) * a2 + R(2) * a3 + a4);
          if constexpr (kPreC)
            Fd[a] = compensated_sum(kHL[k * NF + a], kHL[(RPW + k) * NF + a],
                                    dt * (Vc + R(0.5) * a2) + dt6 * (a1 - R(2) * a2 + a3));
          else if constexpr (COMP)
            Fd[a] = compensated_sum(ldnt<float>(plane<float>(b.F, a, NP), boff),
                                    ldnt<float>(plane<float>(b.Flo, a, NP), boff),
                                    dt * (Vc + R(0.5) * a2) + dt6 * (a1 - R(2) * a2 + a3));
          else if constexpr (FOFF >= 0)
            Fn[a] = q.v[FOFF + a][sl[2]][cy];
          else
            Fn[a] = f[a] + dt6 * (a1 - R(2) * a2 + a3);
          em.add(eg[a]);
        }
        if constexpr (COMP) {
          static_assert(sizeof(R) == 4 && FOFF < 0, "compensated combine: FP32 pass 4 only");
          float lon[NF];
          compensated_split<TH>(tp, g, Fd, Fn, lon);
#pragma unroll
          for (int a = 0; a < NF; ++a) st<float>(b.Flo, a, NP, boff, lon[a]);
          if (TH::kUnitConstraint && g.unit_mag) TH::template project<R>(tp, Vn, Fn);
        } else if (TH::kUnitConstraint && g.unit_mag) {
          if (FOFF < 0) TH::template normalize<R>(tp, Fn);
          TH::template project<R>(tp, Vn, Fn);
        }
#pragma unroll
        for (int a = 0; a < NF; ++a) {
          st<R>(b.Fout, a, NP, boff, Fn[a]);
          st<R>(b.Vout, a, NP, boff, Vn[a]);
        }
      }
    }
  }
  if (PASS == 4) em.fold(acc0, acc1);
}

// One strip of one pass (PASS 0..4 as in stage_tile; 23 = passes 2 and 3 in
// one sweep: both stage fields depend only on F, V and l1, so they are built
// from the same loads into two halves of the ring and give l2 -> lout and
// l3 -> lout2; 41 = pass 4 and pass 1 of the next step in one sweep: the new
// field F' = Temp4 + dt (l1 - 2 l2 + l3)/6 is pointwise in the inputs of pass
// 4, so Temp4 and F' are built from the same loads into the two halves of the
// ring, pass 4 runs on the first half (F', V' -> Fout, Vout) and pass 1 on the
// second (l1 of the next step -> lout, its energy partial -> part2)); 12 =
// passes 1 and 2 (Temp2 = F + dt V/2 does not depend on l1: F and Temp2 in the
// two halves, l1 -> lout with the energy partial, l2 -> lout2; V is read
// without the arrest, which the caller applies afterwards by taking l2 = l1,
// exactly the l2 of an arrested V); 34 = passes 3 and 4 (Temp3 from F, V, l1
// and Temp4 from F, V, l2 in the two halves; l3 stays in registers for the
// combine at the same sites; with COMP the combine is the compensated one).
// Partials: one per strip.
template <class TH, typename R, int PASS, int CW = 1, bool COMP = false>
struct StreamRing {
  // the compensated 1+2 sweep: F, Temp2 and hi + lo joined in FP64 (two
  // float parts) where that fits 64 KiB, else F, Temp2 and the low words
  static constexpr int kSC = (PASS == 41 || PASS == 34 || PASS == 12 ? 1 : 2) | (CW == 2 ? 4 : 0);
  static constexpr bool kJoined =
      PASS == 12 && COMP && sizeof(R) == 4 &&
      StreamShape<R, 4 * TH::NF, kSC>::NR * StreamShape<R, 4 * TH::NF, kSC>::kBytesRow <= 65536;
  static constexpr int NFR =
      kJoined ? 4 * TH::NF
      : (PASS == 12 && COMP) ? 3 * TH::NF
      : (PASS == 23 || PASS == 41 || PASS == 12 || PASS == 34 || (PASS == 1 && COMP)) ? 2 * TH::NF
                                                                                     : TH::NF;
  // the 4+1 and 3+4 sweeps prefetch five and four arrays: one row per wave and iteration
  static constexpr int SC = kSC;
  using shape = StreamShape<R, NFR, SC>;
  using type = Ring<R, NFR, SC>;
};

template <class TH, typename R, int PASS, int CW, bool COMP = false>
__device__ __forceinline__ void stream_strip(typename StreamRing<TH, R, PASS, CW, COMP>::type& q, double* red,
                                             const PassBufs& b, const typename TH::TP& tp, const Geom& g, int xs,
                                             int sx, int y0, bool kill, int t) {
  constexpr int NF = TH::NF, NFR = StreamRing<TH, R, PASS, CW, COMP>::NFR;
  constexpr int SC = StreamRing<TH, R, PASS, CW, COMP>::SC;
  using S = StreamShape<R, NFR, SC>;
  constexpr int RB = S::RB, NR = S::NR, RPW = S::RPW;
  const R dt = R(g.dt);
  const int lane = threadIdx.x & 63;
  const int y = y0 + col_wave<S>() * TY + lane;
  const int cy = col_wave<S>() * TY + lane + 2;
  const int nit = sx / RB;
  (void)kill;
  const StencilK<R> K(g);
  const StencilK<double> Kd(g);
  const unsigned NP = (unsigned)g.N;

  // Prologue: rows xs-2 .. xs+RB+1 straight into ring slots 0 .. RB+3.  Row
  // xs - 2 + j lives in slot j mod NR throughout.
  for (int idx = threadIdx.x; idx < (RB + 4) * S::RY; idx += kThreads) {
    const int r = idx / S::RY, c = idx - r * S::RY;
    const int gx = xs - 2 + r, gy = y0 - 2 + c;
    const bool in = gx >= 0 && gx < g.X && gy >= 0 && gy < g.Y;
    const unsigned boff = in ? (unsigned)(gx * g.Y + gy) * (unsigned)sizeof(R) : 0u;
    if constexpr (PASS == 41) {
      R fo[NF], vv[NF], l1v[NF], l2v[NF], l3v[NF], T4[NF], Fn[NF];
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        const R f0 = ld<R>(plane<R>(b.F, a, NP), boff), v0 = ld<R>(plane<R>(b.V, a, NP), boff);
        const R a1 = ld<R>(plane<R>(b.l1, a, NP), boff), a2 = ld<R>(plane<R>(b.l2, a, NP), boff);
        const R a3 = ld<R>(plane<R>(b.l3, a, NP), boff);
        fo[a] = in ? f0 : R(0);
        vv[a] = (in && !kill) ? v0 : R(0);
        l1v[a] = in ? a1 : R(0);
        l2v[a] = in ? a2 : R(0);
        l3v[a] = in ? a3 : R(0);
      }
      new_field<TH, R>(tp, g, in, fo, vv, l1v, l2v, l3v, dt, T4, Fn);
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        q.v[a][r][c] = T4[a];
        q.v[NF + a][r][c] = Fn[a];
      }
      continue;
    }
#pragma unroll
    for (int a = 0; a < NF; ++a) {
      const R fo = ld<R>(plane<R>(b.F, a, NP), boff);
      const R v = (PASS >= 2 && !kill) ? ld<R>(plane<R>(b.V, a, NP), boff) : R(0);  // arrested: V not read
      const R l = (PASS == 3 || PASS == 23 || PASS == 34) ? ld<R>(plane<R>(b.l1, a, NP), boff)
                : PASS == 4                               ? ld<R>(plane<R>(b.l2, a, NP), boff)
                                                          : R(0);
      if (PASS == 12) {
        q.v[a][r][c] = in ? fo : R(0);
        q.v[NF + a][r][c] = in ? stage_value<2, R>(fo, v, R(0), dt) : R(0);
      } else if (PASS == 34) {
        const R l2 = ld<R>(plane<R>(b.l2, a, NP), boff);
        q.v[a][r][c] = in ? stage_value<3, R>(fo, v, l, dt) : R(0);
        q.v[NF + a][r][c] = in ? stage_value<4, R>(fo, v, l2, dt) : R(0);
      } else if (PASS == 23) {
        q.v[a][r][c] = in ? stage_value<2, R>(fo, v, l, dt) : R(0);
        q.v[NF + a][r][c] = in ? stage_value<3, R>(fo, v, l, dt) : R(0);
      } else {
        q.v[a][r][c] = in ? stage_value<PASS, R>(fo, v, l, dt) : R(0);
      }
      if constexpr ((PASS == 1 || PASS == 12) && COMP) {
        const R lo = ld<R>(plane<R>(b.Flo, a, NP), boff);
        if constexpr (PASS == 12 && NFR == 4 * NF)  // hi + lo joined in FP64
          joined_plane(q, a)[r * S::RY + c] = in ? (double)fo + (double)lo : 0.0;
        else  // low words in the last ring part
          q.v[NFR - NF + a][r][c] = in ? lo : R(0);
      }
    }
  }
  // Pass 4: V and l2 of the sites this thread computes in the current
  // iteration (keep) and, for the first two compute rows, in the next (hold).
  const int w = row_wave<S>();
  R keepV[RPW][NF], keepL[RPW][NF], holdV[RPW][NF], holdL[RPW][NF];
  // 4+1 sweep: l1 and l3 of the same sites (its front loads carry them already)
  R keepA[RPW][NF], keepC[RPW][NF], holdA[RPW][NF], holdC[RPW][NF];
  // compensated 3+4 sweep: hi of the same sites, for the combine (a second
  // read of hi there, rows after its prefetch, misses L2 about half the time)
  constexpr bool kKeepF = PASS == 34 && COMP;
  R keepF[RPW][NF], holdF[RPW][NF];
  if (PASS == 4 || PASS == 41 || PASS == 34) {
#pragma unroll
    for (int k = 0; k < RPW; ++k) {
      const int xk = xs + w * RPW + k, xh = xs + RB + w * RPW + k;
      const bool ink = xk < g.X && y < g.Y, inh = xh < g.X && y < g.Y && (w * RPW + k < 2);
      const unsigned bk = ink ? (unsigned)(xk * g.Y + y) * (unsigned)sizeof(R) : 0u;
      const unsigned bh = inh ? (unsigned)(xh * g.Y + y) * (unsigned)sizeof(R) : 0u;
#pragma unroll
      for (int a = 0; a < NF; ++a) {
        const R v0 = kill ? R(0) : ld<R>(plane<R>(b.V, a, NP), bk);
        const R v1 = kill ? R(0) : ld<R>(plane<R>(b.V, a, NP), bh);
        const R q0 = ld<R>(plane<R>(b.l2, a, NP), bk), q1 = ld<R>(plane<R>(b.l2, a, NP), bh);
        keepV[k][a] = ink ? v0 : R(0);
        keepL[k][a] = ink ? q0 : R(0);
        holdV[k][a] = inh ? v1 : R(0);
        holdL[k][a] = inh ? q1 : R(0);
        if constexpr (kKeepF) {
          const R f0 = ld<R>(plane<R>(b.F, a, NP), bk), f1 = ld<R>(plane<R>(b.F, a, NP), bh);
          keepF[k][a] = ink ? f0 : R(0);
          holdF[k][a] = inh ? f1 : R(0);
        }
        if (PASS == 34) {
          const R a0 = ld<R>(plane<R>(b.l1, a, NP), bk), a1 = ld<R>(plane<R>(b.l1, a, NP), bh);
          keepA[k][a] = ink ? a0 : R(0);
          holdA[k][a] = inh ? a1 : R(0);
        }
        if (PASS == 41) {
          const R a0 = ld<R>(plane<R>(b.l1, a, NP), bk), c0v = ld<R>(plane<R>(b.l3, a, NP), bk);
          const R a1 = ld<R>(plane<R>(b.l1, a, NP), bh), c1v = ld<R>(plane<R>(b.l3, a, NP), bh);
          keepA[k][a] = ink ? a0 : R(0);
          keepC[k][a] = ink ? c0v : R(0);
          holdA[k][a] = inh ? a1 : R(0);
          holdC[k][a] = inh ? c1v : R(0);
        }
      }
    }
  }
  PrefetchFor<R, NF, NFR, SC> pf;
  if (nit > 1) group_load<PASS, R, NF, NFR, SC>(pf, b, g, xs + RB + 2, y0, kill);

  PassBufs b2 = b, b3 = b;
  b3.lout = b.lout2;
  double acc0 = (PASS == 4 || PASS == 41 || PASS == 34) ? -1.0e300 : 0.0;
  double acc1 = 0.0, acc2 = 0.0, unused = 0.0;
  for (int it = 0; it < nit; ++it) {
    __syncthreads();  // ring holds rows xs + it*RB - 2 .. xs + it*RB + RB + 1
    const int c0 = xs + it * RB;
    // every mask true for all rows and columns of this iteration
    const bool interior =
        (c0 > g.halo) && (c0 + RB - 1 < g.X - g.halo) && (y0 > g.halo) && (y0 + S::TYW - 1 < g.Y - g.halo);
    if (PASS == 23) {
      if (interior) {
        stream_rows<TH, R, 2, true, NFR, 0, -1>(q, b2, tp, g, K, Kd, xs, c0, y, cy, nullptr, nullptr, acc0, acc1);
        stream_rows<TH, R, 3, true, NFR, NF, -1>(q, b3, tp, g, K, Kd, xs, c0, y, cy, nullptr, nullptr, acc0, acc1);
      } else {
        stream_rows<TH, R, 2, false, NFR, 0, -1>(q, b2, tp, g, K, Kd, xs, c0, y, cy, nullptr, nullptr, acc0, acc1);
        stream_rows<TH, R, 3, false, NFR, NF, -1>(q, b3, tp, g, K, Kd, xs, c0, y, cy, nullptr, nullptr, acc0, acc1);
      }
    } else if (PASS == 12) {
      // pass 1 from F (ring 0..NF-1: l1 and the energy), pass 2 from Temp2 (ring NF..2NF-1)
      // (compensated form: pass 1 takes the energy of F + its low words, ring 2NF..3NF-1)
      if (interior) {
        stream_rows<TH, R, 1, true, NFR, 0, -1, SC, COMP>(q, b2, tp, g, K, Kd, xs, c0, y, cy, nullptr, nullptr, acc0,
                                                          acc1);
        stream_rows<TH, R, 2, true, NFR, NF, -1>(q, b3, tp, g, K, Kd, xs, c0, y, cy, nullptr, nullptr, unused, unused);
      } else {
        stream_rows<TH, R, 1, false, NFR, 0, -1, SC, COMP>(q, b2, tp, g, K, Kd, xs, c0, y, cy, nullptr, nullptr, acc0,
                                                           acc1);
        stream_rows<TH, R, 2, false, NFR, NF, -1>(q, b3, tp, g, K, Kd, xs, c0, y, cy, nullptr, nullptr, unused,
                                                  unused);
      }
    } else if (PASS == 34) {
      // pass 3 from Temp3 (ring 0..NF-1, l3 into registers), then pass 4 from Temp4 (ring NF..2NF-1)
      R l3r[RPW][NF];
#pragma unroll
      for (int k = 0; k < RPW; ++k)
#pragma unroll
        for (int a = 0; a < NF; ++a) l3r[k][a] = R(0);
      // compensated combine: hi of this iteration's sites (kept from the
      // prefetch) and lo, issued before pass 3 so that its latency hides
      // behind both variations
      float hl[COMP ? 2 * RPW : 1][NF];
      if constexpr (COMP) {
#pragma unroll
        for (int k = 0; k < RPW; ++k) {
          const int x = c0 + w * RPW + k;
          const unsigned boff = (x < g.X && y < g.Y) ? (unsigned)(x * g.Y + y) * (unsigned)sizeof(R) : 0u;
#pragma unroll
          for (int a = 0; a < NF; ++a) {
            hl[k][a] = keepF[k][a];
            hl[RPW + k][a] = ldnt<float>(plane<float>(b.Flo, a, NP), boff);
          }
        }
      }
      if (interior) {
        stream_rows<TH, R, 3, true, NFR, 0, -3>(q, b, tp, g, K, Kd, xs, c0, y, cy, nullptr, nullptr, unused, unused,
                                                nullptr, nullptr, &l3r[0][0]);
        stream_rows<TH, R, 4, true, NFR, NF, -2, SC, COMP>(q, b, tp, g, K, Kd, xs, c0, y, cy, &keepV[0][0],
                                                           &keepL[0][0], acc0, acc1, &keepA[0][0], &l3r[0][0],
                                                           nullptr, &hl[0][0]);
      } else {
        stream_rows<TH, R, 3, false, NFR, 0, -3>(q, b, tp, g, K, Kd, xs, c0, y, cy, nullptr, nullptr, unused, unused,
                                                 nullptr, nullptr, &l3r[0][0]);
        stream_rows<TH, R, 4, false, NFR, NF, -2, SC, COMP>(q, b, tp, g, K, Kd, xs, c0, y, cy, &keepV[0][0],
                                                            &keepL[0][0], acc0, acc1, &keepA[0][0], &l3r[0][0],
                                                            nullptr, &hl[0][0]);
      }
    } else if (PASS == 41) {
      // pass 4 from Temp4 (ring 0..NF-1), then pass 1 of the next step from F' (ring NF..2NF-1)
      if (interior) {
        stream_rows<TH, R, 4, true, NFR, 0, NF>(q, b, tp, g, K, Kd, xs, c0, y, cy, &keepV[0][0], &keepL[0][0], acc0,
                                                acc1, &keepA[0][0], &keepC[0][0]);
        stream_rows<TH, R, 1, true, NFR, NF, -1>(q, b, tp, g, K, Kd, xs, c0, y, cy, nullptr, nullptr, acc2, unused);
      } else {
        stream_rows<TH, R, 4, false, NFR, 0, NF>(q, b, tp, g, K, Kd, xs, c0, y, cy, &keepV[0][0], &keepL[0][0], acc0,
                                                 acc1, &keepA[0][0], &keepC[0][0]);
        stream_rows<TH, R, 1, false, NFR, NF, -1>(q, b, tp, g, K, Kd, xs, c0, y, cy, nullptr, nullptr, acc2, unused);
      }
    } else if (interior) {
      stream_rows<TH, R, PASS, true, NFR, 0, -1, SC, COMP>(q, b, tp, g, K, Kd, xs, c0, y, cy, &keepV[0][0],
                                                           &keepL[0][0], acc0, acc1);
    } else {
      stream_rows<TH, R, PASS, false, NFR, 0, -1, SC, COMP>(q, b, tp, g, K, Kd, xs, c0, y, cy, &keepV[0][0],
                                                            &keepL[0][0], acc0, acc1);
    }
    __syncthreads();  // every read of this iteration's rows is done
    {
      // rows c0+RB+2 .. c0+2RB+1 go to slots (it*RB + RB + 4 ..) mod NR, which
      // no row of iteration it+1 other than themselves maps to.  Done in the
      // last iteration too (nothing reads the ring after it), so that no merge
      // of stored and unstored values follows the block.
      if constexpr (PASS == 41)
        group_store41<TH, R, NFR, SC>(q, pf, (it * RB + RB + 4) & (NR - 1), dt, tp, b, g, c0 + RB + 2, y0, kill);
      else
        group_store<PASS, R, NF, NFR, SC>(q, pf, (it * RB + RB + 4) & (NR - 1), dt);
      if (PASS == 4 || PASS == 41 || PASS == 34) {
        // next iteration's sites: the group just stored, or (first two compute
        // rows) the hold from the group before; the hold takes the new group.
#pragma unroll
        for (int k = 0; k < RPW; ++k) {
          const bool lag = w * RPW + k < 2;
#pragma unroll
          for (int a = 0; a < NF; ++a) {
            // 3+4: the prefetch holds l1 in l and l2 in la
            const R nl = PASS == 34 ? pf.la[k][a] : pf.l[k][a];
            const R nv = pf.v[k][a];
            keepV[k][a] = pinned(lag ? holdV[k][a] : nv);
            keepL[k][a] = pinned(lag ? holdL[k][a] : nl);
            holdV[k][a] = pinned(nv);
            holdL[k][a] = pinned(nl);
            if (PASS == 34) {
              const R na = pf.l[k][a];
              keepA[k][a] = pinned(lag ? holdA[k][a] : na);
              holdA[k][a] = pinned(na);
            }
            if constexpr (kKeepF) {
              const R nf = pf.f[k][a];
              keepF[k][a] = pinned(lag ? holdF[k][a] : nf);
              holdF[k][a] = pinned(nf);
            }
            if (PASS == 41) {
              const R na = pf.la[k][a], nc = pf.lc[k][a];
              keepA[k][a] = pinned(lag ? holdA[k][a] : na);
              keepC[k][a] = pinned(lag ? holdC[k][a] : nc);
              holdA[k][a] = pinned(na);
              holdC[k][a] = pinned(nc);
            }
          }
        }
      }
      // issued in every iteration (after the last group all of it reads as
      // outside, with no memory access), so the prefetch registers carry no
      // merge of loaded and kept values around the loop
      group_load<PASS, R, NF, NFR, SC>(pf, b, g, c0 + 2 * RB + 2, y0, kill, it + 2 < nit);
    }
  }
  if (PASS == 1 || PASS == 12) {
    const double s = block_sum<kThreads>(acc0, red);
    if (threadIdx.x == 0) b.part0[t] = s;
  } else if (PASS == 4 || PASS == 41 || PASS == 34) {
    const double m0 = block_max<kThreads>(acc0, red);
    if (threadIdx.x == 0) b.part0[t] = m0;
    const double m1 = block_max<kThreads>(acc1, red);
    if (threadIdx.x == 0) b.part1[t] = m1;
    if (PASS == 41) {
      const double e = block_sum<kThreads>(acc2, red);
      if (threadIdx.x == 0) b.part2[t] = e;
    }
  }
}

template <class TH, typename R, int PASS, int CW, bool COMP>
__device__ __forceinline__ void stream_pass_body(const PassBufs& b, const typename TH::TP& tp, const Geom& g,
                                                 const SLaunch& L) {
  __shared__ typename StreamRing<TH, R, PASS, CW, COMP>::type q;
  __shared__ double red[kWaves];
  const int t = xcd_tile(blockIdx.x, L.nblocks, L.per_xcd);
  if (t < 0) return;  // whole block exits together, before any barrier
  const int sxi = t / L.nsy, syi = t - sxi * L.nsy;
  bool kill;
  if constexpr (PASS == 23)  // the previous step's finalize folded in, or its decision
    kill = b.fin_n > 0 ? fold_finalize<kThreads>(b, t == 0, red) : (b.st->kill != 0);
  else
    kill = (PASS >= 2 && PASS != 12) ? (b.st->kill != 0) : false;  // 1+2 reads V before the arrest decision
  PassBufs ba = b;
  if (PASS == 34 && kill) ba.l2 = b.l1;  // l2 of an arrested V is l1 (Temp2 = F)
  stream_strip<TH, R, PASS, CW, COMP>(q, red, ba, tp, g, sxi * L.sx, L.sx,
                                      syi * StreamRing<TH, R, PASS, CW, COMP>::shape::TYW, kill, t);
}

template <class TH, typename R, int PASS, int CW>
__global__ void __launch_bounds__(kThreads) stream_pass(PassBufs b, typename TH::TP tp, Geom g, SLaunch L) {
  stream_pass_body<TH, R, PASS, CW, false>(b, tp, g, L);
}

// Pass 1 or 4 of the compensated FP32 form, held to the register budget of 3
// waves per SIMD (the plain form's pass 4 budget).
template <class TH, typename R, int PASS, int CW>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3)))
stream_pass_comp(PassBufs b, typename TH::TP tp, Geom g, SLaunch L) {
  stream_pass_body<TH, R, PASS, CW, true>(b, tp, g, L);
}

}  // namespace ss
