// ORACLE — test infrastructure only; the product never includes or links this file.
//
// A forward-mode (dual-number) build of the portable step: the reference's single-source step in Tree order
// with the portable math policy (engine/include/sinksim/kernel/step.hpp:41-54, Order::Tree, MathPortable),
// restated over a scalar type T that carries, beside each value, its derivatives along up to N directions.
// Geometry and constants stay double; the dynamic state, the frame, every partial, pass, head, flow, load and
// the parameters are T. Branches are taken on values, so the values are those of the portable oracle bit for
// bit, and the tangents are the exact derivatives of the discrete program wherever no branch flips.
//
// Parameters (no counterpart in the reference): a vector theta of scale factors. Connection c's area is
// area[c] * theta[pa[c]] and its discharge coefficient coef[c] * theta[pc[c]] (index -1: no factor), the
// product the sweep's kind scales form on the area (engine/include/sinksim/sweep.hpp:49-65).
//
// Derivative conventions (the derivative of the discrete program, not of a smoothed one):
//   - sin/cos: the derivative of sin is the fdlibm cos of the same argument and vice versa;
//   - pow(x, y) with constant y: y * pow(x, y) / x, and 0 at x == 0;
//   - sqrt: 0.5 / sqrt(x), and 0 at x == 0; fabs: the sign of the value (+ at 0);
//   - min/max selections and clamps carry the derivative of the selected operand.
//
// Parity pins (tests/): the values equal the portable oracle record for record; the tangents equal central
// finite differences of the oracle itself on horizons without branch flips.
#pragma once

#include <cmath>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "oracle_step.hpp"

namespace orc {
namespace dual {

template <int N>
struct D {
  double v;
  double d[N];
  D() : v(0) { for (int i = 0; i < N; ++i) d[i] = 0; }
  D(double x) : v(x) { for (int i = 0; i < N; ++i) d[i] = 0; }  // NOLINT: constants lift implicitly
};

template <int N> inline D<N> operator-(const D<N>& a) { D<N> r; r.v = -a.v; for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r; }
template <int N> inline D<N> operator+(const D<N>& a, const D<N>& b) { D<N> r; r.v = a.v + b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> inline D<N> operator-(const D<N>& a, const D<N>& b) { D<N> r; r.v = a.v - b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> inline D<N> operator*(const D<N>& a, const D<N>& b) {
  D<N> r; r.v = a.v * b.v;
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N> inline D<N> operator/(const D<N>& a, const D<N>& b) {
  D<N> r; r.v = a.v / b.v;
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
  return r;
}
template <int N> inline D<N> operator+(const D<N>& a, double b) { D<N> r = a; r.v = a.v + b; return r; }
template <int N> inline D<N> operator+(double a, const D<N>& b) { D<N> r = b; r.v = a + b.v; return r; }
template <int N> inline D<N> operator-(const D<N>& a, double b) { D<N> r = a; r.v = a.v - b; return r; }
template <int N> inline D<N> operator-(double a, const D<N>& b) { D<N> r; r.v = a - b.v; for (int i = 0; i < N; ++i) r.d[i] = -b.d[i]; return r; }
template <int N> inline D<N> operator*(const D<N>& a, double b) { D<N> r; r.v = a.v * b; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b; return r; }
template <int N> inline D<N> operator*(double a, const D<N>& b) { D<N> r; r.v = a * b.v; for (int i = 0; i < N; ++i) r.d[i] = a * b.d[i]; return r; }
template <int N> inline D<N> operator/(const D<N>& a, double b) { D<N> r; r.v = a.v / b; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / b; return r; }
template <int N> inline D<N> operator/(double a, const D<N>& b) {
  D<N> r; r.v = a / b.v;
  for (int i = 0; i < N; ++i) r.d[i] = -r.v * b.d[i] / b.v;
  return r;
}
template <int N> inline D<N>& operator+=(D<N>& a, const D<N>& b) { a = a + b; return a; }
template <int N> inline D<N>& operator-=(D<N>& a, const D<N>& b) { a = a - b; return a; }
template <int N> inline D<N>& operator+=(D<N>& a, double b) { a.v += b; return a; }

#define ORC_DUAL_CMP(op)                                                                     \
  template <int N> inline bool operator op(const D<N>& a, const D<N>& b) { return a.v op b.v; } \
  template <int N> inline bool operator op(const D<N>& a, double b) { return a.v op b; }        \
  template <int N> inline bool operator op(double a, const D<N>& b) { return a op b.v; }
ORC_DUAL_CMP(<)
ORC_DUAL_CMP(>)
ORC_DUAL_CMP(<=)
ORC_DUAL_CMP(>=)
ORC_DUAL_CMP(==)
ORC_DUAL_CMP(!=)
#undef ORC_DUAL_CMP

inline double val(double x) { return x; }
template <int N> inline double val(const D<N>& x) { return x.v; }

inline double dsqrt(double x) { return std::sqrt(x); }
template <int N> inline D<N> dsqrt(const D<N>& x) {
  D<N> r; r.v = std::sqrt(x.v);
  const double k = r.v > 0 ? 0.5 / r.v : 0.0;
  for (int i = 0; i < N; ++i) r.d[i] = x.d[i] * k;
  return r;
}
inline double dfabs(double x) { return std::fabs(x); }
template <int N> inline D<N> dfabs(const D<N>& x) { return x.v < 0 ? -x : x; }

inline double dsin(double x) { return orc_fdl_sin(x); }
inline double dcos(double x) { return orc_fdl_cos(x); }
template <int N> inline D<N> dsin(const D<N>& x) {
  D<N> r; r.v = orc_fdl_sin(x.v);
  const double c = orc_fdl_cos(x.v);
  for (int i = 0; i < N; ++i) r.d[i] = x.d[i] * c;
  return r;
}
template <int N> inline D<N> dcos(const D<N>& x) {
  D<N> r; r.v = orc_fdl_cos(x.v);
  const double s = orc_fdl_sin(x.v);
  for (int i = 0; i < N; ++i) r.d[i] = -(x.d[i] * s);
  return r;
}
inline double dpow(double x, double y) { return orc_fdl_pow(x, y); }
template <int N> inline D<N> dpow(const D<N>& x, double y) {
  D<N> r; r.v = orc_fdl_pow(x.v, y);
  const double k = x.v != 0 ? y * r.v / x.v : 0.0;
  for (int i = 0; i < N; ++i) r.d[i] = x.d[i] * k;
  return r;
}

template <class T> inline T tmax(const T& a, const T& b) { return (b > a) ? b : a; }  // math.hpp:29
template <class T> inline T tmin(const T& a, const T& b) { return (b < a) ? b : a; }  // math.hpp:30

// ---------------------------------------------------------------- records over T (types.hpp:159-182)

template <class T>
struct FrameT {
  T R00, R01, R02, R10, R11, R12, R20, R21, R22, tx, ty, tz, cth;
};
template <class T>
struct BuoyT {
  T V, x, y, z, top;
};
template <class T>
struct PassT {
  T vol, dv, mx, my, mz, zmin, zmax;
};
template <class T>
struct LoadsT {
  T Fz, Mth, Mph, mw, Iw_p, Iw_r;
};

// frame.hpp:10-21
template <class T>
inline FrameT<T> pose(const Body& b, const T& zO, const T& th, const T& ph) {
  const T cth = dcos(th), sth = dsin(th), cph = dcos(ph), sph = dsin(ph);
  FrameT<T> F;
  F.R00 = cth;
  F.R01 = sth * sph;
  F.R02 = sth * cph;
  F.R10 = T(0);
  F.R11 = cph;
  F.R12 = -sph;
  F.R20 = -sth;
  F.R21 = cth * sph;
  F.R22 = cth * cph;
  F.tx = b.XO - (F.R00 * b.g0x + F.R01 * b.g0y + F.R02 * b.g0z);
  F.ty = b.YO - (F.R10 * b.g0x + F.R11 * b.g0y + F.R12 * b.g0z);
  F.tz = zO - (F.R20 * b.g0x + F.R21 * b.g0y + F.R22 * b.g0z);
  F.cth = cth;
  return F;
}

template <class T>
inline T world_z(const FrameT<T>& F, double x, double y, double z) { return F.R20 * x + F.R21 * y + F.R22 * z + F.tz; }

template <class T>
inline HP<T> hp_zero_t() { return HP<T>{T(0), T(0), T(0), T(0), T(-1e9)}; }
template <class T>
inline PP<T> pp_zero_t() { return PP<T>{T(0), T(0), T(0), T(0), T(0), T(1e9), T(-1e9)}; }

template <class T>
inline void absorb_t(HP<T>& a, const HP<T>& b) {
  a.V += b.V; a.Mx += b.Mx; a.My += b.My; a.Mz += b.Mz;
  if (b.top > a.top) a.top = b.top;
}
template <class T>
inline void absorb_t(PP<T>& a, const PP<T>& b) {
  a.vol += b.vol; a.dv += b.dv; a.mx += b.mx; a.my += b.my; a.mz += b.mz;
  if (b.zmin < a.zmin) a.zmin = b.zmin;
  if (b.zmax > a.zmax) a.zmax = b.zmax;
}
template <class P, int L>
inline P halving_t(P (&v)[L]) {  // reduce.hpp:74-79
  for (int off = L / 2; off > 0; off >>= 1)
    for (int l = 0; l < off; ++l) absorb_t(v[l], v[l + off]);
  return v[0];
}

// one column of the buoyancy integral (reduce.hpp:39-49)
template <class T>
inline void column(const Ship& S, const FrameT<T>& F, const T& iR, Index i, HP<T>& p) {
  const double x = S.x[i], y = S.y[i], lo = S.zlo[i], hi = S.zhi[i];
  const T base = F.R20 * x + F.R21 * y + F.tz;
  const T wt = base + F.R22 * hi;
  if (wt > p.top) p.top = wt;
  const T zp = -base * iR;
  if (zp <= lo) return;
  const T s = zp >= hi ? T(hi - lo) : zp - lo;
  const T v = S.dxdy[i] * s;
  p.V += v; p.Mx += v * x; p.My += v * y; p.Mz += v * (lo + 0.5 * s);
}

// one segment of a node pass (reduce.hpp:52-71)
template <class T>
inline void segment(const Ship& S, const FrameT<T>& F, const T& iR, const T& h, Index j, PP<T>& p) {
  const Index i = S.segCol[j];
  const double x = S.x[i], y = S.y[i], a = S.segA[j], b = S.segE[j];
  const T base = F.R20 * x + F.R21 * y + F.tz;
  const T wa = base + F.R22 * a, wb = base + F.R22 * b;
  if (wa < p.zmin) p.zmin = wa;
  if (wb > p.zmax) p.zmax = wb;
  const T zp = (h - base) * iR;
  if (zp <= a) return;
  const double A = S.dxdy[i];
  T L;
  if (zp >= b) {
    L = T(b - a);
  } else {
    L = zp - a;
    p.dv += A;
  }
  const T v = A * L;
  p.vol += v; p.mx += v * x; p.my += v * y; p.mz += v * (a + 0.5 * L);
}

template <class T>
inline BuoyT<T> finish_buoy(const HP<T>& p) {  // hydro.hpp:37-45
  BuoyT<T> o;
  o.V = p.V;
  o.x = p.V > 0 ? p.Mx / p.V : T(0);
  o.y = p.V > 0 ? p.My / p.V : T(0);
  o.z = p.V > 0 ? p.Mz / p.V : T(0);
  o.top = p.top;
  return o;
}

template <class T>
inline PassT<T> finish_pass(const Ship& S, Index n, const T& iR, const PP<T>& p) {  // hydro.hpp:101-109
  const double mu = S.mu[n];
  PassT<T> r;
  r.vol = p.vol * mu;
  r.dv = p.dv * mu * iR;
  r.mx = p.mx * mu; r.my = p.my * mu; r.mz = p.mz * mu;
  r.zmin = p.zmin; r.zmax = p.zmax;
  return r;
}

// hydro.hpp:48-64 (Tree order)
template <class T>
inline BuoyT<T> buoyancy(const Ship& S, const FrameT<T>& F) {
  const T iR = 1.0 / F.R22;
  HP<T> warps[kThreads / kLanes];
  for (int w = 0; w < kThreads / kLanes; ++w) {
    HP<T> ln[kLanes];
    for (int l = 0; l < kLanes; ++l) {
      ln[l] = hp_zero_t<T>();
      for (Index i = w * kLanes + l; i < S.ncol; i += kThreads) column(S, F, iR, i, ln[l]);
    }
    warps[w] = halving_t(ln);
  }
  HP<T> total = warps[0];
  for (int w = 1; w < kThreads / kLanes; ++w) absorb_t(total, warps[w]);
  return finish_buoy(total);
}

// hydro.hpp:112-122 (Tree order)
template <class T>
inline PassT<T> node_pass(const Ship& S, const FrameT<T>& F, Index n, const T& h) {
  const Index s0 = S.segStart[n], s1 = s0 + S.segCount[n];
  const T iR = 1.0 / F.R22;
  PP<T> ln[kLanes];
  for (int l = 0; l < kLanes; ++l) {
    ln[l] = pp_zero_t<T>();
    for (Index j = s0 + l; j < s1; j += kLanes) segment(S, F, iR, h, j, ln[l]);
  }
  return finish_pass(S, n, iR, halving_t(ln));
}

// ---------------------------------------------------------------- the simulation over T

// Model constants a dual simulation needs beyond the oracle's views, all double.
struct Params {
  Index np;
  const Index* pa;  // per connection: index of its area factor, or -1
  const Index* pc;  // per connection: index of its coefficient factor, or -1
  int relaxed;      // 1: the level's derivative as if each solve converged (see relax_level)
};

template <class T>
class Sim {
 public:
  // S, M: the oracle's views over the model; area/tOn/en: the instance's connection tables (sweep applied).
  Sim(const Ship& S, const Net& M, const Params& P, const double* coef, const Index* gmax)
      : S_(S), M_(M), P_(P), coef_(coef) {
    const Index NN = S.nn, NC = M.nc;
    vol.assign(NN, T(0)); level.assign(NN, T(0)); cx.assign(NN, T(0)); cy.assign(NN, T(0)); cz.assign(NN, T(0));
    area_.assign(NN, T(0)); zmin_.assign(NN, T(0)); zmax_.assign(NN, T(0)); hEff_.assign(NN, T(0)); aEff_.assign(NN, T(0));
    acc_.assign(NN, T(0)); excess_.assign(NN, T(0)); deg_.assign(NN, 0); merged_.assign(std::max<Index>(M.ng, 1), 0);
    passes_.assign(*gmax > 0 ? *gmax : 1, PassT<T>{});
    dV_.assign(NC, T(0)); q_.assign(NC, T(0)); over_.assign(std::max<Index>(M.nmon, 1), T(0));
    markZ_.assign(std::max<Index>(M.nmark, 1), T(0));
    ca_.assign(NC, T(0)); cc_.assign(NC, T(0));
    overT_.assign(std::max<Index>(M.nmon, 1), -1); markT_.assign(std::max<Index>(M.nmark, 1), -1);
  }

  // area factors and coefficient factors from theta (sweep.hpp:49-65 for the area; the coefficient likewise)
  void set_params(const T* theta) {
    for (Index c = 0; c < M_.nc; ++c) {
      ca_[c] = P_.pa[c] >= 0 ? M_.area[c] * theta[P_.pa[c]] : T(M_.area[c]);
      cc_[c] = P_.pc[c] >= 0 ? coef_[c] * theta[P_.pc[c]] : T(coef_[c]);
    }
  }

  void restore(double t0, const T* s6, const T* v, const T* l, const T* x, const T* y, const T* z) {
    t = t0; zO = s6[0]; pitch = s6[1]; roll = s6[2]; vz = s6[3]; wth = s6[4]; wph = s6[5];
    for (Index n = 0; n < S_.nn; ++n) { vol[n] = v[n]; level[n] = l[n]; cx[n] = x[n]; cy[n] = y[n]; cz[n] = z[n]; }
    for (auto& a : area_) a = T(0);
    foundered = false; founderT = -1;
    for (auto& o : overT_) o = -1;
    for (auto& m : markT_) m = -1;
  }

  // step.hpp:41-54 with MathPortable and Order::Tree; returns nothing, fills the event-signal hooks
  void step() {
    const FrameT<T> F = pose(M_.body, zO, pitch, roll);
    const BuoyT<T> hb = buoyancy(S_, F);
    solve_levels(F);
    flows_and_volumes(F);
    const LoadsT<T> L = loads(F, hb);
    integrate(L);
    Vb = hb.V; hullTop = hb.top;
    track_events(F);
    frame = F;
  }

  // dynamic state
  double t = 0;
  T zO, pitch, roll, vz, wth, wph;
  std::vector<T> vol, level, cx, cy, cz;
  // outputs of the last step
  T Vb, hullTop;
  FrameT<T> frame;
  bool foundered = false;
  double founderT = -1;
  std::vector<double> overT_, markT_;
  std::vector<T> over_;   // monitor flow sums of the last step
  std::vector<T> markZ_;  // world z of each mark in the last step's frame
  std::vector<T> q_;     // connection flows of the last step

 private:
  Ship S_;
  Net M_;
  Params P_;
  const double* coef_;
  std::vector<T> area_, zmin_, zmax_, hEff_, aEff_, acc_, excess_, dV_, ca_, cc_;
  std::vector<Index> deg_;
  std::vector<Byte> merged_;
  std::vector<PassT<T>> passes_;

  // The relaxed derivative of a level: the Newton solve stops within a volume tolerance, so a level's derivative
  // in the discrete program is that of the last iterate (often the previous step's level, carried unchanged);
  // relaxed, it is that of the root of vol(h) = V at the final level h: the tangent gains one Newton step,
  // d h += (dV - dvol) / a, with a the pass's slope at h floored at the space's minimum effective area amin (the
  // floor the flow limiter puts under the free-surface area, flows.hpp:15-18: near the top or bottom of a space
  // the slope vanishes and 1 / dv would make the linearised flows unstable). The value of h is untouched.
  static bool relax_level(T& h, const T& resid, const T& dv, double amin) {
    if constexpr (std::is_same_v<T, double>) {
      return false;
    } else {
      const double a = jmax(dv.v, amin);
      if (!(a > kSlopeMin)) return false;
      for (int i = 0; i < (int)(sizeof(h.d) / sizeof(double)); ++i) h.d[i] += resid.d[i] / a;
      return true;
    }
  }

  void set_centroid(Index n, const PassT<T>& p) {
    if (p.vol > kCentroidMinVolume) { cx[n] = p.mx / p.vol; cy[n] = p.my / p.vol; cz[n] = p.mz / p.vol; }
  }

  // levels.hpp:13-49 (full = false)
  void solve_single(const FrameT<T>& F, Index n) {
    const T V = vol[n];
    const double Vmax = S_.vmax[n];
    T h = level[n];
    PassT<T> p = node_pass(S_, F, n, h);
    const T lo = p.zmin, hi = p.zmax;
    zmin_[n] = lo; zmax_[n] = hi;
    if (V <= kEmptyVolume) { level[n] = lo; area_[n] = T(0); return; }
    if (V >= Vmax * kFullFraction) {
      p = node_pass(S_, F, n, hi);
      level[n] = hi + tmax(T(0), V - Vmax) / S_.aov[n];
      area_[n] = T(S_.aov[n]);
      set_centroid(n, p);
      return;
    }
    if (h < lo || h > hi) {
      h = tmin(hi, tmax(lo, h));
      p = node_pass(S_, F, n, h);
    }
    const double tol = jmax(1e-3 * Vmax, Real(0.5));
    T a = lo, b = hi;
    for (int it = 0; it < kNewtonFast; ++it) {
      const T err = p.vol - V;
      if (dfabs(err) < tol) break;
      if (err > 0) b = h; else a = h;
      T hn = p.dv > kSlopeMin ? h - err / p.dv : 0.5 * (a + b);
      if (!(hn > a && hn < b)) hn = 0.5 * (a + b);
      h = hn;
      p = node_pass(S_, F, n, h);
    }
    if (P_.relaxed && relax_level(h, V - p.vol, p.dv, S_.amin[n])) p = node_pass(S_, F, n, h);
    level[n] = h; area_[n] = p.dv;
    set_centroid(n, p);
  }

  // levels.hpp:54-162 (full = false)
  void solve_group(const FrameT<T>& F, const Index* members, Index N) {
    PassT<T>* P = passes_.data();
    T V = T(0);
    double Vmax = 0;
    for (Index k = 0; k < N; ++k) { V += vol[members[k]]; Vmax += S_.vmax[members[k]]; }
    bool allWet = true;
    Index firstWet = -1;
    for (Index k = 0; k < N; ++k) {
      if (vol[members[k]] > 0) { if (firstWet < 0) firstWet = k; }
      else allWet = false;
    }
    T h;
    if (allWet) {
      T s = T(0);
      for (Index k = 0; k < N; ++k) s += level[members[k]];
      h = s / double(N);
    } else {
      h = level[members[firstWet >= 0 ? firstWet : N - 1]];
    }
    for (Index k = 0; k < N; ++k) P[k] = node_pass(S_, F, members[k], h);
    for (Index k = 0; k < N; ++k) { zmin_[members[k]] = P[k].zmin; zmax_[members[k]] = P[k].zmax; }
    T lo = P[0].zmin, hi = P[0].zmax;
    for (Index k = 1; k < N; ++k) { if (P[k].zmin < lo) lo = P[k].zmin; if (P[k].zmax > hi) hi = P[k].zmax; }

    if (V <= kEmptyVolume) {
      for (Index k = 0; k < N; ++k) { const Index n = members[k]; vol[n] = T(0); level[n] = P[k].zmin; area_[n] = T(0); }
      return;
    }
    if (V >= Vmax * kFullFraction) {
      const T over = tmax(T(0), V - Vmax);
      double aSum = 0;
      for (Index k = 0; k < N; ++k) aSum += S_.aov[members[k]];
      const T dh = over / aSum;
      T assigned = T(0);
      for (Index k = 0; k < N - 1; ++k) {
        const T v = S_.vmax[members[k]] + dh * S_.aov[members[k]];
        vol[members[k]] = v;
        assigned += v;
      }
      vol[members[N - 1]] = V - assigned;
      for (Index k = 0; k < N; ++k) {
        const Index n = members[k];
        P[k] = node_pass(S_, F, n, hi);
        level[n] = hi + dh; area_[n] = T(S_.aov[n]);
        set_centroid(n, P[k]);
      }
      return;
    }
    if (!(h >= lo && h <= hi)) {
      h = tmin(hi, tmax(lo, h));
      for (Index k = 0; k < N; ++k) P[k] = node_pass(S_, F, members[k], h);
    }
    const double tol = jmax(1e-3 * Vmax, Real(0.5));
    T a = lo, b = hi;
    for (int it = 0; it < kNewtonFast; ++it) {
      T sum = T(0);
      for (Index k = 0; k < N; ++k) sum += P[k].vol;
      const T err = sum - V;
      if (dfabs(err) < tol) break;
      if (err > 0) b = h; else a = h;
      T dv = T(0);
      for (Index k = 0; k < N; ++k) dv += P[k].dv;
      T hn = dv > kSlopeMin ? h - err / dv : 0.5 * (a + b);
      if (!(hn > a && hn < b)) hn = 0.5 * (a + b);
      h = hn;
      for (Index k = 0; k < N; ++k) P[k] = node_pass(S_, F, members[k], h);
    }
    if (P_.relaxed) {
      T sum = T(0), dv = T(0);
      double amin = 0;
      for (Index k = 0; k < N; ++k) { sum += P[k].vol; dv += P[k].dv; amin += S_.amin[members[k]]; }
      if (relax_level(h, V - sum, dv, amin))
        for (Index k = 0; k < N; ++k) P[k] = node_pass(S_, F, members[k], h);
    }
    T tot = T(0);
    for (Index k = 0; k < N; ++k) tot += P[k].vol;
    T assigned = T(0);
    for (Index k = 0; k < N - 1; ++k) {
      T v = tot > 0 ? V * P[k].vol / tot : V / double(N);
      if (v > S_.vmax[members[k]]) v = T(S_.vmax[members[k]]);
      vol[members[k]] = v;
      assigned += v;
    }
    const double vmaxLast = S_.vmax[members[N - 1]];
    if (V - assigned > vmaxLast) {
      const T need = V - vmaxLast;
      T capped = T(0);
      for (Index k = N - 2; k >= 0; --k) {
        T prior = T(0);
        for (Index j = 0; j < k; ++j) prior += vol[members[j]];
        const T v = (need - capped) - prior;
        if (v > S_.vmax[members[k]] && k > 0) {
          vol[members[k]] = T(S_.vmax[members[k]]);
          capped += S_.vmax[members[k]];
          continue;
        }
        vol[members[k]] = v;
        break;
      }
      assigned = T(0);
      for (Index k = 0; k < N - 1; ++k) assigned += vol[members[k]];
    }
    vol[members[N - 1]] = V - assigned;
    for (Index k = 0; k < N; ++k) {
      const Index n = members[k];
      level[n] = h; area_[n] = P[k].dv;
      set_centroid(n, P[k]);
    }
  }

  // levels.hpp:166-194
  void solve_levels(const FrameT<T>& F) {
    for (Index g = 0; g < M_.ng; ++g) {
      const Index N = M_.gCount[g];
      const Index* members = M_.gNodes + M_.gStart[g];
      const Byte mode = M_.gMode[g];
      bool merged = false;
      if (N >= 2) {
        if (mode == 1) {
          merged = true;
        } else if (mode == 2) {
          const T zMw = world_z(F, M_.gpx[g], M_.gpy[g], M_.gpz[g]);
          merged = true;
          for (Index k = 0; k < N; ++k) {
            const Index n = members[k];
            if (!(vol[n] > kEmptyVolume && level[n] > zMw + M_.gMargin)) { merged = false; break; }
          }
        }
      }
      merged_[g] = merged ? 1 : 0;
      if (merged) solve_group(F, members, N);
      else for (Index k = 0; k < N; ++k) solve_single(F, members[k]);
    }
  }

  bool counts(Index c) const {  // flows.hpp:21-26
    if (!M_.en[c]) return false;
    const Index a = M_.ca[c], b = M_.cb[c];
    if (a >= 0 && hEff_[a] < kNoHeadTest && hEff_[b] < kNoHeadTest) return false;
    return true;
  }

  // flows.hpp:38-78 with the connection's area and coefficient carrying their factors
  T flow(const FrameT<T>& F, Index c) {
    if (!M_.en[c] || M_.tOn[c] > t) { q_[c] = T(0); return T(0); }
    if (M_.skipGroup[c] >= 0 && merged_[M_.skipGroup[c]]) { q_[c] = T(0); return T(0); }
    const Index a = M_.ca[c], b = M_.cb[c];
    const T zo = F.R20 * M_.cx[c] + F.R21 * M_.cy[c] + F.R22 * M_.cz[c] + F.tz;
    const T ha = a < 0 ? T(0) : hEff_[a], hbv = hEff_[b];
    const T Ha = ha > zo ? ha - zo : T(0), Hb = hbv > zo ? hbv - zo : T(0);
    const double sq2g = M_.sq2g;
    T q = T(0);
    if (M_.law[c] == 0) {
      const T d = Ha - Hb;
      if (d != 0) q = cc_[c] * ca_[c] * sq2g * d / dsqrt(tmax(dfabs(d), T(kOrificeHeadFloor)));
    } else {
      const T H1 = Ha > Hb ? Ha : Hb;
      if (H1 > 0) {
        const T H2 = Ha > Hb ? Hb : Ha;
        T f = T(1);
        if (H2 > 0) {
          const T r = H2 / H1;
          f = dpow(tmax(T(0), 1.0 - dpow(r, kWeirExponent)), kVillemonteExponent);
        }
        q = cc_[c] * ca_[c] * kWeirCoefficient * sq2g * H1 * dsqrt(H1) * f * (Ha >= Hb ? 1.0 : -1.0);
      }
    }
    if (q == 0) { q_[c] = T(0); return T(0); }
    T dV = q * M_.dt;
    const Index src = dV > 0 ? a : b, dst = dV > 0 ? b : a;
    const T hs = src < 0 ? T(0) : hEff_[src];
    const T hd = dst < 0 ? T(0) : hEff_[dst];
    const T inv = (src < 0 ? T(0) : 1.0 / aEff_[src]) + (dst < 0 ? T(0) : 1.0 / aEff_[dst]);
    const Index degRef = src < 0 ? deg_[dst] : (dst < 0 ? deg_[src] : (deg_[src] > deg_[dst] ? deg_[src] : deg_[dst]));
    const double share = kLimiterShare / double(degRef > 1 ? degRef : 1);
    T lim = inv > 0 ? share * (hs - tmax(hd, zo)) / inv : dfabs(dV);
    if (lim < 0) lim = T(0);
    if (dfabs(dV) > lim) dV = dV > 0 ? lim : -lim;
    q_[c] = dV / M_.dt;
    return dV;
  }

  // flows.hpp:121-134 and 99-111
  void flows_and_volumes(const FrameT<T>& F) {
    const Index NN = S_.nn;
    for (Index n = 0; n < NN; ++n) {
      hEff_[n] = vol[n] > kEmptyVolume ? level[n] : T(kNoHead);
      aEff_[n] = vol[n] >= S_.vmax[n] ? T(S_.aov[n]) : tmax(area_[n], T(S_.amin[n]));
    }
    for (Index n = 0; n < NN; ++n) {
      const Index s = M_.incStart[n], e = s + M_.incCount[n];
      Index d = 0;
      for (Index k = s; k < e; ++k)
        if (counts(M_.incConn[k])) ++d;
      deg_[n] = d;
    }
    for (Index c = 0; c < M_.nc; ++c) dV_[c] = flow(F, c);
    for (Index n = 0; n < NN; ++n) {
      const Index s = M_.incStart[n], e = s + M_.incCount[n];
      T a = T(0);
      for (Index k = s; k < e; ++k) {
        const T& d = dV_[M_.incConn[k]];
        a = M_.incSide[k] ? a + d : a - d;
      }
      acc_[n] = a;
    }
    for (Index n = 0; n < NN; ++n) {
      T v = vol[n] + acc_[n];
      if (v < 0) v = T(0);
      const double vm = S_.vmax[n];
      if (v > vm) {
        const T head = tmax(T(0), -zmax_[n]);
        const T cap = vm + S_.aov[n] * head;
        if (v > cap) v = cap;
      }
      vol[n] = v;
    }
  }

  // body.hpp:11-40
  LoadsT<T> loads(const FrameT<T>& F, const BuoyT<T>& hb) const {
    const double W = M_.rho * M_.g;
    const T Fb = W * hb.V;
    const T bxw = F.R00 * hb.x + F.R01 * hb.y + F.R02 * hb.z + F.tx;
    const T byw = F.R10 * hb.x + F.R11 * hb.y + F.R12 * hb.z + F.ty;
    T tauX = (byw - M_.body.YO) * Fb, tauY = -(bxw - M_.body.XO) * Fb;
    T Fz = Fb - M_.g * M_.body.ms;
    T mw = T(0), Iw_p = T(0), Iw_r = T(0);
    for (Index n = 0; n < S_.nn; ++n) {
      const T& V = vol[n];
      if (V <= 0) continue;
      const T x = cx[n], y = cy[n], z = cz[n];
      const T wx = F.R00 * x + F.R01 * y + F.R02 * z + F.tx - M_.body.XO;
      const T wy = F.R10 * x + F.R11 * y + F.R12 * z + F.ty - M_.body.YO;
      const T wz = F.R20 * x + F.R21 * y + F.R22 * z + F.tz - zO;
      const T f = -W * V;
      tauX += wy * f; tauY -= wx * f; Fz += f;
      const T m = M_.rho * V;
      mw += m;
      Iw_p += m * (wx * wx + wz * wz);
      Iw_r += m * (wy * wy + wz * wz);
    }
    LoadsT<T> L;
    L.Fz = Fz; L.Mth = tauY; L.Mph = tauX * F.cth;
    L.mw = mw; L.Iw_p = Iw_p; L.Iw_r = Iw_r;
    return L;
  }

  // body.hpp:42-57
  void integrate(const LoadsT<T>& L) {
    const double dt = M_.dt;
    const Body& b = M_.body;
    const T mh = b.mh0 + L.mw;
    const T Ith = b.Ith0 + L.Iw_p, Iph = b.Iph0 + L.Iw_r;
    const T az = (L.Fz - b.cDz * vz - b.qz * dfabs(vz) * vz) / mh;
    const T ath = (L.Mth - b.cDth * wth - b.qth * dfabs(wth) * wth) / Ith;
    const T aph = (L.Mph - b.cDph * wph - b.qph * dfabs(wph) * wph) / Iph;
    vz += az * dt; wth += ath * dt; wph += aph * dt;
    zO += vz * dt; pitch += wth * dt; roll += wph * dt;
    if (pitch > kPitchClamp) { pitch = T(kPitchClamp); wth = T(0); }
    if (pitch < -kPitchClamp) { pitch = T(-kPitchClamp); wth = T(0); }
    if (roll > kRollClamp) { roll = T(kRollClamp); wph = T(0); }
    if (roll < -kRollClamp) { roll = T(-kRollClamp); wph = T(0); }
    t += dt;
  }

  // events.hpp:11-38: the monitor sums are kept (over_) for the event-time loss
  void track_events(const FrameT<T>& F) {
    for (Index m = 0; m < M_.nmon; ++m) over_[m] = T(0);
    for (Index c = 0; c < M_.nc; ++c)
      if (M_.monitor[c] >= 0 && q_[c] > 0) over_[M_.monitor[c]] += q_[c];
    for (Index m = 0; m < M_.nmon; ++m)
      if (over_[m] > M_.monThreshold && overT_[m] < 0) overT_[m] = t;
    for (Index i = 0; i < M_.nmark; ++i) {
      const T wz = world_z(F, M_.mkx[i], M_.mky[i], M_.mkz[i]);
      markZ_[i] = wz;
      if (markT_[i] >= 0) continue;
      const bool under = M_.mkWhen[i] == 0;
      if ((under && wz < 0) || (!under && wz > 0)) markT_[i] = t;
    }
    if (!foundered && (hullTop < M_.hullTopBelow || std::fabs(val(pitch)) > M_.pitchAbsAbove ||
                       (Vb < M_.buoyancyBelow && t > M_.minT))) {
      foundered = true;
      founderT = t;
    }
  }
};

}  // namespace dual
}  // namespace orc
