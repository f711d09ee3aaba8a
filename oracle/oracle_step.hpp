// ORACLE — test infrastructure only; the product never includes or links this file.
//
// A CPU restatement of the reference's single-source flooding step, operation for operation, so that the
// HIP path can be held to it bit for bit. Three summation orders exist, as in the reference
// (engine/include/sinksim/kernel/step.hpp:18):
//   Serial     the oracle's serial loops (hydro.hpp:14-35, 67-99)
//   Tree       256 thread partials / 32-lane warp trees in double ("portable", reduce.hpp:12-79)
//   TreeMixed  the same with single-precision partials widened to double ("portable32", reduce.hpp:81-164)
// and three math policies (math.hpp:27-75): Portable (fdlibm sin/cos/pow), Oracle (fdlibm sin/cos and
// V8 13.6 pow semantics over the platform pow, third_party/v8math/v8_math.hpp:26-36) and Std (platform).
//
// Parity pins (tests/): the reference's golden traces for the three numerics (dense single steps, sampled
// windows and end to end at tolerance 0), the box-barge closed forms of tests/test_kernel.cpp:163-227,
// mass conservation, and the reference's own CPU build (oracle/_ref) on sweeps.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

extern "C" {
double orc_fdl_sin(double x);
double orc_fdl_cos(double x);
double orc_fdl_pow(double x, double y);
}

namespace orc {

using Real = double;
using Index = std::int32_t;
using Byte = std::uint8_t;

// scheme v1 constants (kernel/scheme.hpp:11-28)
constexpr Real kEmptyVolume = 1e-6;
constexpr Real kFullFraction = 1.0 - 1e-7;
constexpr Real kLimiterShare = 0.5;
constexpr Real kOrificeHeadFloor = 0.01;
constexpr Real kWeirCoefficient = 2.0 / 3.0;
constexpr Real kWeirExponent = 1.5;
constexpr Real kVillemonteExponent = 0.385;
constexpr int kNewtonFast = 6;
constexpr int kNewtonFull = 40;
constexpr Real kSlopeMin = 1e-9;
constexpr Real kCentroidMinVolume = 1e-9;
constexpr Real kNoHead = -1e9;
constexpr Real kNoHeadTest = -1e8;
constexpr Real kPitchClamp = 1.35;
constexpr Real kRollClamp = 1.4;
constexpr Real kPi = 3.141592653589793;

constexpr int kLanes = 32;    // logical warp width of the canonical tree (reduce.hpp:12)
constexpr int kThreads = 256; // logical block width (reduce.hpp:13)

// JavaScript Math.max / Math.min for ordinary numbers (math.hpp:29-30)
inline Real jmax(Real a, Real b) { return (b > a) ? b : a; }
inline Real jmin(Real a, Real b) { return (b < a) ? b : a; }

struct MathPortable {
  static Real sin(Real x) { return orc_fdl_sin(x); }
  static Real cos(Real x) { return orc_fdl_cos(x); }
  static Real pow(Real x, Real y) { return orc_fdl_pow(x, y); }
};

// V8 13.6 Math.pow: ECMAScript special cases, y == 2 -> x*x, y == 0.5 -> sqrt, else the platform pow
// (v8_math.hpp:26-36). sin/cos are fdlibm's, which the oracle build takes from the vendored netlib library.
struct MathOracle {
  static Real sin(Real x) { return orc_fdl_sin(x); }
  static Real cos(Real x) { return orc_fdl_cos(x); }
  static Real pow(Real x, Real y) {
    if (std::isnan(y)) return NAN;
    if (std::isinf(y) && (x == 1 || x == -1)) return NAN;
    if (y == 2) return x * x;
    if (y == 0.5) {
      if (std::isinf(x)) return INFINITY;
      return std::sqrt(x + 0);
    }
    return std::pow(x, y);
  }
};

struct MathStd {
  static Real sin(Real x) { return std::sin(x); }
  static Real cos(Real x) { return std::cos(x); }
  static Real pow(Real x, Real y) { return std::pow(x, y); }
};

enum class Order { Serial, Tree, TreeMixed };

// Operation counters of the algorithmic work (per thread), read by the benchmark's roofline: column evaluations
// wet/dry, segment evaluations partial/full/dry, node passes, steps, active connections, pow calls, wet nodes.
struct OpCount {
  long long cwet, cdry, spart, sfull, sdry, passes, steps, active, pows, wetNodes;
};
inline thread_local OpCount g_ops{};
inline thread_local bool g_counting = false;
#define ORC_COUNT(field) \
  do {                   \
    if (g_counting) ++g_ops.field; \
  } while (0)

// ---------------------------------------------------------------- views (kernel/types.hpp:13-185)

struct Ship {
  Index ncol;
  const Real *x, *y, *dxdy, *zlo, *zhi;
  Index nseg;
  const Index* segCol;
  const Real *segA, *segE;
  Index nn;
  const Real *mu, *vmax, *amin, *aov;
  const Index *segStart, *segCount;
};

struct Body {
  Real ms, KG, g0x, g0y, g0z, XO, YO, mh0, Ith0, Iph0, cDz, cDth, cDph, qz, qth, qph;
};

struct Net {
  Real rho, g, sq2g, dt;
  Index ng;
  const Index *gStart, *gCount, *gNodes;
  const Byte* gMode;
  const Real *gpx, *gpy, *gpz;
  Real gMargin;
  Index nc;
  const Index *ca, *cb;
  const Byte* law;
  const Real *cx, *cy, *cz, *area, *coef;
  const Byte* en;
  const Real* tOn;
  const Index *skipGroup, *monitor;
  Real* q;  // per connection, written by the step
  // incidence (model.cpp:26-56)
  const Index *incStart, *incCount, *incConn;
  const Byte* incSide;
  Index seaCount;
  const Index* sea;
  Index nmon;
  Real monThreshold;
  Index nmark;
  const Real *mkx, *mky, *mkz;
  const Byte* mkWhen;
  Real hullTopBelow, pitchAbsAbove, buoyancyBelow, minT;
  Body body;
};

struct Frame {
  Real R00, R01, R02, R10, R11, R12, R20, R21, R22, tx, ty, tz, cth;
};
struct Buoy {
  Real V, x, y, z, top;
};
struct Pass {
  Real vol, dv, mx, my, mz, zmin, zmax;
};
struct Loads {
  Real Fz, Mth, Mph, mw, Iw_p, Iw_r, bxw, byw;
};

struct State {
  Real t, zO, pitch, roll, vz, wth, wph;
  Real* vol;
  Real* level;
};

struct Scratch {
  Real *area, *cx, *cy, *cz, *zmin, *zmax, *hEff, *aEff, *acc, *excess;
  Index* deg;
  Byte* merged;
  Pass* passes;
  Real* over;
  Real* dV;
};

struct Outputs {
  Real inflow, Vb, hullTop, mw, bx, by, bz;
};

struct Event {
  Real t;
  Index id;
  Byte kind;  // 0 overflow, 1 mark, 2 founder
};

struct Events {
  Event* items;
  Index capacity, count;
  Real* overT;
  Real* markT;
  Byte foundered;
  Real founderT;
};

inline void push_event(Events& ev, Real t, Index id, Byte kind) {  // types.hpp:240-247
  if (ev.count < ev.capacity) {
    ev.items[ev.count] = Event{t, id, kind};
    ev.count++;
  }
}

// ---------------------------------------------------------------- pose (frame.hpp:9-26)

template <class M>
inline Frame pose(const Body& b, Real zO, Real th, Real ph) {
  const Real cth = M::cos(th), sth = M::sin(th), cph = M::cos(ph), sph = M::sin(ph);
  Frame F;
  F.R00 = cth;
  F.R01 = sth * sph;
  F.R02 = sth * cph;
  F.R10 = 0;
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

inline Real world_z(const Frame& F, Real x, Real y, Real z) { return F.R20 * x + F.R21 * y + F.R22 * z + F.tz; }

// ---------------------------------------------------------------- partials (reduce.hpp:16-164)

template <class T>
struct HP {
  T V, Mx, My, Mz, top;
};
template <class T>
struct PP {
  T vol, dv, mx, my, mz, zmin, zmax;
};

template <class T>
inline HP<T> hp_zero() { return HP<T>{T(0), T(0), T(0), T(0), T(-1e9)}; }
template <class T>
inline PP<T> pp_zero() { return PP<T>{T(0), T(0), T(0), T(0), T(0), T(1e9), T(-1e9)}; }

template <class T>
inline void absorb(HP<T>& a, const HP<T>& b) {
  a.V += b.V; a.Mx += b.Mx; a.My += b.My; a.Mz += b.Mz;
  if (b.top > a.top) a.top = b.top;
}
template <class T>
inline void absorb(PP<T>& a, const PP<T>& b) {
  a.vol += b.vol; a.dv += b.dv; a.mx += b.mx; a.my += b.my; a.mz += b.mz;
  if (b.zmin < a.zmin) a.zmin = b.zmin;
  if (b.zmax > a.zmax) a.zmax = b.zmax;
}

// the canonical halving tree over L partials: l absorbs l + off for off = L/2 ... 1 (reduce.hpp:74-79)
template <class P, int L>
inline P halving(P (&v)[L]) {
  for (int off = L / 2; off > 0; off >>= 1)
    for (int l = 0; l < off; ++l) absorb(v[l], v[l + off]);
  return v[0];
}

struct FrameF {
  float R20, R21, R22, tz, iR;
};
inline FrameF frame_f(const Frame& F) {
  FrameF f;
  f.R20 = static_cast<float>(F.R20);
  f.R21 = static_cast<float>(F.R21);
  f.R22 = static_cast<float>(F.R22);
  f.tz = static_cast<float>(F.tz);
  f.iR = 1.0f / f.R22;
  return f;
}

// one column of the buoyancy integral, double (reduce.hpp:39-49)
inline void column(const Ship& S, const Frame& F, Real iR, Index i, HP<Real>& p) {
  const Real x = S.x[i], y = S.y[i], lo = S.zlo[i], hi = S.zhi[i];
  const Real base = F.R20 * x + F.R21 * y + F.tz;
  const Real wt = base + F.R22 * hi;
  if (wt > p.top) p.top = wt;
  const Real zp = -base * iR;
  if (zp <= lo) { ORC_COUNT(cdry); return; }
  ORC_COUNT(cwet);
  const Real s = zp >= hi ? hi - lo : zp - lo;
  const Real v = S.dxdy[i] * s;
  p.V += v; p.Mx += v * x; p.My += v * y; p.Mz += v * (lo + 0.5 * s);
}

// one column, single precision (reduce.hpp:131-142)
inline void column_f(const Ship& S, const FrameF& f, Index i, HP<float>& p) {
  const float x = static_cast<float>(S.x[i]), y = static_cast<float>(S.y[i]);
  const float lo = static_cast<float>(S.zlo[i]), hi = static_cast<float>(S.zhi[i]);
  const float base = f.R20 * x + f.R21 * y + f.tz;
  const float wt = base + f.R22 * hi;
  if (wt > p.top) p.top = wt;
  const float zp = -base * f.iR;
  if (zp <= lo) { ORC_COUNT(cdry); return; }
  ORC_COUNT(cwet);
  const float s = zp >= hi ? hi - lo : zp - lo;
  const float v = static_cast<float>(S.dxdy[i]) * s;
  p.V += v; p.Mx += v * x; p.My += v * y; p.Mz += v * (lo + 0.5f * s);
}

// one segment of a node pass, double (reduce.hpp:52-71)
inline void segment(const Ship& S, const Frame& F, Real iR, Real h, Index j, PP<Real>& p) {
  const Index i = S.segCol[j];
  const Real x = S.x[i], y = S.y[i], a = S.segA[j], b = S.segE[j];
  const Real base = F.R20 * x + F.R21 * y + F.tz;
  const Real wa = base + F.R22 * a, wb = base + F.R22 * b;
  if (wa < p.zmin) p.zmin = wa;
  if (wb > p.zmax) p.zmax = wb;
  const Real zp = (h - base) * iR;
  if (zp <= a) { ORC_COUNT(sdry); return; }
  const Real A = S.dxdy[i];
  Real L;
  if (zp >= b) {
    ORC_COUNT(sfull);
    L = b - a;
  } else {
    ORC_COUNT(spart);
    L = zp - a;
    p.dv += A;
  }
  const Real v = A * L;
  p.vol += v; p.mx += v * x; p.my += v * y; p.mz += v * (a + 0.5 * L);
}

// one segment, single precision (reduce.hpp:144-164)
inline void segment_f(const Ship& S, const FrameF& f, float h, Index j, PP<float>& p) {
  const Index i = S.segCol[j];
  const float x = static_cast<float>(S.x[i]), y = static_cast<float>(S.y[i]);
  const float a = static_cast<float>(S.segA[j]), b = static_cast<float>(S.segE[j]);
  const float base = f.R20 * x + f.R21 * y + f.tz;
  const float wa = base + f.R22 * a, wb = base + f.R22 * b;
  if (wa < p.zmin) p.zmin = wa;
  if (wb > p.zmax) p.zmax = wb;
  const float zp = (h - base) * f.iR;
  if (zp <= a) { ORC_COUNT(sdry); return; }
  const float A = static_cast<float>(S.dxdy[i]);
  float L;
  if (zp >= b) {
    ORC_COUNT(sfull);
    L = b - a;
  } else {
    ORC_COUNT(spart);
    L = zp - a;
    p.dv += A;
  }
  const float v = A * L;
  p.vol += v; p.mx += v * x; p.my += v * y; p.mz += v * (a + 0.5f * L);
}

inline HP<Real> widen(const HP<float>& p) {
  return HP<Real>{static_cast<Real>(p.V), static_cast<Real>(p.Mx), static_cast<Real>(p.My), static_cast<Real>(p.Mz), static_cast<Real>(p.top)};
}
inline PP<Real> widen(const PP<float>& p) {
  return PP<Real>{static_cast<Real>(p.vol), static_cast<Real>(p.dv), static_cast<Real>(p.mx), static_cast<Real>(p.my),
                  static_cast<Real>(p.mz), static_cast<Real>(p.zmin), static_cast<Real>(p.zmax)};
}

inline Buoy finish_buoy(const HP<Real>& p) {  // hydro.hpp:37-45
  Buoy o;
  o.V = p.V;
  o.x = p.V > 0 ? p.Mx / p.V : 0;
  o.y = p.V > 0 ? p.My / p.V : 0;
  o.z = p.V > 0 ? p.Mz / p.V : 0;
  o.top = p.top;
  return o;
}

inline Pass finish_pass(const Ship& S, Index n, Real iR, const PP<Real>& p) {  // hydro.hpp:101-109
  const Real mu = S.mu[n];
  Pass r;
  r.vol = p.vol * mu;
  r.dv = p.dv * mu * iR;
  r.mx = p.mx * mu; r.my = p.my * mu; r.mz = p.mz * mu;
  r.zmin = p.zmin; r.zmax = p.zmax;
  return r;
}

// ---------------------------------------------------------------- (a) column integrals (hydro.hpp)

template <Order O>
inline Buoy buoyancy(const Ship& S, const Frame& F) {
  if constexpr (O == Order::Serial) {
    // hydro.hpp:14-35: a single running sum; the same arithmetic as one partial holding every column
    const Real iR = 1 / F.R22;
    HP<Real> p = hp_zero<Real>();
    for (Index i = 0; i < S.ncol; ++i) column(S, F, iR, i, p);
    return finish_buoy(p);
  } else if constexpr (O == Order::Tree) {
    // hydro.hpp:48-64: partial t sums columns t, t+256, ...; 8 warp trees; warps combined 0..7 in order
    const Real iR = 1 / F.R22;
    HP<Real> warps[kThreads / kLanes];
    for (int w = 0; w < kThreads / kLanes; ++w) {
      HP<Real> lanes[kLanes];
      for (int l = 0; l < kLanes; ++l) {
        lanes[l] = hp_zero<Real>();
        for (Index i = w * kLanes + l; i < S.ncol; i += kThreads) column(S, F, iR, i, lanes[l]);
      }
      warps[w] = halving(lanes);
    }
    HP<Real> total = warps[0];
    for (int w = 1; w < kThreads / kLanes; ++w) absorb(total, warps[w]);
    return finish_buoy(total);
  } else {
    // hydro.hpp:125-141: single-precision partials and warp trees, widened before the in-order combine
    const FrameF f = frame_f(F);
    HP<float> warps[kThreads / kLanes];
    for (int w = 0; w < kThreads / kLanes; ++w) {
      HP<float> lanes[kLanes];
      for (int l = 0; l < kLanes; ++l) {
        lanes[l] = hp_zero<float>();
        for (Index i = w * kLanes + l; i < S.ncol; i += kThreads) column_f(S, f, i, lanes[l]);
      }
      warps[w] = halving(lanes);
    }
    HP<Real> total = widen(warps[0]);
    for (int w = 1; w < kThreads / kLanes; ++w) absorb(total, widen(warps[w]));
    return finish_buoy(total);
  }
}

template <Order O>
inline Pass node_pass(const Ship& S, const Frame& F, Index n, Real h) {
  ORC_COUNT(passes);
  const Index s0 = S.segStart[n], s1 = s0 + S.segCount[n];
  if constexpr (O == Order::Serial) {
    // hydro.hpp:67-99
    const Real iR = 1 / F.R22;
    PP<Real> p = pp_zero<Real>();
    for (Index j = s0; j < s1; ++j) segment(S, F, iR, h, j, p);
    return finish_pass(S, n, iR, p);
  } else if constexpr (O == Order::Tree) {
    // hydro.hpp:112-122
    const Real iR = 1 / F.R22;
    PP<Real> lanes[kLanes];
    for (int l = 0; l < kLanes; ++l) {
      lanes[l] = pp_zero<Real>();
      for (Index j = s0 + l; j < s1; j += kLanes) segment(S, F, iR, h, j, lanes[l]);
    }
    return finish_pass(S, n, iR, halving(lanes));
  } else {
    // hydro.hpp:144-155
    const FrameF f = frame_f(F);
    const float hf = static_cast<float>(h);
    PP<float> lanes[kLanes];
    for (int l = 0; l < kLanes; ++l) {
      lanes[l] = pp_zero<float>();
      for (Index j = s0 + l; j < s1; j += kLanes) segment_f(S, f, hf, j, lanes[l]);
    }
    return finish_pass(S, n, 1 / F.R22, widen(halving(lanes)));
  }
}

// ---------------------------------------------------------------- (b) levels (levels.hpp)

inline void set_centroid(Scratch& sc, Index n, const Pass& p) {
  if (p.vol > kCentroidMinVolume) { sc.cx[n] = p.mx / p.vol; sc.cy[n] = p.my / p.vol; sc.cz[n] = p.mz / p.vol; }
}

// levels.hpp:13-49
template <Order O>
inline void solve_single(const Ship& S, const Frame& F, Index n, bool full, State& st, Scratch& sc) {
  const Real V = st.vol[n], Vmax = S.vmax[n];
  Real h = st.level[n];
  Pass p = node_pass<O>(S, F, n, h);
  const Real lo = p.zmin, hi = p.zmax;
  sc.zmin[n] = lo; sc.zmax[n] = hi;
  if (V <= kEmptyVolume) {
    st.level[n] = lo; sc.area[n] = 0;
    return;
  }
  if (V >= Vmax * kFullFraction) {
    p = node_pass<O>(S, F, n, hi);
    st.level[n] = hi + jmax(Real(0), V - Vmax) / S.aov[n];
    sc.area[n] = S.aov[n];
    set_centroid(sc, n, p);
    return;
  }
  if (h < lo || h > hi) {
    h = jmin(hi, jmax(lo, h));
    p = node_pass<O>(S, F, n, h);
  }
  const Real tol = full ? 1e-6 * Vmax + 1e-3 : jmax(1e-3 * Vmax, Real(0.5));
  Real a = lo, b = hi;
  const int iters = full ? kNewtonFull : kNewtonFast;
  for (int it = 0; it < iters; ++it) {
    const Real err = p.vol - V;
    if (std::fabs(err) < tol) break;
    if (err > 0) b = h; else a = h;
    Real hn = p.dv > kSlopeMin ? h - err / p.dv : 0.5 * (a + b);
    if (!(hn > a && hn < b)) hn = 0.5 * (a + b);
    h = hn;
    p = node_pass<O>(S, F, n, h);
  }
  st.level[n] = h; sc.area[n] = p.dv;
  set_centroid(sc, n, p);
}

// levels.hpp:54-162
template <Order O>
inline void solve_group(const Ship& S, const Frame& F, const Index* members, Index N, bool full, State& st, Scratch& sc) {
  Pass* P = sc.passes;
  Real V = 0, Vmax = 0;
  for (Index k = 0; k < N; ++k) { V += st.vol[members[k]]; Vmax += S.vmax[members[k]]; }
  bool allWet = true;
  Index firstWet = -1;
  for (Index k = 0; k < N; ++k) {
    if (st.vol[members[k]] > 0) { if (firstWet < 0) firstWet = k; }
    else allWet = false;
  }
  Real h;
  if (allWet) {
    Real s = 0;
    for (Index k = 0; k < N; ++k) s += st.level[members[k]];
    h = s / Real(N);
  } else {
    h = st.level[members[firstWet >= 0 ? firstWet : N - 1]];
  }
  for (Index k = 0; k < N; ++k) P[k] = node_pass<O>(S, F, members[k], h);
  for (Index k = 0; k < N; ++k) { sc.zmin[members[k]] = P[k].zmin; sc.zmax[members[k]] = P[k].zmax; }
  Real lo = P[0].zmin, hi = P[0].zmax;
  for (Index k = 1; k < N; ++k) { if (P[k].zmin < lo) lo = P[k].zmin; if (P[k].zmax > hi) hi = P[k].zmax; }

  if (V <= kEmptyVolume) {
    for (Index k = 0; k < N; ++k) { const Index n = members[k]; st.vol[n] = 0; st.level[n] = P[k].zmin; sc.area[n] = 0; }
    return;
  }
  if (V >= Vmax * kFullFraction) {
    const Real over = jmax(Real(0), V - Vmax);
    Real aSum = 0;
    for (Index k = 0; k < N; ++k) aSum += S.aov[members[k]];
    const Real dh = over / aSum;
    Real assigned = 0;
    for (Index k = 0; k < N - 1; ++k) {
      const Real v = S.vmax[members[k]] + dh * S.aov[members[k]];
      st.vol[members[k]] = v;
      assigned += v;
    }
    st.vol[members[N - 1]] = V - assigned;
    for (Index k = 0; k < N; ++k) {
      const Index n = members[k];
      P[k] = node_pass<O>(S, F, n, hi);
      st.level[n] = hi + dh; sc.area[n] = S.aov[n];
      set_centroid(sc, n, P[k]);
    }
    return;
  }
  if (!(h >= lo && h <= hi)) {
    h = jmin(hi, jmax(lo, h));
    for (Index k = 0; k < N; ++k) P[k] = node_pass<O>(S, F, members[k], h);
  }
  const Real tol = full ? 1e-6 * Vmax + 1e-3 : jmax(1e-3 * Vmax, Real(0.5));
  Real a = lo, b = hi;
  const int iters = full ? kNewtonFull : kNewtonFast;
  for (int it = 0; it < iters; ++it) {
    Real sum = 0;
    for (Index k = 0; k < N; ++k) sum += P[k].vol;
    const Real err = sum - V;
    if (std::fabs(err) < tol) break;
    if (err > 0) b = h; else a = h;
    Real dv = 0;
    for (Index k = 0; k < N; ++k) dv += P[k].dv;
    Real hn = dv > kSlopeMin ? h - err / dv : 0.5 * (a + b);
    if (!(hn > a && hn < b)) hn = 0.5 * (a + b);
    h = hn;
    for (Index k = 0; k < N; ++k) P[k] = node_pass<O>(S, F, members[k], h);
  }
  Real tot = 0;
  for (Index k = 0; k < N; ++k) tot += P[k].vol;
  Real assigned = 0;
  for (Index k = 0; k < N - 1; ++k) {
    Real v = tot > 0 ? V * P[k].vol / tot : V / Real(N);
    if (v > S.vmax[members[k]]) v = S.vmax[members[k]];
    st.vol[members[k]] = v;
    assigned += v;
  }
  const Real vmaxLast = S.vmax[members[N - 1]];
  if (V - assigned > vmaxLast) {
    const Real need = V - vmaxLast;
    Real capped = 0;
    for (Index k = N - 2; k >= 0; --k) {
      Real prior = 0;
      for (Index j = 0; j < k; ++j) prior += st.vol[members[j]];
      const Real v = (need - capped) - prior;
      if (v > S.vmax[members[k]] && k > 0) {
        st.vol[members[k]] = S.vmax[members[k]];
        capped += S.vmax[members[k]];
        continue;
      }
      st.vol[members[k]] = v;
      break;
    }
    assigned = 0;
    for (Index k = 0; k < N - 1; ++k) assigned += st.vol[members[k]];
  }
  st.vol[members[N - 1]] = V - assigned;
  for (Index k = 0; k < N; ++k) {
    const Index n = members[k];
    st.level[n] = h; sc.area[n] = P[k].dv;
    set_centroid(sc, n, P[k]);
  }
}

// levels.hpp:166-194
template <Order O>
inline void solve_levels(const Ship& S, const Net& M, const Frame& F, bool full, State& st, Scratch& sc) {
  for (Index g = 0; g < M.ng; ++g) {
    const Index N = M.gCount[g];
    const Index* members = M.gNodes + M.gStart[g];
    const Byte mode = M.gMode[g];
    bool merged = false;
    if (N >= 2) {
      if (mode == 1) {
        merged = true;
      } else if (mode == 2) {
        const Real zMw = world_z(F, M.gpx[g], M.gpy[g], M.gpz[g]);
        merged = true;
        for (Index k = 0; k < N; ++k) {
          const Index n = members[k];
          if (!(st.vol[n] > kEmptyVolume && st.level[n] > zMw + M.gMargin)) { merged = false; break; }
        }
      }
    }
    sc.merged[g] = merged ? 1 : 0;
    if (merged) solve_group<O>(S, F, members, N, full, st, sc);
    else for (Index k = 0; k < N; ++k) solve_single<O>(S, F, members[k], full, st, sc);
  }
}

// ---------------------------------------------------------------- (c) network (flows.hpp)

inline bool counts(const Net& M, const Scratch& sc, Index c) {  // flows.hpp:21-26
  if (!M.en[c]) return false;
  const Index a = M.ca[c], b = M.cb[c];
  if (a >= 0 && sc.hEff[a] < kNoHeadTest && sc.hEff[b] < kNoHeadTest) return false;
  return true;
}

template <class Mth>
inline Real flow(const Net& M, const Frame& F, const State& st, const Scratch& sc, Index c) {  // flows.hpp:38-78
  if (!M.en[c] || M.tOn[c] > st.t) { M.q[c] = 0; return 0; }
  if (M.skipGroup[c] >= 0 && sc.merged[M.skipGroup[c]]) { M.q[c] = 0; return 0; }
  const Index a = M.ca[c], b = M.cb[c];
  const Real zo = F.R20 * M.cx[c] + F.R21 * M.cy[c] + F.R22 * M.cz[c] + F.tz;
  const Real ha = a < 0 ? Real(0) : sc.hEff[a], hbv = sc.hEff[b];
  const Real Ha = ha > zo ? ha - zo : Real(0), Hb = hbv > zo ? hbv - zo : Real(0);
  const Real sq2g = M.sq2g;
  Real q = 0;
  if (M.law[c] == 0) {
    const Real d = Ha - Hb;
    if (d != 0) q = M.coef[c] * M.area[c] * sq2g * d / std::sqrt(jmax(std::fabs(d), kOrificeHeadFloor));
  } else {
    const Real H1 = Ha > Hb ? Ha : Hb;
    if (H1 > 0) {
      const Real H2 = Ha > Hb ? Hb : Ha;
      Real f = 1;
      if (H2 > 0) {
        const Real r = H2 / H1;
        if (g_counting) g_ops.pows += 2;
        f = Mth::pow(jmax(Real(0), 1 - Mth::pow(r, kWeirExponent)), kVillemonteExponent);
      }
      q = M.coef[c] * M.area[c] * kWeirCoefficient * sq2g * H1 * std::sqrt(H1) * f * (Ha >= Hb ? Real(1) : Real(-1));
    }
  }
  if (q == 0) { M.q[c] = 0; return 0; }
  ORC_COUNT(active);
  Real dV = q * M.dt;
  const Index src = dV > 0 ? a : b, dst = dV > 0 ? b : a;
  const Real hs = src < 0 ? Real(0) : sc.hEff[src];
  const Real hd = dst < 0 ? Real(0) : sc.hEff[dst];
  const Real inv = (src < 0 ? Real(0) : 1 / sc.aEff[src]) + (dst < 0 ? Real(0) : 1 / sc.aEff[dst]);
  const Index degRef = src < 0 ? sc.deg[dst] : (dst < 0 ? sc.deg[src] : (sc.deg[src] > sc.deg[dst] ? sc.deg[src] : sc.deg[dst]));
  const Real share = kLimiterShare / Real(degRef > 1 ? degRef : 1);
  Real lim = inv > 0 ? share * (hs - jmax(hd, zo)) / inv : std::fabs(dV);
  if (lim < 0) lim = 0;
  if (std::fabs(dV) > lim) dV = dV > 0 ? lim : -lim;
  M.q[c] = dV / M.dt;
  return dV;
}

// flows.hpp:121-134, the serial driver
template <class Mth>
inline Real flows_and_volumes(const Ship& S, const Net& M, const Frame& F, State& st, Scratch& sc) {
  const Index NN = S.nn;
  for (Index n = 0; n < NN; ++n) {
    sc.hEff[n] = st.vol[n] > kEmptyVolume ? st.level[n] : kNoHead;
    sc.aEff[n] = st.vol[n] >= S.vmax[n] ? S.aov[n] : jmax(sc.area[n], S.amin[n]);
  }
  for (Index n = 0; n < NN; ++n) {
    const Index s = M.incStart[n], e = s + M.incCount[n];
    Index d = 0;
    for (Index k = s; k < e; ++k)
      if (counts(M, sc, M.incConn[k])) ++d;
    sc.deg[n] = d;
  }
  for (Index c = 0; c < M.nc; ++c) sc.dV[c] = flow<Mth>(M, F, st, sc, c);
  for (Index n = 0; n < NN; ++n) {
    const Index s = M.incStart[n], e = s + M.incCount[n];
    Real a = 0;
    for (Index k = s; k < e; ++k) {
      const Real d = sc.dV[M.incConn[k]];
      a = M.incSide[k] ? a + d : a - d;
    }
    sc.acc[n] = a;
  }
  Real seaIn = 0;
  for (Index k = 0; k < M.seaCount; ++k) seaIn += sc.dV[M.sea[k]];
  for (Index n = 0; n < NN; ++n) {  // flows.hpp:99-111
    Real v = st.vol[n] + sc.acc[n];
    if (v < 0) v = 0;
    const Real vm = S.vmax[n];
    Real excess = 0;
    if (v > vm) {
      const Real head = jmax(Real(0), -sc.zmax[n]);
      const Real cap = vm + S.aov[n] * head;
      if (v > cap) { excess = v - cap; v = cap; }
    }
    sc.excess[n] = excess;
    st.vol[n] = v;
  }
  for (Index n = 0; n < NN; ++n)  // flows.hpp:114-118
    if (sc.excess[n] > 0) seaIn -= sc.excess[n];
  return seaIn;
}

// ---------------------------------------------------------------- (d) rigid body (body.hpp)

inline Loads loads(const Ship& S, const Net& M, const Frame& F, const Buoy& hb, const State& st, const Scratch& sc) {
  const Real W = M.rho * M.g;
  const Real Fb = W * hb.V;
  const Real bxw = F.R00 * hb.x + F.R01 * hb.y + F.R02 * hb.z + F.tx;
  const Real byw = F.R10 * hb.x + F.R11 * hb.y + F.R12 * hb.z + F.ty;
  Real tauX = (byw - M.body.YO) * Fb, tauY = -(bxw - M.body.XO) * Fb;
  Real Fz = Fb - M.g * M.body.ms;
  Real mw = 0, Iw_p = 0, Iw_r = 0;
  for (Index n = 0; n < S.nn; ++n) {
    const Real V = st.vol[n];
    if (V <= 0) continue;
    ORC_COUNT(wetNodes);
    const Real x = sc.cx[n], y = sc.cy[n], z = sc.cz[n];
    const Real wx = F.R00 * x + F.R01 * y + F.R02 * z + F.tx - M.body.XO;
    const Real wy = F.R10 * x + F.R11 * y + F.R12 * z + F.ty - M.body.YO;
    const Real wz = F.R20 * x + F.R21 * y + F.R22 * z + F.tz - st.zO;
    const Real f = -W * V;
    tauX += wy * f; tauY -= wx * f; Fz += f;
    const Real m = M.rho * V;
    mw += m;
    Iw_p += m * (wx * wx + wz * wz);
    Iw_r += m * (wy * wy + wz * wz);
  }
  Loads L;
  L.Fz = Fz; L.Mth = tauY; L.Mph = tauX * F.cth;
  L.mw = mw; L.Iw_p = Iw_p; L.Iw_r = Iw_r;
  L.bxw = bxw; L.byw = byw;
  return L;
}

inline void integrate(const Net& M, const Loads& L, State& st) {
  const Real dt = M.dt;
  const Body& b = M.body;
  const Real mh = b.mh0 + L.mw;
  const Real Ith = b.Ith0 + L.Iw_p, Iph = b.Iph0 + L.Iw_r;
  const Real az = (L.Fz - b.cDz * st.vz - b.qz * std::fabs(st.vz) * st.vz) / mh;
  const Real ath = (L.Mth - b.cDth * st.wth - b.qth * std::fabs(st.wth) * st.wth) / Ith;
  const Real aph = (L.Mph - b.cDph * st.wph - b.qph * std::fabs(st.wph) * st.wph) / Iph;
  st.vz += az * dt; st.wth += ath * dt; st.wph += aph * dt;
  st.zO += st.vz * dt; st.pitch += st.wth * dt; st.roll += st.wph * dt;
  if (st.pitch > kPitchClamp) { st.pitch = kPitchClamp; st.wth = 0; }
  if (st.pitch < -kPitchClamp) { st.pitch = -kPitchClamp; st.wth = 0; }
  if (st.roll > kRollClamp) { st.roll = kRollClamp; st.wph = 0; }
  if (st.roll < -kRollClamp) { st.roll = -kRollClamp; st.wph = 0; }
  st.t += dt;
}

// ---------------------------------------------------------------- events (events.hpp:11-38)

inline void track_events(const Net& M, const Frame& F, const State& st, const Outputs& out, Scratch& sc, Events& ev) {
  for (Index m = 0; m < M.nmon; ++m) sc.over[m] = 0;
  for (Index c = 0; c < M.nc; ++c)
    if (M.monitor[c] >= 0 && M.q[c] > 0) sc.over[M.monitor[c]] += M.q[c];
  for (Index m = 0; m < M.nmon; ++m) {
    if (sc.over[m] > M.monThreshold && ev.overT[m] < 0) {
      ev.overT[m] = st.t;
      push_event(ev, st.t, m, 0);
    }
  }
  for (Index i = 0; i < M.nmark; ++i) {
    if (ev.markT[i] >= 0) continue;
    const Real wz = world_z(F, M.mkx[i], M.mky[i], M.mkz[i]);
    const bool under = M.mkWhen[i] == 0;
    if ((under && wz < 0) || (!under && wz > 0)) {
      ev.markT[i] = st.t;
      push_event(ev, st.t, i, 1);
    }
  }
  if (!ev.foundered && (out.hullTop < M.hullTopBelow || std::fabs(st.pitch) > M.pitchAbsAbove ||
                        (out.Vb < M.buoyancyBelow && st.t > M.minT))) {
    ev.foundered = 1;
    ev.founderT = st.t;
    push_event(ev, st.t, -1, 2);
  }
}

// ---------------------------------------------------------------- the step (step.hpp:41-54)

template <class Mth, Order O>
inline void step(const Ship& S, const Net& M, State& st, Scratch& sc, Outputs& out, Events& ev) {
  ORC_COUNT(steps);
  const Frame F = pose<Mth>(M.body, st.zO, st.pitch, st.roll);
  const Buoy hb = buoyancy<O>(S, F);
  solve_levels<O>(S, M, F, false, st, sc);
  const Real seaIn = flows_and_volumes<Mth>(S, M, F, st, sc);
  out.inflow = seaIn / M.dt;
  const Loads L = loads(S, M, F, hb, st, sc);
  integrate(M, L, st);
  out.Vb = hb.V; out.hullTop = hb.top; out.mw = L.mw;
  out.bx = hb.x; out.by = hb.y; out.bz = hb.z;
  track_events(M, F, st, out, sc, ev);
}

// readouts.hpp:20-36
struct Readouts {
  Real t, trimDeg, listDeg, draftF, draftA, waterM3, waterT, inflowTpm;
};

template <class Mth>
inline Readouts readouts(const Ship& S, const Net& M, Real xFP, Real xAP, const State& st, const Outputs& out) {
  const Frame F = pose<Mth>(M.body, st.zO, st.pitch, st.roll);
  Readouts r;
  r.t = st.t;
  r.trimDeg = st.pitch * 180 / kPi;
  r.listDeg = st.roll * 180 / kPi;
  r.draftF = -world_z(F, xFP, 0, 0);
  r.draftA = -world_z(F, xAP, 0, 0);
  Real water = 0;
  for (Index n = 0; n < S.nn; ++n) water += st.vol[n];
  r.waterM3 = water;
  r.waterT = water * M.rho / 1000;
  r.inflowTpm = out.inflow * M.rho / 1000 * 60;
  return r;
}

}  // namespace orc
