// The host CPU engine of libsinksim_hip.so: the reference's four numerics on the CPU, for the CPU tools
// (sinksim_sweep, whose reference runs every numerics on the host) and for callers without a device.
//
//   oracle      serial sums (the oracle's order), fdlibm sin/cos, V8 13.6 pow semantics over the platform pow
//               (third_party/v8math/v8_math.hpp:26-36)
//   portable    the GPU's canonical tree order in double (256 partials, 32-lane halving trees), fdlibm sin/cos/pow
//   portable32  the same trees over single-precision partials, widened after each tree
//   std         serial sums, the platform's sin/cos/pow
//
// Reference: Simulation::step / run / record_trace (engine/src/simulation.cpp:147-221) over kernel::step<Math, Order>
// (engine/include/sinksim/kernel/step.hpp:41-54), and sweep::run_cpu (engine/src/sweep.cpp:121-161). The step is
// restated phase by phase in the reference's association order; this translation unit is compiled with
// -ffp-contract=off like the device code, and the fdlibm it calls is the device's own (fdlibm_dev.hpp, host and
// device), so that portable bits are the same on both sides.
//
// Layout: a model holds the ship and scenario (copied, so a handle owns its data), its incidence lists and segment
// records pre-gathered per segment (x, y, bottom, top, footprint; float copies for portable32). A worker thread owns
// one instance's working storage and reuses it across the instances it takes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sinksim_hip.h"
#include "batch_host.hpp"
#include "fdlibm_dev.hpp"

namespace {

// ------------------------------------------------------------------------------------------- math policies

struct PortableMath {
  static double sin(double x) { return ssa::fdl::sin(x); }
  static double cos(double x) { return ssa::fdl::cos(x); }
  static double pow(double x, double y) { return ssa::fdl::pow(x, y); }
};

// V8 13.6's Math.pow: the ECMAScript special cases, y == 2 as x * x, y == 0.5 as sqrt(x + 0), the platform pow
// otherwise; sin and cos are fdlibm's (V8's ieee754 port of it).
struct OracleMath {
  static double sin(double x) { return ssa::fdl::sin(x); }
  static double cos(double x) { return ssa::fdl::cos(x); }
  static double pow(double x, double y) {
    if (std::isnan(y)) return std::numeric_limits<double>::quiet_NaN();
    if (std::isinf(y) && (x == 1 || x == -1)) return std::numeric_limits<double>::quiet_NaN();
    if (y == 2) return x * x;
    if (y == 0.5) return std::isinf(x) ? std::numeric_limits<double>::infinity() : std::sqrt(x + 0);
    return std::pow(x, y);
  }
};

struct StdMath {
  static double sin(double x) { return std::sin(x); }
  static double cos(double x) { return std::cos(x); }
  static double pow(double x, double y) { return std::pow(x, y); }
};

// scheme v1 (engine/include/sinksim/kernel/scheme.hpp:11-28) and JavaScript's Math.max / Math.min for ordinary
// numbers (kernel/math.hpp:29-30)
constexpr double kEmpty = 1e-6, kFull = 1.0 - 1e-7, kShare = 0.5, kHeadFloor = 0.01, kWeirC = 2.0 / 3.0, kWeirExp = 1.5, kVilExp = 0.385;
constexpr int kNewton = 6;
constexpr double kSlope = 1e-9, kCentroidVol = 1e-9, kNoHead = -1e9, kNoHeadTest = -1e8, kPitchMax = 1.35, kRollMax = 1.4,
                 kPi = 3.141592653589793;
inline double jmax(double a, double b) { return (b > a) ? b : a; }
inline double jmin(double a, double b) { return (b < a) ? b : a; }

// ------------------------------------------------------------------------------------------- model

template <class T>
struct Rec {
  T x, y, lo, hi, A;  // a column, or a segment with its column's x, y, footprint and its own bottom and top
};

struct HostModel {
  int ncol = 0, nseg = 0, nn = 0, ng = 0, nc = 0, nmon = 0, nmark = 0;
  std::vector<Rec<double>> col, seg;
  std::vector<Rec<float>> colF, segF;
  std::vector<double> mu, vmax, amin, aov;
  std::vector<int32_t> segStart, segCount;
  double rho = 0, g = 0, sq2g = 0, dt = 0, gMargin = 0, monThreshold = 0;
  std::vector<int32_t> gStart, gCount, gNodes;
  std::vector<uint8_t> gMode;
  std::vector<double> gpx, gpy, gpz;
  std::vector<int32_t> ca, cb, skipGroup, monitor;
  std::vector<uint8_t> law, en0;
  std::vector<double> cx, cy, cz, area0, coef, tOn0;
  std::vector<int32_t> incStart, incCount, incConn, sea;  // node -> connections in increasing order (model.cpp:26-56)
  std::vector<uint8_t> incSide;                           // 1: the node is the connection's b end
  std::vector<double> mkx, mky, mkz;
  std::vector<uint8_t> mkWhen;
  double hullTopBelow = 0, pitchAbsAbove = 0, buoyancyBelow = 0, minT = 0;
  double ms = 0, g0x = 0, g0y = 0, g0z = 0, XO = 0, YO = 0, mh0 = 0, Ith0 = 0, Iph0 = 0, cDz = 0, cDth = 0, cDph = 0, qz = 0, qth = 0, qph = 0;
  double xFP = 0, xAP = 0;
  double state0[7] = {};
  std::vector<double> vol0, level0, cx0, cy0, cz0;
  int maxGroup = 1, evCap = 1;
};

template <class T>
std::vector<T> copy_of(const T* p, size_t n) {
  return p && n ? std::vector<T>(p, p + n) : std::vector<T>(n);
}

HostModel make_model(const ss_ship& S, const ss_sim& M) {
  ssa::validate(S, M, "ss_cpu");  // every array present and every index in range
  HostModel m;
  m.ncol = S.ncol; m.nseg = S.nseg; m.nn = S.nnode; m.ng = M.ngroup; m.nc = M.nconn; m.nmon = M.nmonitor; m.nmark = M.nmark;
  const size_t nc = static_cast<size_t>(S.ncol), ns = static_cast<size_t>(S.nseg), nn = static_cast<size_t>(S.nnode);
  m.col.resize(nc); m.colF.resize(nc);
  for (size_t i = 0; i < nc; ++i) {
    m.col[i] = Rec<double>{S.x[i], S.y[i], S.zlo[i], S.zhi[i], S.dxdy[i]};
    m.colF[i] = Rec<float>{static_cast<float>(S.x[i]), static_cast<float>(S.y[i]), static_cast<float>(S.zlo[i]), static_cast<float>(S.zhi[i]),
                           static_cast<float>(S.dxdy[i])};
  }
  m.seg.resize(ns); m.segF.resize(ns);
  for (size_t j = 0; j < ns; ++j) {
    const int32_t c = S.segCol[j];
    m.seg[j] = Rec<double>{S.x[c], S.y[c], S.segA[j], S.segE[j], S.dxdy[c]};
    m.segF[j] = Rec<float>{static_cast<float>(S.x[c]), static_cast<float>(S.y[c]), static_cast<float>(S.segA[j]), static_cast<float>(S.segE[j]),
                           static_cast<float>(S.dxdy[c])};
  }
  m.mu = copy_of(S.mu, nn); m.vmax = copy_of(S.vmax, nn); m.amin = copy_of(S.amin, nn); m.aov = copy_of(S.aov, nn);
  m.segStart = copy_of(S.segStart, nn); m.segCount = copy_of(S.segCount, nn);
  m.rho = M.rho; m.g = M.g; m.sq2g = std::sqrt(2 * M.g); m.dt = M.dt; m.gMargin = M.gMargin; m.monThreshold = M.monitorThreshold;
  const size_t ng = static_cast<size_t>(M.ngroup), ncn = static_cast<size_t>(M.nconn);
  m.gStart = copy_of(M.gNodeStart, ng); m.gCount = copy_of(M.gNodeCount, ng); m.gMode = copy_of(M.gMode, ng);
  m.gpx = copy_of(M.gpx, ng); m.gpy = copy_of(M.gpy, ng); m.gpz = copy_of(M.gpz, ng);
  size_t members = 0;
  for (size_t g = 0; g < ng; ++g) {
    members = std::max(members, static_cast<size_t>(m.gStart[g]) + static_cast<size_t>(m.gCount[g]));
    m.maxGroup = std::max(m.maxGroup, m.gCount[g]);
  }
  m.gNodes = copy_of(M.gNodes, members);
  m.ca = copy_of(M.ca, ncn); m.cb = copy_of(M.cb, ncn); m.law = copy_of(M.law, ncn); m.cx = copy_of(M.cx, ncn); m.cy = copy_of(M.cy, ncn);
  m.cz = copy_of(M.cz, ncn); m.area0 = copy_of(M.area, ncn); m.coef = copy_of(M.coef, ncn); m.en0 = copy_of(M.en, ncn);
  m.tOn0 = copy_of(M.tOn, ncn); m.skipGroup = copy_of(M.skipGroup, ncn); m.monitor = copy_of(M.monitor, ncn);
  // incidence: every node's connections in increasing connection order; a connection whose a end is the sea goes to
  // the sea list instead
  m.incCount.assign(nn, 0);
  for (size_t c = 0; c < ncn; ++c) {
    if (m.ca[c] >= 0) m.incCount[static_cast<size_t>(m.ca[c])]++;
    m.incCount[static_cast<size_t>(m.cb[c])]++;
  }
  m.incStart.assign(nn, 0);
  for (size_t n = 1; n < nn; ++n) m.incStart[n] = m.incStart[n - 1] + m.incCount[n - 1];
  const size_t total = nn ? static_cast<size_t>(m.incStart[nn - 1] + m.incCount[nn - 1]) : 0;
  m.incConn.assign(total, 0);
  m.incSide.assign(total, 0);
  std::vector<int32_t> at(m.incStart);
  for (size_t c = 0; c < ncn; ++c) {
    const int32_t ci = static_cast<int32_t>(c);
    if (m.ca[c] >= 0) {
      const size_t k = static_cast<size_t>(at[static_cast<size_t>(m.ca[c])]++);
      m.incConn[k] = ci; m.incSide[k] = 0;
    } else {
      m.sea.push_back(ci);
    }
    const size_t k = static_cast<size_t>(at[static_cast<size_t>(m.cb[c])]++);
    m.incConn[k] = ci; m.incSide[k] = 1;
  }
  m.mkx = copy_of(M.mkx, static_cast<size_t>(M.nmark)); m.mky = copy_of(M.mky, static_cast<size_t>(M.nmark));
  m.mkz = copy_of(M.mkz, static_cast<size_t>(M.nmark)); m.mkWhen = copy_of(M.mkWhen, static_cast<size_t>(M.nmark));
  m.hullTopBelow = M.founder[0]; m.pitchAbsAbove = M.founder[1]; m.buoyancyBelow = M.founder[2]; m.minT = M.founder[3];
  const double* b = M.body;
  m.ms = b[0]; m.g0x = b[2]; m.g0y = b[3]; m.g0z = b[4]; m.XO = b[5]; m.YO = b[6]; m.mh0 = b[7]; m.Ith0 = b[8]; m.Iph0 = b[9];
  m.cDz = b[10]; m.cDth = b[11]; m.cDph = b[12]; m.qz = b[13]; m.qth = b[14]; m.qph = b[15];
  m.xFP = M.readout[0]; m.xAP = M.readout[1];
  std::copy(M.state0, M.state0 + 7, m.state0);
  m.vol0 = copy_of(M.vol0, nn); m.level0 = copy_of(M.level0, nn); m.cx0 = copy_of(M.cx0, nn); m.cy0 = copy_of(M.cy0, nn);
  m.cz0 = copy_of(M.cz0, nn);
  m.evCap = m.nmon + m.nmark + 1;
  return m;
}

// ------------------------------------------------------------------------------------------- one instance

struct Frame {
  double R00, R01, R02, R10, R11, R12, R20, R21, R22, tx, ty, tz, cth;
};

struct Event {
  double t;
  int32_t id, kind;  // kind 0 overflow (monitor), 1 mark, 2 founder (id -1)
};

struct Pass {
  double vol, dv, mx, my, mz, zmin, zmax;
};

struct Inst {
  // state (the centroids cx, cy, cz are state too)
  double t = 0, zO = 0, pitch = 0, roll = 0, vz = 0, wth = 0, wph = 0;
  std::vector<double> vol, level, cx, cy, cz;
  // per-instance connection tables
  std::vector<uint8_t> en;
  std::vector<double> area, tOn;
  // working storage of a step
  std::vector<double> sarea, zmin, zmax, hEff, aEff, acc, excess, dV, q, over;
  std::vector<int32_t> deg;
  std::vector<uint8_t> merged;
  std::vector<Pass> passes;
  // outputs of the last step
  double inflow = 0, Vb = 0, hullTop = 0, mw = 0, bx = 0, by = 0, bz = 0;
  // events
  std::vector<double> overT, markT;
  std::vector<Event> events;
  bool foundered = false;
  double founderT = -1;

  void size_for(const HostModel& m) {
    const size_t nn = static_cast<size_t>(m.nn), nc = static_cast<size_t>(m.nc);
    for (std::vector<double>* v : {&vol, &level, &cx, &cy, &cz, &sarea, &zmin, &zmax, &hEff, &aEff, &acc, &excess}) v->assign(nn, 0.0);
    deg.assign(nn, 0);
    merged.assign(static_cast<size_t>(m.ng), 0);
    dV.assign(nc, 0.0); q.assign(nc, 0.0);
    over.assign(static_cast<size_t>(m.nmon), 0.0);
    passes.assign(static_cast<size_t>(m.maxGroup), Pass{});
    events.reserve(static_cast<size_t>(m.evCap));
  }

  // Simulation::reset_to_initial (engine/src/simulation.cpp:83-90) with the instance's connection tables
  void reset(const HostModel& m, const uint8_t* en_, const double* area_, const double* tOn_) {
    size_for(m);
    const size_t nc = static_cast<size_t>(m.nc);
    en.assign(en_ ? en_ : m.en0.data(), (en_ ? en_ : m.en0.data()) + nc);
    area.assign(area_ ? area_ : m.area0.data(), (area_ ? area_ : m.area0.data()) + nc);
    tOn.assign(tOn_ ? tOn_ : m.tOn0.data(), (tOn_ ? tOn_ : m.tOn0.data()) + nc);
    t = m.state0[0]; zO = m.state0[1]; pitch = m.state0[2]; roll = m.state0[3]; vz = m.state0[4]; wth = m.state0[5]; wph = m.state0[6];
    vol = m.vol0; level = m.level0; cx = m.cx0; cy = m.cy0; cz = m.cz0;
    overT.assign(static_cast<size_t>(m.nmon), -1.0);
    markT.assign(static_cast<size_t>(m.nmark), -1.0);
    events.clear();
    foundered = false;
    founderT = -1;
  }

  void push_event(const HostModel& m, double at, int32_t id, int32_t kind) {
    if (static_cast<int>(events.size()) < m.evCap) events.push_back(Event{at, id, kind});
  }
};

// frame.hpp:9-21
template <class Math>
Frame pose(const HostModel& m, double zO, double th, double ph) {
  const double cth = Math::cos(th), sth = Math::sin(th), cph = Math::cos(ph), sph = Math::sin(ph);
  Frame F;
  F.R00 = cth; F.R01 = sth * sph; F.R02 = sth * cph;
  F.R10 = 0; F.R11 = cph; F.R12 = -sph;
  F.R20 = -sth; F.R21 = cth * sph; F.R22 = cth * cph;
  F.tx = m.XO - (F.R00 * m.g0x + F.R01 * m.g0y + F.R02 * m.g0z);
  F.ty = m.YO - (F.R10 * m.g0x + F.R11 * m.g0y + F.R12 * m.g0z);
  F.tz = zO - (F.R20 * m.g0x + F.R21 * m.g0y + F.R22 * m.g0z);
  F.cth = cth;
  return F;
}

inline double world_z(const Frame& F, double x, double y, double z) { return F.R20 * x + F.R21 * y + F.R22 * z + F.tz; }

// ------------------------------------------------------------------------------------------- column integrals

// The pose as the integrals see it in precision T: double takes the frame's own values and iR = 1 / R22; float
// rounds them once and forms 1.0f / R22 in float (reduce.hpp:90-98).
template <class T>
struct Plane {
  T R20, R21, R22, tz, iR;
};
template <class T>
Plane<T> plane(const Frame& F) {
  if constexpr (sizeof(T) == sizeof(double)) {
    return Plane<T>{F.R20, F.R21, F.R22, F.tz, 1 / F.R22};
  } else {
    const float r22 = static_cast<float>(F.R22);
    return Plane<T>{static_cast<float>(F.R20), static_cast<float>(F.R21), r22, static_cast<float>(F.tz), 1.0f / r22};
  }
}

template <class T>
struct Buoy {
  T V, Mx, My, Mz, top;
  void add(const Buoy& o) {  // reduce.hpp:27-30
    V += o.V; Mx += o.Mx; My += o.My; Mz += o.Mz;
    if (o.top > top) top = o.top;
  }
};
template <class T>
struct Wet {
  T vol, dv, mx, my, mz, zmin, zmax;
  void add(const Wet& o) {  // reduce.hpp:32-36
    vol += o.vol; dv += o.dv; mx += o.mx; my += o.my; mz += o.mz;
    if (o.zmin < zmin) zmin = o.zmin;
    if (o.zmax > zmax) zmax = o.zmax;
  }
};

// one column into a buoyancy partial (hydro.hpp:17-27, reduce.hpp:39-49 and 131-142)
template <class T>
inline void column(const Plane<T>& f, const Rec<T>& r, Buoy<T>& p) {
  const T base = f.R20 * r.x + f.R21 * r.y + f.tz;
  const T wt = base + f.R22 * r.hi;
  if (wt > p.top) p.top = wt;
  const T zp = -base * f.iR;
  if (zp <= r.lo) return;
  const T s = zp >= r.hi ? r.hi - r.lo : zp - r.lo;
  const T v = r.A * s;
  p.V += v; p.Mx += v * r.x; p.My += v * r.y; p.Mz += v * (r.lo + T(0.5) * s);
}

// one segment into a node-pass partial (hydro.hpp:73-91, reduce.hpp:52-71 and 144-164)
template <class T>
inline void segment(const Plane<T>& f, T h, const Rec<T>& r, Wet<T>& p) {
  const T base = f.R20 * r.x + f.R21 * r.y + f.tz;
  const T wa = base + f.R22 * r.lo, wb = base + f.R22 * r.hi;
  if (wa < p.zmin) p.zmin = wa;
  if (wb > p.zmax) p.zmax = wb;
  const T zp = (h - base) * f.iR;
  if (zp <= r.lo) return;
  T L;
  if (zp >= r.hi) {
    L = r.hi - r.lo;
  } else {
    L = zp - r.lo;
    p.dv += r.A;
  }
  const T v = r.A * L;
  p.vol += v; p.mx += v * r.x; p.my += v * r.y; p.mz += v * (r.lo + T(0.5) * L);
}

// the canonical halving tree over L partials, its result in v[0] (reduce.hpp:74-79)
template <class P, int L>
P halve(P* v) {
  for (int off = L / 2; off > 0; off >>= 1)
    for (int l = 0; l < off; ++l) v[l].add(v[l + off]);
  return v[0];
}

template <class T>
Buoy<double> widen(const Buoy<T>& p) {
  return Buoy<double>{static_cast<double>(p.V), static_cast<double>(p.Mx), static_cast<double>(p.My), static_cast<double>(p.Mz),
                      static_cast<double>(p.top)};
}
template <class T>
Wet<double> widen(const Wet<T>& p) {
  return Wet<double>{static_cast<double>(p.vol), static_cast<double>(p.dv), static_cast<double>(p.mx), static_cast<double>(p.my),
                     static_cast<double>(p.mz), static_cast<double>(p.zmin), static_cast<double>(p.zmax)};
}

const std::vector<Rec<double>>& columns_of(const HostModel& m, double) { return m.col; }
const std::vector<Rec<float>>& columns_of(const HostModel& m, float) { return m.colF; }
const std::vector<Rec<double>>& segments_of(const HostModel& m, double) { return m.seg; }
const std::vector<Rec<float>>& segments_of(const HostModel& m, float) { return m.segF; }

// The two orders of summation. Serial: one running sum over the columns / segments in index order (the oracle's).
// Tree<T>: partial t of 256 takes the columns i with i mod 256 == t in order (a node pass: lane l of 32 the segments
// s0 + l, s0 + l + 32, ...), each 32 partials are halved into one, and the 8 results of the buoyancy are added in
// order (hydro.hpp:47-64, 113-155). Mixed precision keeps partials and trees in float and widens each tree's result.
struct Serial {
  static Buoy<double> buoyancy(const HostModel& m, const Frame& F) {
    const Plane<double> f = plane<double>(F);
    Buoy<double> p{0, 0, 0, 0, -1e9};
    for (const Rec<double>& r : m.col) column(f, r, p);
    return p;
  }
  static Wet<double> pass(const HostModel& m, const Frame& F, int n, double h) {
    const Plane<double> f = plane<double>(F);
    Wet<double> p{0, 0, 0, 0, 0, 1e9, -1e9};
    const size_t s0 = static_cast<size_t>(m.segStart[n]), s1 = s0 + static_cast<size_t>(m.segCount[n]);
    for (size_t j = s0; j < s1; ++j) segment(f, h, m.seg[j], p);
    return p;
  }
};

template <class T>
struct Tree {
  static Buoy<double> buoyancy(const HostModel& m, const Frame& F) {
    const Plane<T> f = plane<T>(F);
    Buoy<T> part[256];
    for (Buoy<T>& p : part) p = Buoy<T>{T(0), T(0), T(0), T(0), T(-1e9)};
    const std::vector<Rec<T>>& cols = columns_of(m, T(0));
    for (size_t i = 0; i < cols.size(); ++i) column(f, cols[i], part[i & 255]);
    Buoy<double> total = widen(halve<Buoy<T>, 32>(part));
    for (int w = 1; w < 8; ++w) total.add(widen(halve<Buoy<T>, 32>(part + 32 * w)));
    return total;
  }
  static Wet<double> pass(const HostModel& m, const Frame& F, int n, double h) {
    const Plane<T> f = plane<T>(F);
    const T hh = static_cast<T>(h);
    Wet<T> lane[32];
    for (Wet<T>& p : lane) p = Wet<T>{T(0), T(0), T(0), T(0), T(0), T(1e9), T(-1e9)};
    const std::vector<Rec<T>>& segs = segments_of(m, T(0));
    const size_t s0 = static_cast<size_t>(m.segStart[n]), s1 = s0 + static_cast<size_t>(m.segCount[n]);
    for (size_t j = s0; j < s1; ++j) segment(f, hh, segs[j], lane[(j - s0) & 31]);
    return widen(halve<Wet<T>, 32>(lane));
  }
};

// finish_hydro / finish_pass (hydro.hpp:37-45, 101-109): moments over volume, the permeability, dv over R22
template <class Sums>
void buoyancy(const HostModel& m, const Frame& F, double& V, double& x, double& y, double& z, double& top) {
  const Buoy<double> p = Sums::buoyancy(m, F);
  V = p.V;
  x = p.V > 0 ? p.Mx / p.V : 0;
  y = p.V > 0 ? p.My / p.V : 0;
  z = p.V > 0 ? p.Mz / p.V : 0;
  top = p.top;
}

template <class Sums>
Pass node_pass(const HostModel& m, const Frame& F, int n, double h) {
  const Wet<double> p = Sums::pass(m, F, n, h);
  const double mu = m.mu[static_cast<size_t>(n)], iR = 1 / F.R22;
  return Pass{p.vol * mu, p.dv * mu * iR, p.mx * mu, p.my * mu, p.mz * mu, p.zmin, p.zmax};
}

// ------------------------------------------------------------------------------------------- (b) levels

inline void set_centroid(Inst& s, size_t n, const Pass& p) {
  if (p.vol > kCentroidVol) { s.cx[n] = p.mx / p.vol; s.cy[n] = p.my / p.vol; s.cz[n] = p.mz / p.vol; }
}

// levels.hpp:13-49 (a step's solve: 6 Newton iterations, the looser tolerance)
template <class Sums>
void level_single(const HostModel& m, const Frame& F, Inst& s, int n) {
  const size_t k = static_cast<size_t>(n);
  const double V = s.vol[k], Vmax = m.vmax[k];
  double h = s.level[k];
  Pass p = node_pass<Sums>(m, F, n, h);
  const double lo = p.zmin, hi = p.zmax;
  s.zmin[k] = lo; s.zmax[k] = hi;
  if (V <= kEmpty) {
    s.level[k] = lo; s.sarea[k] = 0;
    return;
  }
  if (V >= Vmax * kFull) {
    p = node_pass<Sums>(m, F, n, hi);
    s.level[k] = hi + jmax(0.0, V - Vmax) / m.aov[k];
    s.sarea[k] = m.aov[k];
    set_centroid(s, k, p);
    return;
  }
  if (h < lo || h > hi) {
    h = jmin(hi, jmax(lo, h));
    p = node_pass<Sums>(m, F, n, h);
  }
  const double tol = jmax(1e-3 * Vmax, 0.5);
  double a = lo, b = hi;
  for (int it = 0; it < kNewton; ++it) {
    const double err = p.vol - V;
    if (std::fabs(err) < tol) break;
    if (err > 0) b = h; else a = h;
    double hn = p.dv > kSlope ? h - err / p.dv : 0.5 * (a + b);
    if (!(hn > a && hn < b)) hn = 0.5 * (a + b);
    h = hn;
    p = node_pass<Sums>(m, F, n, h);
  }
  s.level[k] = h; s.sarea[k] = p.dv;
  set_centroid(s, k, p);
}

// levels.hpp:54-162: N members under one free surface, the volume conserved exactly
template <class Sums>
void level_group(const HostModel& m, const Frame& F, Inst& s, const int32_t* mem, int N) {
  std::vector<Pass>& P = s.passes;
  auto node = [&](int k) { return static_cast<size_t>(mem[k]); };
  double V = 0, Vmax = 0;
  for (int k = 0; k < N; ++k) { V += s.vol[node(k)]; Vmax += m.vmax[node(k)]; }
  bool allWet = true;
  int firstWet = -1;
  for (int k = 0; k < N; ++k) {
    if (s.vol[node(k)] > 0) { if (firstWet < 0) firstWet = k; }
    else allWet = false;
  }
  double h;
  if (allWet) {
    double sum = 0;
    for (int k = 0; k < N; ++k) sum += s.level[node(k)];
    h = sum / static_cast<double>(N);
  } else {
    h = s.level[node(firstWet >= 0 ? firstWet : N - 1)];
  }
  for (int k = 0; k < N; ++k) P[static_cast<size_t>(k)] = node_pass<Sums>(m, F, mem[k], h);
  for (int k = 0; k < N; ++k) { s.zmin[node(k)] = P[static_cast<size_t>(k)].zmin; s.zmax[node(k)] = P[static_cast<size_t>(k)].zmax; }
  double lo = P[0].zmin, hi = P[0].zmax;
  for (int k = 1; k < N; ++k) { if (P[static_cast<size_t>(k)].zmin < lo) lo = P[static_cast<size_t>(k)].zmin; if (P[static_cast<size_t>(k)].zmax > hi) hi = P[static_cast<size_t>(k)].zmax; }
  if (V <= kEmpty) {
    for (int k = 0; k < N; ++k) { const size_t n = node(k); s.vol[n] = 0; s.level[n] = P[static_cast<size_t>(k)].zmin; s.sarea[n] = 0; }
    return;
  }
  if (V >= Vmax * kFull) {
    const double overV = jmax(0.0, V - Vmax);
    double aSum = 0;
    for (int k = 0; k < N; ++k) aSum += m.aov[node(k)];
    const double dh = overV / aSum;
    double assigned = 0;
    for (int k = 0; k < N - 1; ++k) {
      const double v = m.vmax[node(k)] + dh * m.aov[node(k)];
      s.vol[node(k)] = v;
      assigned += v;
    }
    s.vol[node(N - 1)] = V - assigned;
    for (int k = 0; k < N; ++k) {
      const size_t n = node(k);
      P[static_cast<size_t>(k)] = node_pass<Sums>(m, F, mem[k], hi);
      s.level[n] = hi + dh; s.sarea[n] = m.aov[n];
      set_centroid(s, n, P[static_cast<size_t>(k)]);
    }
    return;
  }
  if (!(h >= lo && h <= hi)) {
    h = jmin(hi, jmax(lo, h));
    for (int k = 0; k < N; ++k) P[static_cast<size_t>(k)] = node_pass<Sums>(m, F, mem[k], h);
  }
  const double tol = jmax(1e-3 * Vmax, 0.5);
  double a = lo, b = hi;
  for (int it = 0; it < kNewton; ++it) {
    double sum = 0;
    for (int k = 0; k < N; ++k) sum += P[static_cast<size_t>(k)].vol;
    const double err = sum - V;
    if (std::fabs(err) < tol) break;
    if (err > 0) b = h; else a = h;
    double dv = 0;
    for (int k = 0; k < N; ++k) dv += P[static_cast<size_t>(k)].dv;
    double hn = dv > kSlope ? h - err / dv : 0.5 * (a + b);
    if (!(hn > a && hn < b)) hn = 0.5 * (a + b);
    h = hn;
    for (int k = 0; k < N; ++k) P[static_cast<size_t>(k)] = node_pass<Sums>(m, F, mem[k], h);
  }
  // the water shared in proportion to the geometry at this level, capped by each member's capacity; when the last
  // member would overflow, the members before it take V - vmax(last), filled from the one nearest the end
  double tot = 0;
  for (int k = 0; k < N; ++k) tot += P[static_cast<size_t>(k)].vol;
  double assigned = 0;
  for (int k = 0; k < N - 1; ++k) {
    double v = tot > 0 ? V * P[static_cast<size_t>(k)].vol / tot : V / static_cast<double>(N);
    if (v > m.vmax[node(k)]) v = m.vmax[node(k)];
    s.vol[node(k)] = v;
    assigned += v;
  }
  const double vmaxLast = m.vmax[node(N - 1)];
  if (V - assigned > vmaxLast) {
    const double needV = V - vmaxLast;
    double capped = 0;
    for (int k = N - 2; k >= 0; --k) {
      double prior = 0;
      for (int j = 0; j < k; ++j) prior += s.vol[node(j)];
      const double v = (needV - capped) - prior;
      if (v > m.vmax[node(k)] && k > 0) {
        s.vol[node(k)] = m.vmax[node(k)];
        capped += m.vmax[node(k)];
        continue;
      }
      s.vol[node(k)] = v;
      break;
    }
    assigned = 0;
    for (int k = 0; k < N - 1; ++k) assigned += s.vol[node(k)];
  }
  s.vol[node(N - 1)] = V - assigned;
  for (int k = 0; k < N; ++k) {
    const size_t n = node(k);
    s.level[n] = h; s.sarea[n] = P[static_cast<size_t>(k)].dv;
    set_centroid(s, n, P[static_cast<size_t>(k)]);
  }
}

// levels.hpp:166-194: per group, the merge decision (Always, or Threshold: every member wet above the group's
// reference point plus the margin), then a shared or a per-node solve
template <class Sums>
void levels(const HostModel& m, const Frame& F, Inst& s) {
  for (int g = 0; g < m.ng; ++g) {
    const size_t gi = static_cast<size_t>(g);
    const int N = m.gCount[gi];
    const int32_t* mem = m.gNodes.data() + m.gStart[gi];
    bool merged = false;
    if (N >= 2) {
      if (m.gMode[gi] == 1) {
        merged = true;
      } else if (m.gMode[gi] == 2) {
        const double zMw = world_z(F, m.gpx[gi], m.gpy[gi], m.gpz[gi]);
        merged = true;
        for (int k = 0; k < N; ++k) {
          const size_t n = static_cast<size_t>(mem[k]);
          if (!(s.vol[n] > kEmpty && s.level[n] > zMw + m.gMargin)) { merged = false; break; }
        }
      }
    }
    s.merged[gi] = merged ? 1 : 0;
    if (merged) level_group<Sums>(m, F, s, mem, N);
    else
      for (int k = 0; k < N; ++k) level_single<Sums>(m, F, s, mem[k]);
  }
}

// ------------------------------------------------------------------------------------------- (c) network

// flows.hpp:38-78: the volume moved through connection c this step (a -> b positive) after the limiter; q[c] too
template <class Math>
double flow(const HostModel& m, const Frame& F, Inst& s, size_t c) {
  if (!s.en[c] || s.tOn[c] > s.t) { s.q[c] = 0; return 0; }
  if (m.skipGroup[c] >= 0 && s.merged[static_cast<size_t>(m.skipGroup[c])]) { s.q[c] = 0; return 0; }
  const int a = m.ca[c], b = m.cb[c];
  const double zo = F.R20 * m.cx[c] + F.R21 * m.cy[c] + F.R22 * m.cz[c] + F.tz;
  const double ha = a < 0 ? 0.0 : s.hEff[static_cast<size_t>(a)], hb = s.hEff[static_cast<size_t>(b)];
  const double Ha = ha > zo ? ha - zo : 0.0, Hb = hb > zo ? hb - zo : 0.0;
  double q = 0;
  if (m.law[c] == 0) {
    const double d = Ha - Hb;
    if (d != 0) q = m.coef[c] * s.area[c] * m.sq2g * d / std::sqrt(jmax(std::fabs(d), kHeadFloor));
  } else {
    const double H1 = Ha > Hb ? Ha : Hb;
    if (H1 > 0) {
      const double H2 = Ha > Hb ? Hb : Ha;
      double f = 1;
      if (H2 > 0) f = Math::pow(jmax(0.0, 1 - Math::pow(H2 / H1, kWeirExp)), kVilExp);
      q = m.coef[c] * s.area[c] * kWeirC * m.sq2g * H1 * std::sqrt(H1) * f * (Ha >= Hb ? 1.0 : -1.0);
    }
  }
  if (q == 0) { s.q[c] = 0; return 0; }
  double dV = q * m.dt;
  const int src = dV > 0 ? a : b, dst = dV > 0 ? b : a;
  const double hs = src < 0 ? 0.0 : s.hEff[static_cast<size_t>(src)];
  const double hd = dst < 0 ? 0.0 : s.hEff[static_cast<size_t>(dst)];
  const double inv = (src < 0 ? 0.0 : 1 / s.aEff[static_cast<size_t>(src)]) + (dst < 0 ? 0.0 : 1 / s.aEff[static_cast<size_t>(dst)]);
  const int dS = src < 0 ? 0 : s.deg[static_cast<size_t>(src)], dD = dst < 0 ? 0 : s.deg[static_cast<size_t>(dst)];
  const int degRef = src < 0 ? dD : (dst < 0 ? dS : (dS > dD ? dS : dD));
  const double share = kShare / static_cast<double>(degRef > 1 ? degRef : 1);
  double lim = inv > 0 ? share * (hs - jmax(hd, zo)) / inv : std::fabs(dV);
  if (lim < 0) lim = 0;
  if (std::fabs(dV) > lim) dV = dV > 0 ? lim : -lim;
  s.q[c] = dV / m.dt;
  return dV;
}

// ------------------------------------------------------------------------------------------- the step

// step.hpp:41-54: pose, buoyancy, levels, flows (heads, degrees, connections, gathers, sea inflow), volume
// updates with the excess returned in node order, loads, integration, events
template <class Math, class Sums>
void step(const HostModel& m, Inst& s) {
  const Frame F = pose<Math>(m, s.zO, s.pitch, s.roll);
  double hV, hx, hy, hz, htop;
  buoyancy<Sums>(m, F, hV, hx, hy, hz, htop);
  levels<Sums>(m, F, s);
  const size_t NN = static_cast<size_t>(m.nn), NC = static_cast<size_t>(m.nc);
  // flows.hpp:15-35: effective heads, then the limiter degrees (enabled and not between two empty spaces)
  for (size_t n = 0; n < NN; ++n) {
    s.hEff[n] = s.vol[n] > kEmpty ? s.level[n] : kNoHead;
    s.aEff[n] = s.vol[n] >= m.vmax[n] ? m.aov[n] : jmax(s.sarea[n], m.amin[n]);
  }
  for (size_t n = 0; n < NN; ++n) {
    int d = 0;
    const size_t e = static_cast<size_t>(m.incStart[n] + m.incCount[n]);
    for (size_t k = static_cast<size_t>(m.incStart[n]); k < e; ++k) {
      const size_t c = static_cast<size_t>(m.incConn[k]);
      if (!s.en[c]) continue;
      const int a = m.ca[c], b = m.cb[c];
      if (a >= 0 && s.hEff[static_cast<size_t>(a)] < kNoHeadTest && s.hEff[static_cast<size_t>(b)] < kNoHeadTest) continue;
      ++d;
    }
    s.deg[n] = d;
  }
  for (size_t c = 0; c < NC; ++c) s.dV[c] = flow<Math>(m, F, s, c);
  // flows.hpp:81-96: gathers in increasing connection order, the sea's connections in increasing order
  for (size_t n = 0; n < NN; ++n) {
    double acc = 0;
    const size_t e = static_cast<size_t>(m.incStart[n] + m.incCount[n]);
    for (size_t k = static_cast<size_t>(m.incStart[n]); k < e; ++k) {
      const double d = s.dV[static_cast<size_t>(m.incConn[k])];
      acc = m.incSide[k] ? acc + d : acc - d;
    }
    s.acc[n] = acc;
  }
  double seaIn = 0;
  for (int32_t c : m.sea) seaIn += s.dV[static_cast<size_t>(c)];
  // flows.hpp:99-118: clip at 0, cap a pressed-up space, the excess back to the sea in node order
  for (size_t n = 0; n < NN; ++n) {
    double v = s.vol[n] + s.acc[n];
    if (v < 0) v = 0;
    double excess = 0;
    if (v > m.vmax[n]) {
      const double cap = m.vmax[n] + m.aov[n] * jmax(0.0, -s.zmax[n]);
      if (v > cap) { excess = v - cap; v = cap; }
    }
    s.excess[n] = excess;
    s.vol[n] = v;
  }
  for (size_t n = 0; n < NN; ++n)
    if (s.excess[n] > 0) seaIn -= s.excess[n];
  s.inflow = seaIn / m.dt;
  // body.hpp:11-40: buoyancy and the flood water at their centroids, nodes in order
  const double W = m.rho * m.g;
  const double Fb = W * hV;
  const double bxw = F.R00 * hx + F.R01 * hy + F.R02 * hz + F.tx;
  const double byw = F.R10 * hx + F.R11 * hy + F.R12 * hz + F.ty;
  double tauX = (byw - m.YO) * Fb, tauY = -(bxw - m.XO) * Fb, Fz = Fb - m.g * m.ms, mw = 0, Ip = 0, Ir = 0;
  for (size_t n = 0; n < NN; ++n) {
    const double V = s.vol[n];
    if (V <= 0) continue;
    const double x = s.cx[n], y = s.cy[n], z = s.cz[n];
    const double wx = F.R00 * x + F.R01 * y + F.R02 * z + F.tx - m.XO;
    const double wy = F.R10 * x + F.R11 * y + F.R12 * z + F.ty - m.YO;
    const double wz = F.R20 * x + F.R21 * y + F.R22 * z + F.tz - s.zO;
    const double f = -W * V;
    tauX += wy * f; tauY -= wx * f; Fz += f;
    const double mm = m.rho * V;
    mw += mm;
    Ip += mm * (wx * wx + wz * wz);
    Ir += mm * (wy * wy + wz * wz);
  }
  // body.hpp:42-57: semi-implicit Euler with linear and quadratic damping, the clamps
  const double Mth = tauY, Mph = tauX * F.cth;
  const double mh = m.mh0 + mw, Ith = m.Ith0 + Ip, Iph = m.Iph0 + Ir, dt = m.dt;
  const double az = (Fz - m.cDz * s.vz - m.qz * std::fabs(s.vz) * s.vz) / mh;
  const double ath = (Mth - m.cDth * s.wth - m.qth * std::fabs(s.wth) * s.wth) / Ith;
  const double aph = (Mph - m.cDph * s.wph - m.qph * std::fabs(s.wph) * s.wph) / Iph;
  s.vz += az * dt; s.wth += ath * dt; s.wph += aph * dt;
  s.zO += s.vz * dt; s.pitch += s.wth * dt; s.roll += s.wph * dt;
  if (s.pitch > kPitchMax) { s.pitch = kPitchMax; s.wth = 0; }
  if (s.pitch < -kPitchMax) { s.pitch = -kPitchMax; s.wth = 0; }
  if (s.roll > kRollMax) { s.roll = kRollMax; s.wph = 0; }
  if (s.roll < -kRollMax) { s.roll = -kRollMax; s.wph = 0; }
  s.t += dt;
  s.Vb = hV; s.hullTop = htop; s.mw = mw; s.bx = hx; s.by = hy; s.bz = hz;
  // events.hpp:11-38: monitor sums in connection order, first crossings, marks, the founder rule
  std::fill(s.over.begin(), s.over.end(), 0.0);
  for (size_t c = 0; c < NC; ++c)
    if (m.monitor[c] >= 0 && s.q[c] > 0) s.over[static_cast<size_t>(m.monitor[c])] += s.q[c];
  for (int k = 0; k < m.nmon; ++k) {
    const size_t i = static_cast<size_t>(k);
    if (s.over[i] > m.monThreshold && s.overT[i] < 0) { s.overT[i] = s.t; s.push_event(m, s.t, k, 0); }
  }
  for (int k = 0; k < m.nmark; ++k) {
    const size_t i = static_cast<size_t>(k);
    if (s.markT[i] >= 0) continue;
    const double wz = world_z(F, m.mkx[i], m.mky[i], m.mkz[i]);
    const bool under = m.mkWhen[i] == 0;
    if ((under && wz < 0) || (!under && wz > 0)) { s.markT[i] = s.t; s.push_event(m, s.t, k, 1); }
  }
  if (!s.foundered && (s.hullTop < m.hullTopBelow || std::fabs(s.pitch) > m.pitchAbsAbove || (s.Vb < m.buoyancyBelow && s.t > m.minT))) {
    s.foundered = true;
    s.founderT = s.t;
    s.push_event(m, s.t, -1, 2);
  }
}

// Simulation::evaluate_hydro (simulation.cpp:137-145): the outputs of the current pose before any step
template <class Math, class Sums>
void evaluate(const HostModel& m, Inst& s) {
  const Frame F = pose<Math>(m, s.zO, s.pitch, s.roll);
  buoyancy<Sums>(m, F, s.Vb, s.bx, s.by, s.bz, s.hullTop);
  s.inflow = 0;
  double mw = 0;
  for (size_t n = 0; n < static_cast<size_t>(m.nn); ++n) mw += m.rho * s.vol[n];
  s.mw = mw;
}

// readouts.hpp:20-36 as a curve record: t trim list water(t) inflow(t/min) draftF draftA
template <class Math>
void readouts(const HostModel& m, const Inst& s, double* r) {
  const Frame F = pose<Math>(m, s.zO, s.pitch, s.roll);
  double water = 0;
  for (size_t n = 0; n < static_cast<size_t>(m.nn); ++n) water += s.vol[n];
  r[0] = s.t;
  r[1] = s.pitch * 180 / kPi;
  r[2] = s.roll * 180 / kPi;
  r[3] = water * m.rho / 1000;
  r[4] = s.inflow * m.rho / 1000 * 60;
  r[5] = -world_z(F, m.xFP, 0, 0);
  r[6] = -world_z(F, m.xAP, 0, 0);
}

// the trace layout: t zO pitch roll vz wth wph inflow Vb hullTop, then vol level cx cy cz per node
void record(const HostModel& m, const Inst& s, double* r) {
  const double head[10] = {s.t, s.zO, s.pitch, s.roll, s.vz, s.wth, s.wph, s.inflow, s.Vb, s.hullTop};
  std::copy(head, head + 10, r);
  const size_t nn = static_cast<size_t>(m.nn);
  double* p = r + 10;
  for (const std::vector<double>* v : {&s.vol, &s.level, &s.cx, &s.cy, &s.cz}) { std::copy(v->begin(), v->end(), p); p += nn; }
}

// The four numerics as (math, order) pairs (simulation.cpp:147-154).
struct Engine {
  void (*step)(const HostModel&, Inst&);
  void (*evaluate)(const HostModel&, Inst&);
  void (*readouts)(const HostModel&, const Inst&, double*);
};

Engine engine(int numerics) {
  switch (numerics) {
    case SS_NUMERICS_ORACLE: return Engine{step<OracleMath, Serial>, evaluate<OracleMath, Serial>, readouts<OracleMath>};
    case SS_NUMERICS_PORTABLE: return Engine{step<PortableMath, Tree<double>>, evaluate<PortableMath, Tree<double>>, readouts<PortableMath>};
    case SS_NUMERICS_PORTABLE32: return Engine{step<PortableMath, Tree<float>>, evaluate<PortableMath, Tree<float>>, readouts<PortableMath>};
    case SS_NUMERICS_STD: return Engine{step<StdMath, Serial>, evaluate<StdMath, Serial>, readouts<StdMath>};
    default: ssa::fail(SS_INVALID_ARGUMENT, "ss_cpu: numerics must be oracle (0), portable (1), portable32 (2) or std (3)");
  }
}

void fill_result(const Inst& s, int64_t steps, ss_instance_result* r) {
  r->foundered = s.foundered ? 1 : 0;
  r->founderT = s.founderT;
  r->tEnd = s.t;
  r->steps = steps;
  const double st[7] = {s.t, s.zO, s.pitch, s.roll, s.vz, s.wth, s.wph}, out[7] = {s.inflow, s.Vb, s.hullTop, s.mw, s.bx, s.by, s.bz};
  std::copy(st, st + 7, r->state);
  std::copy(out, out + 7, r->outputs);
  r->events = static_cast<int32_t>(s.events.size());
}

void copy_events(const Inst& s, ss_event* out, int32_t cap) {
  if (!out || cap <= 0) return;
  const size_t n = std::min(s.events.size(), static_cast<size_t>(cap));
  for (size_t k = 0; k < n; ++k) out[k] = ss_event{s.events[k].t, s.events[k].id, s.events[k].kind};
}

}  // namespace

extern "C" {

ss_status ss_cpu_sweep(const ss_ship* ship, const ss_sim* sim, const ss_cpu_options* opt, int32_t count, const uint8_t* en, const double* area,
                       const double* tOn, ss_instance_result* results, ss_event* events, int32_t evCap, double* curves, int32_t curveCap,
                       int32_t* curveCount, double* final7, double* seconds) {
  return ssa::guarded([&] {
    using ssa::need;
    need(ship && sim && opt, "ss_cpu_sweep: null argument");
    need(count >= 0, "ss_cpu_sweep: negative count");
    need(opt->tMax >= 0 && std::isfinite(opt->tMax), "ss_cpu_sweep: tMax must be finite and not negative");
    need(opt->curveEvery > 0, "ss_cpu_sweep: curveEvery must be positive");
    need(evCap >= 0 && curveCap >= 0, "ss_cpu_sweep: negative capacity");
    const Engine E = engine(opt->numerics);
    const HostModel m = make_model(*ship, *sim);
    const size_t nc = static_cast<size_t>(m.nc);
    int threads = opt->threads > 0 ? opt->threads : static_cast<int>(std::thread::hardware_concurrency());
    threads = std::max(1, std::min(threads, std::max<int32_t>(count, 1)));
    std::atomic<int32_t> next{0};
    std::atomic<bool> failed{false};
    // Simulation::run (simulation.cpp:167-192): a curve record whenever t reaches the next multiple of every, the step,
    // the calm rule past 30 min; a closing record and the final readouts
    auto work = [&] {
      try {
        Inst s;
        for (int32_t i = next.fetch_add(1); i < count && !failed.load(); i = next.fetch_add(1)) {
          const size_t k = static_cast<size_t>(i);
          s.reset(m, en ? en + k * nc : nullptr, area ? area + k * nc : nullptr, tOn ? tOn + k * nc : nullptr);
          E.evaluate(m, s);
          double* cur = curves ? curves + k * static_cast<size_t>(curveCap) * 7 : nullptr;
          int32_t nrec = 0;
          auto keep = [&] {
            double r[7];
            E.readouts(m, s, r);
            if (cur && nrec < curveCap) std::copy(r, r + 7, cur + static_cast<size_t>(nrec) * 7);
            ++nrec;
          };
          double due = 0, calm = 0;
          int64_t steps = 0;
          while (s.t < opt->tMax && !s.foundered) {
            if (s.t >= due) { keep(); due += opt->curveEvery; }
            E.step(m, s);
            ++steps;
            if (opt->stopWhenStable && s.t > 1800) {
              calm = (std::fabs(s.inflow) < 0.02 && std::fabs(s.wth) < 1e-5 && std::fabs(s.vz) < 1e-3) ? calm + m.dt : 0;
              if (calm > 1200) break;
            }
          }
          keep();
          if (curveCount) curveCount[k] = nrec;
          if (final7) E.readouts(m, s, final7 + k * 7);
          if (results) fill_result(s, steps, results + k);
          copy_events(s, events ? events + k * static_cast<size_t>(evCap) : nullptr, evCap);
        }
      } catch (...) {
        failed.store(true);
      }
    };
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (std::thread& t : pool) t.join();
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (failed.load()) ssa::fail(SS_RUNTIME, "ss_cpu_sweep: out of host memory in a worker");
  });
}

ss_status ss_cpu_trace(const ss_ship* ship, const ss_sim* sim, int32_t numerics, double tMax, int32_t denseSteps, double sampleEvery,
                       const uint8_t* en, const double* area, const double* tOn, double* dense, int32_t denseCap, int32_t* denseCount,
                       double* samples, int32_t samplesCap, int32_t* samplesCount, ss_instance_result* result, ss_event* events,
                       int32_t evCap) {
  return ssa::guarded([&] {
    using ssa::need;
    need(ship && sim, "ss_cpu_trace: null argument");
    need(tMax > 0 && std::isfinite(tMax) && sampleEvery > 0 && denseSteps >= 0, "ss_cpu_trace: bad trace options");
    need(denseCap >= 0 && samplesCap >= 0 && evCap >= 0, "ss_cpu_trace: negative capacity");
    const Engine E = engine(numerics);
    const HostModel m = make_model(*ship, *sim);
    const size_t rs = static_cast<size_t>(10 + 5 * m.nn);
    Inst s;
    s.reset(m, en, area, tOn);
    // Simulation::record_trace (simulation.cpp:194-221): the initial record, a dense record after each of the first
    // denseSteps steps, a sample whenever t reaches the next multiple of sampleEvery
    E.evaluate(m, s);
    int32_t nd = 0, ns = 0;
    auto put = [&](double* buf, int32_t cap, int32_t& n) {
      if (buf && n < cap) record(m, s, buf + static_cast<size_t>(n) * rs);
      ++n;
    };
    put(dense, denseCap, nd);
    put(samples, samplesCap, ns);
    double due = sampleEvery;
    int64_t steps = 0;
    while (!s.foundered && s.t < tMax) {
      E.step(m, s);
      ++steps;
      if (steps <= denseSteps) put(dense, denseCap, nd);
      if (s.t >= due - 1e-9) { put(samples, samplesCap, ns); due += sampleEvery; }
    }
    if (denseCount) *denseCount = nd;
    if (samplesCount) *samplesCount = ns;
    if (result) fill_result(s, steps, result);
    copy_events(s, events, evCap);
  });
}

}  // extern "C"
