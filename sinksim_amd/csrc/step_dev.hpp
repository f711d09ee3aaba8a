// The flooding step on gfx950: one 256-thread workgroup (four wave64s) per simulation, the simulation's state,
// scratch and connection overrides in LDS, the shared hull geometry streamed from L2.
//
// Bits: the step reproduces the reference's portable numerics (double) and portable32 numerics (mixed) exactly
// (reference: engine/include/sinksim/kernel/step.hpp:140-153 with Order::Tree / Order::TreeMixed). Three sums
// have a canonical order and are reproduced as such:
//   - buoyancy: 256 logical partials, partial t holding columns t, t+256, ...; each 32-partial group reduced by
//     the halving tree (l absorbs l+16, l+8, ... l+1); the 8 group results combined 0..7 in order
//     (reduce.hpp:12-107, hydro.hpp:242-258). Here a logical 32-lane group is one half of a wave64.
//   - node pass: 32 lane partials over the node's segments with stride 32, then the same tree (hydro.hpp:306-316).
//   - every other sum (sea inflow, node gathers, loads, monitors) is serial in the reference's index order.
// All other arithmetic is element-wise and is written operation for operation as in the reference; compile with
// -ffp-contract=off so that nothing is fused.
//
// Work layout per step (barriers marked |):
//   pose (4 lanes evaluate cos/sin of pitch and roll) | buoyancy partials + one level-solve unit per 32 lanes
//   | per node: effective heads, degrees | per connection: flows | per node: gather, volume update, load terms;
//   monitor sums; sea inflow | lanes 0-7 of wave 0: the serial tail (loads, integration, events) |
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fdlibm_dev.hpp"

namespace ssa {

constexpr int kThreads = 256;           // logical block of the canonical buoyancy order
constexpr int kLanes = 32;              // logical warp of the canonical trees
constexpr int kUnits = kThreads / kLanes;

// scheme v1 (reference: engine/include/sinksim/kernel/scheme.hpp:11-28)
constexpr double kEmptyVolume = 1e-6;
constexpr double kFullFraction = 1.0 - 1e-7;
constexpr double kLimiterShare = 0.5;
constexpr double kOrificeHeadFloor = 0.01;
constexpr double kWeirCoefficient = 2.0 / 3.0;
constexpr double kWeirExponent = 1.5;
constexpr double kVillemonteExponent = 0.385;
constexpr int kNewtonFast = 6;
constexpr double kSlopeMin = 1e-9;
constexpr double kCentroidMinVolume = 1e-9;
constexpr double kNoHead = -1e9;
constexpr double kNoHeadTest = -1e8;
constexpr double kPitchClamp = 1.35;
constexpr double kRollClamp = 1.4;
constexpr double kPi = 3.141592653589793;

// JavaScript Math.max / Math.min for ordinary numbers (reference: kernel/math.hpp:338-339)
__device__ __forceinline__ double jmax(double a, double b) { return (b > a) ? b : a; }
__device__ __forceinline__ double jmin(double a, double b) { return (b < a) ? b : a; }

// ------------------------------------------------------------------------------------------- device views

// The hull, shared by every simulation. Segment arrays are pre-gathered per segment (x, y and footprint of the
// segment's column) so that a node pass reads five contiguous streams; the f32 copies are the same values
// rounded once on the host, which is what the portable32 arithmetic converts them to anyway.
struct DevShip {
  int32_t ncol, nseg, nn;
  const double *x, *y, *dxdy, *zlo, *zhi;
  const float *xf, *yf, *dxdyf, *zlof, *zhif;
  const int32_t *segStart, *segCount;
  const double *sx, *sy, *sA, *sa, *se;
  const float *sxf, *syf, *sAf, *saf, *sef;
  const double *mu, *vmax, *amin, *aov;
};

// The scenario's immutable network and constants.
struct DevNet {
  double rho, g, sq2g, dt;
  int32_t ng;
  const int32_t *gStart, *gCount, *gNodes;
  const uint8_t* gMode;
  const double *gpx, *gpy, *gpz;
  double gMargin;
  int32_t nc;
  const int32_t *ca, *cb;
  const uint8_t* law;
  const double *cx, *cy, *cz, *coef;
  const int32_t* skipGroup;
  const int32_t *incStart, *incCount, *incConn;
  const uint8_t* incSide;
  int32_t seaCount;
  const int32_t* sea;
  int32_t nmon;
  double monThreshold;
  const int32_t *monStart, *monCount, *monConn;  // per monitor, its connections in increasing order
  int32_t nmark;
  const double *mkx, *mky, *mkz;
  const uint8_t* mkWhen;
  double hullTopBelow, pitchAbsAbove, buoyancyBelow, minT;
  double ms, g0x, g0y, g0z, XO, YO, mh0, Ith0, Iph0, cDz, cDth, cDph, qz, qth, qph;
  double xFP, xAP;
  int32_t maxGroup, evCap;
};

// Per-instance global arrays of the batch.
struct DevBatch {
  int32_t count;
  const uint8_t* en;
  const double *area, *tOn;  // [count * nc]
  double* st;                // [count * 7]: t zO pitch roll vz wth wph
  double* out;               // [count * 7]: inflow Vb hullTop mw bx by bz
  double *vol, *level, *cx, *cy, *cz;  // [count * nn]
  double* evT;
  int32_t* evId;
  uint8_t* evKind;
  int32_t* evCount;
  double *overT, *markT;
  uint8_t* foundered;
  double* founderT;
  long long* steps;
  double* final7;            // [count * 7] readouts at the end of the last launch
  int32_t traced, denseSteps, maxSamples, recSize;
  double sampleEvery;
  double *denseRec, *sampleRec;
  int32_t *denseCount, *sampleCount;
  double* nextSample;
  double curveEvery;
  int32_t curveMax;
  double* curveRec;
  int32_t* curveCount;
  double* nextCurve;
};

// LDS layout of one simulation, as offsets in doubles (the host computes it: lds_layout in batch.hip).
struct LdsLayout {
  int32_t vol, level, cx, cy, cz, area, zmin, zmax, hEff, aEff, excess;
  int32_t tX, tY, tF, tM, tIp, tIr;
  int32_t dV, q, carea, ctOn, passes, warpHyd, trig, scal, over, overT, markT, evT;
  int32_t deg, evId;      // int32 arrays, offsets in doubles
  int32_t merged, cen, evKind;  // byte arrays, offsets in doubles
  int32_t total;          // doubles
};

// scalar slots of the `scal` block
enum : int { S_T = 0, S_ZO, S_PITCH, S_ROLL, S_VZ, S_WTH, S_WPH, O_INFLOW, O_VB, O_TOP, O_MW, O_BX, O_BY, O_BZ, S_SEA, S_FOUNDERT, S_EVCOUNT,
             S_FOUNDERED, S_COUNT };

struct Blk {
  double *vol, *level, *cx, *cy, *cz, *area, *zmin, *zmax, *hEff, *aEff, *excess;
  double *tX, *tY, *tF, *tM, *tIp, *tIr;
  double *dV, *q, *carea, *ctOn, *passes, *warpHyd, *trig, *scal, *over, *overT, *markT, *evT;
  int32_t *deg, *evId;
  uint8_t *merged, *cen, *evKind;
};

__device__ inline Blk carve(double* base, const LdsLayout& L) {
  Blk b;
  b.vol = base + L.vol; b.level = base + L.level; b.cx = base + L.cx; b.cy = base + L.cy; b.cz = base + L.cz;
  b.area = base + L.area; b.zmin = base + L.zmin; b.zmax = base + L.zmax; b.hEff = base + L.hEff; b.aEff = base + L.aEff;
  b.excess = base + L.excess;
  b.tX = base + L.tX; b.tY = base + L.tY; b.tF = base + L.tF; b.tM = base + L.tM; b.tIp = base + L.tIp; b.tIr = base + L.tIr;
  b.dV = base + L.dV; b.q = base + L.q; b.carea = base + L.carea; b.ctOn = base + L.ctOn; b.passes = base + L.passes;
  b.warpHyd = base + L.warpHyd; b.trig = base + L.trig; b.scal = base + L.scal; b.over = base + L.over; b.overT = base + L.overT;
  b.markT = base + L.markT; b.evT = base + L.evT;
  b.deg = reinterpret_cast<int32_t*>(base + L.deg); b.evId = reinterpret_cast<int32_t*>(base + L.evId);
  b.merged = reinterpret_cast<uint8_t*>(base + L.merged); b.cen = reinterpret_cast<uint8_t*>(base + L.cen);
  b.evKind = reinterpret_cast<uint8_t*>(base + L.evKind);
  return b;
}

// ------------------------------------------------------------------------------------------- pose

struct Frame {
  double R00, R01, R02, R10, R11, R12, R20, R21, R22, tx, ty, tz, cth, iR;
};
struct FrameF {
  float R20, R21, R22, tz, iR;
};

// reference: kernel/frame.hpp:10-21, from the four trigonometric values
__device__ inline Frame make_frame(const DevNet& M, double zO, double cth, double sth, double cph, double sph) {
  Frame F;
  F.R00 = cth; F.R01 = sth * sph; F.R02 = sth * cph;
  F.R10 = 0;   F.R11 = cph;       F.R12 = -sph;
  F.R20 = -sth; F.R21 = cth * sph; F.R22 = cth * cph;
  F.tx = M.XO - (F.R00 * M.g0x + F.R01 * M.g0y + F.R02 * M.g0z);
  F.ty = M.YO - (F.R10 * M.g0x + F.R11 * M.g0y + F.R12 * M.g0z);
  F.tz = zO - (F.R20 * M.g0x + F.R21 * M.g0y + F.R22 * M.g0z);
  F.cth = cth;
  F.iR = 1 / F.R22;  // hydro.hpp:209,245,264,309
  return F;
}

// reference: kernel/reduce.hpp:118-126
__device__ inline FrameF frame_f(const Frame& F) {
  FrameF f;
  f.R20 = static_cast<float>(F.R20);
  f.R21 = static_cast<float>(F.R21);
  f.R22 = static_cast<float>(F.R22);
  f.tz = static_cast<float>(F.tz);
  f.iR = 1.0f / f.R22;
  return f;
}

__device__ inline double world_z(const Frame& F, double x, double y, double z) { return F.R20 * x + F.R21 * y + F.R22 * z + F.tz; }

// cos(pitch), sin(pitch), cos(roll), sin(roll) for slot 0..3 (frame.hpp:11)
__device__ inline double pose_trig(int slot, double pitch, double roll) {
  return fdl::sin_or_cos(slot < 2 ? pitch : roll, (slot & 1) == 0 ? 1 : 0);
}

// ------------------------------------------------------------------------------------------- partials

template <class T>
struct HP {
  T V, Mx, My, Mz, top;
};
template <class T>
struct PP {
  T vol, dv, mx, my, mz, zmin, zmax;
};

template <class T>
__device__ __forceinline__ void absorb(HP<T>& a, const HP<T>& b) {
  a.V += b.V; a.Mx += b.Mx; a.My += b.My; a.Mz += b.Mz;
  if (b.top > a.top) a.top = b.top;
}
template <class T>
__device__ __forceinline__ void absorb(PP<T>& a, const PP<T>& b) {
  a.vol += b.vol; a.dv += b.dv; a.mx += b.mx; a.my += b.my; a.mz += b.mz;
  if (b.zmin < a.zmin) a.zmin = b.zmin;
  if (b.zmax > a.zmax) a.zmax = b.zmax;
}

// The halving tree of a logical 32-lane group (one half of the wave64): for off = 16 ... 1 lane l < off
// absorbs lane l + off. Lanes >= off keep stale values that the tree never reads again.
template <class T>
__device__ __forceinline__ T down(T v, int off) { return __shfl_down(v, off, kLanes); }

template <class T>
__device__ __forceinline__ void tree32(HP<T>& p, int lane) {
#pragma unroll
  for (int off = kLanes / 2; off > 0; off >>= 1) {
    HP<T> o;
    o.V = down(p.V, off); o.Mx = down(p.Mx, off); o.My = down(p.My, off); o.Mz = down(p.Mz, off); o.top = down(p.top, off);
    if (lane < off) absorb(p, o);
  }
}
template <class T>
__device__ __forceinline__ void tree32(PP<T>& p, int lane) {
#pragma unroll
  for (int off = kLanes / 2; off > 0; off >>= 1) {
    PP<T> o;
    o.vol = down(p.vol, off); o.dv = down(p.dv, off); o.mx = down(p.mx, off); o.my = down(p.my, off); o.mz = down(p.mz, off);
    o.zmin = down(p.zmin, off); o.zmax = down(p.zmax, off);
    if (lane < off) absorb(p, o);
  }
}

// lane 0 of the logical group to every lane of it
template <class T>
__device__ __forceinline__ T from_lead(T v) { return __shfl(v, 0, kLanes); }

// one column into a buoyancy partial (reduce.hpp:67-77 / 159-170)
__device__ __forceinline__ void column(const DevShip& S, const Frame& F, int i, HP<double>& p) {
  const double x = S.x[i], y = S.y[i], lo = S.zlo[i], hi = S.zhi[i];
  const double base = F.R20 * x + F.R21 * y + F.tz;
  const double wt = base + F.R22 * hi;
  if (wt > p.top) p.top = wt;
  const double zp = -base * F.iR;
  if (zp <= lo) return;
  const double s = zp >= hi ? hi - lo : zp - lo;
  const double v = S.dxdy[i] * s;
  p.V += v; p.Mx += v * x; p.My += v * y; p.Mz += v * (lo + 0.5 * s);
}
__device__ __forceinline__ void column(const DevShip& S, const FrameF& f, int i, HP<float>& p) {
  const float x = S.xf[i], y = S.yf[i], lo = S.zlof[i], hi = S.zhif[i];
  const float base = f.R20 * x + f.R21 * y + f.tz;
  const float wt = base + f.R22 * hi;
  if (wt > p.top) p.top = wt;
  const float zp = -base * f.iR;
  if (zp <= lo) return;
  const float s = zp >= hi ? hi - lo : zp - lo;
  const float v = S.dxdyf[i] * s;
  p.V += v; p.Mx += v * x; p.My += v * y; p.Mz += v * (lo + 0.5f * s);
}

// one segment into a node-pass partial (reduce.hpp:80-99 / 172-192)
__device__ __forceinline__ void segment(const DevShip& S, const Frame& F, double h, int j, PP<double>& p) {
  const double x = S.sx[j], y = S.sy[j], a = S.sa[j], b = S.se[j];
  const double base = F.R20 * x + F.R21 * y + F.tz;
  const double wa = base + F.R22 * a, wb = base + F.R22 * b;
  if (wa < p.zmin) p.zmin = wa;
  if (wb > p.zmax) p.zmax = wb;
  const double zp = (h - base) * F.iR;
  if (zp <= a) return;
  const double A = S.sA[j];
  double L;
  if (zp >= b) {
    L = b - a;
  } else {
    L = zp - a;
    p.dv += A;
  }
  const double v = A * L;
  p.vol += v; p.mx += v * x; p.my += v * y; p.mz += v * (a + 0.5 * L);
}
__device__ __forceinline__ void segment(const DevShip& S, const FrameF& f, float h, int j, PP<float>& p) {
  const float x = S.sxf[j], y = S.syf[j], a = S.saf[j], b = S.sef[j];
  const float base = f.R20 * x + f.R21 * y + f.tz;
  const float wa = base + f.R22 * a, wb = base + f.R22 * b;
  if (wa < p.zmin) p.zmin = wa;
  if (wb > p.zmax) p.zmax = wb;
  const float zp = (h - base) * f.iR;
  if (zp <= a) return;
  const float A = S.sAf[j];
  float L;
  if (zp >= b) {
    L = b - a;
  } else {
    L = zp - a;
    p.dv += A;
  }
  const float v = A * L;
  p.vol += v; p.mx += v * x; p.my += v * y; p.mz += v * (a + 0.5f * L);
}

struct Pass {
  double vol, dv, mx, my, mz, zmin, zmax;
};

// The node pass done by one logical 32-lane group; the result is in every lane of the group.
// (hydro.hpp:306-316 / 338-349; finish_pass hydro.hpp:295-303)
template <bool Mixed>
struct UnitPass {
  const DevShip& S;
  const Frame& F;
  const FrameF& f;
  int lane;
  __device__ __forceinline__ Pass operator()(int n, double h) const {
    const int s0 = S.segStart[n], s1 = s0 + S.segCount[n];
    double vol, dv, mx, my, mz, zmin, zmax;
    if constexpr (Mixed) {
      const float hf = static_cast<float>(h);
      PP<float> p{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1e9f, -1e9f};
      for (int j = s0 + lane; j < s1; j += kLanes) segment(S, f, hf, j, p);
      tree32(p, lane);
      vol = static_cast<double>(from_lead(p.vol)); dv = static_cast<double>(from_lead(p.dv));
      mx = static_cast<double>(from_lead(p.mx)); my = static_cast<double>(from_lead(p.my)); mz = static_cast<double>(from_lead(p.mz));
      zmin = static_cast<double>(from_lead(p.zmin)); zmax = static_cast<double>(from_lead(p.zmax));
    } else {
      PP<double> p{0.0, 0.0, 0.0, 0.0, 0.0, 1e9, -1e9};
      for (int j = s0 + lane; j < s1; j += kLanes) segment(S, F, h, j, p);
      tree32(p, lane);
      vol = from_lead(p.vol); dv = from_lead(p.dv); mx = from_lead(p.mx); my = from_lead(p.my); mz = from_lead(p.mz);
      zmin = from_lead(p.zmin); zmax = from_lead(p.zmax);
    }
    const double mu = S.mu[n];
    Pass r;
    r.vol = vol * mu;
    r.dv = dv * mu * F.iR;
    r.mx = mx * mu; r.my = my * mu; r.mz = mz * mu;
    r.zmin = zmin; r.zmax = zmax;
    return r;
  }
};

// ------------------------------------------------------------------------------------------- (b) levels

__device__ __forceinline__ void put_centroid(const Blk& B, int n, const Pass& p) {
  if (p.vol > kCentroidMinVolume) { B.cx[n] = p.mx / p.vol; B.cy[n] = p.my / p.vol; B.cz[n] = p.mz / p.vol; }
}

__device__ __forceinline__ void put_pass(double* P, int k, const Pass& p) {
  double* q = P + 7 * k;
  q[0] = p.vol; q[1] = p.dv; q[2] = p.mx; q[3] = p.my; q[4] = p.mz; q[5] = p.zmin; q[6] = p.zmax;
}
__device__ __forceinline__ Pass get_pass(const double* P, int k) {
  const double* q = P + 7 * k;
  return Pass{q[0], q[1], q[2], q[3], q[4], q[5], q[6]};
}

// levels.hpp:13-49 with full = false (the step never asks for the 40-iteration solve)
template <class PassFn>
__device__ void solve_single(const DevShip& S, const PassFn& pass, int n, const Blk& B, bool lead) {
  const double V = B.vol[n], Vmax = S.vmax[n];
  double h = B.level[n];
  Pass p = pass(n, h);
  const double lo = p.zmin, hi = p.zmax;
  if (lead) { B.zmin[n] = lo; B.zmax[n] = hi; }
  if (V <= kEmptyVolume) {
    if (lead) { B.level[n] = lo; B.area[n] = 0; }
    return;
  }
  if (V >= Vmax * kFullFraction) {
    p = pass(n, hi);
    if (lead) {
      B.level[n] = hi + jmax(0.0, V - Vmax) / S.aov[n];
      B.area[n] = S.aov[n];
      put_centroid(B, n, p);
    }
    return;
  }
  if (h < lo || h > hi) {
    h = jmin(hi, jmax(lo, h));
    p = pass(n, h);
  }
  const double tol = jmax(1e-3 * Vmax, 0.5);
  double a = lo, b = hi;
  for (int it = 0; it < kNewtonFast; ++it) {
    const double err = p.vol - V;
    if (__builtin_fabs(err) < tol) break;
    if (err > 0) b = h; else a = h;
    double hn = p.dv > kSlopeMin ? h - err / p.dv : 0.5 * (a + b);
    if (!(hn > a && hn < b)) hn = 0.5 * (a + b);
    h = hn;
    p = pass(n, h);
  }
  if (lead) {
    B.level[n] = h; B.area[n] = p.dv;
    put_centroid(B, n, p);
  }
}

// levels.hpp:54-162 with full = false. P is this unit's pass scratch in LDS (maxGroup entries).
template <class PassFn>
__device__ void solve_group(const DevShip& S, const PassFn& pass, const int32_t* members, int N, const Blk& B, double* P, bool lead) {
  double V = 0, Vmax = 0;
  for (int k = 0; k < N; ++k) { V += B.vol[members[k]]; Vmax += S.vmax[members[k]]; }
  bool allWet = true;
  int firstWet = -1;
  for (int k = 0; k < N; ++k) {
    if (B.vol[members[k]] > 0) { if (firstWet < 0) firstWet = k; }
    else allWet = false;
  }
  double h;
  if (allWet) {
    double s = 0;
    for (int k = 0; k < N; ++k) s += B.level[members[k]];
    h = s / static_cast<double>(N);
  } else {
    h = B.level[members[firstWet >= 0 ? firstWet : N - 1]];
  }
  double lo = 0, hi = 0;
  for (int k = 0; k < N; ++k) {
    const Pass p = pass(members[k], h);
    if (lead) { put_pass(P, k, p); B.zmin[members[k]] = p.zmin; B.zmax[members[k]] = p.zmax; }
    if (k == 0) { lo = p.zmin; hi = p.zmax; }
    else { if (p.zmin < lo) lo = p.zmin; if (p.zmax > hi) hi = p.zmax; }
  }
  if (V <= kEmptyVolume) {
    if (lead) {
      for (int k = 0; k < N; ++k) { const int n = members[k]; B.vol[n] = 0; B.level[n] = get_pass(P, k).zmin; B.area[n] = 0; }
    }
    return;
  }
  if (V >= Vmax * kFullFraction) {
    const double over = jmax(0.0, V - Vmax);
    double aSum = 0;
    for (int k = 0; k < N; ++k) aSum += S.aov[members[k]];
    const double dh = over / aSum;
    double assigned = 0;
    for (int k = 0; k < N - 1; ++k) {
      const double v = S.vmax[members[k]] + dh * S.aov[members[k]];
      if (lead) B.vol[members[k]] = v;
      assigned += v;
    }
    if (lead) B.vol[members[N - 1]] = V - assigned;
    for (int k = 0; k < N; ++k) {
      const int n = members[k];
      const Pass p = pass(n, hi);
      if (lead) {
        B.level[n] = hi + dh; B.area[n] = S.aov[n];
        put_centroid(B, n, p);
      }
    }
    return;
  }
  if (!(h >= lo && h <= hi)) {
    h = jmin(hi, jmax(lo, h));
    for (int k = 0; k < N; ++k) {
      const Pass p = pass(members[k], h);
      if (lead) put_pass(P, k, p);
    }
  }
  const double tol = jmax(1e-3 * Vmax, 0.5);
  double a = lo, b = hi;
  for (int it = 0; it < kNewtonFast; ++it) {
    double sum = 0;
    for (int k = 0; k < N; ++k) sum += P[7 * k + 0];
    const double err = sum - V;
    if (__builtin_fabs(err) < tol) break;
    if (err > 0) b = h; else a = h;
    double dv = 0;
    for (int k = 0; k < N; ++k) dv += P[7 * k + 1];
    double hn = dv > kSlopeMin ? h - err / dv : 0.5 * (a + b);
    if (!(hn > a && hn < b)) hn = 0.5 * (a + b);
    h = hn;
    for (int k = 0; k < N; ++k) {
      const Pass p = pass(members[k], h);
      if (lead) put_pass(P, k, p);
    }
  }
  // share the water in proportion to the geometry at this level (levels.hpp:125-161)
  if (!lead) return;
  double tot = 0;
  for (int k = 0; k < N; ++k) tot += P[7 * k + 0];
  double assigned = 0;
  for (int k = 0; k < N - 1; ++k) {
    double v = tot > 0 ? V * P[7 * k + 0] / tot : V / static_cast<double>(N);
    if (v > S.vmax[members[k]]) v = S.vmax[members[k]];
    B.vol[members[k]] = v;
    assigned += v;
  }
  const double vmaxLast = S.vmax[members[N - 1]];
  if (V - assigned > vmaxLast) {
    const double need = V - vmaxLast;
    double capped = 0;
    for (int k = N - 2; k >= 0; --k) {
      double prior = 0;
      for (int j = 0; j < k; ++j) prior += B.vol[members[j]];
      const double v = (need - capped) - prior;
      if (v > S.vmax[members[k]] && k > 0) {
        B.vol[members[k]] = S.vmax[members[k]];
        capped += S.vmax[members[k]];
        continue;
      }
      B.vol[members[k]] = v;
      break;
    }
    assigned = 0;
    for (int k = 0; k < N - 1; ++k) assigned += B.vol[members[k]];
  }
  B.vol[members[N - 1]] = V - assigned;
  for (int k = 0; k < N; ++k) {
    const int n = members[k];
    const Pass p = get_pass(P, k);
    B.level[n] = h; B.area[n] = p.dv;
    put_centroid(B, n, p);
  }
}

// levels.hpp:166-188: the merge decision, then the solve, for group g
template <class PassFn>
__device__ void solve_group_entry(const DevShip& S, const DevNet& M, const Frame& F, const PassFn& pass, int g, const Blk& B, double* P,
                                  bool lead) {
  const int N = M.gCount[g];
  const int32_t* members = M.gNodes + M.gStart[g];
  const uint8_t mode = M.gMode[g];
  bool merged = false;
  if (N >= 2) {
    if (mode == 1) {
      merged = true;
    } else if (mode == 2) {
      const double zMw = world_z(F, M.gpx[g], M.gpy[g], M.gpz[g]);
      merged = true;
      for (int k = 0; k < N; ++k) {
        const int n = members[k];
        if (!(B.vol[n] > kEmptyVolume && B.level[n] > zMw + M.gMargin)) { merged = false; break; }
      }
    }
  }
  if (lead) B.merged[g] = merged ? 1 : 0;
  if (merged) solve_group(S, pass, members, N, B, P, lead);
  else for (int k = 0; k < N; ++k) solve_single(S, pass, members[k], B, lead);
}

// ------------------------------------------------------------------------------------------- (c) network

__device__ __forceinline__ double head_of(const Blk& B, int n) { return B.vol[n] > kEmptyVolume ? B.level[n] : kNoHead; }

// flows.hpp:211-231: effective head and area of node n, and the number of connections that count toward the
// limiter there (an enabled connection not between two empty spaces). The heads of the other ends are
// recomputed here from vol and level, so that one pass over the nodes suffices.
__device__ __forceinline__ void node_heads(const DevShip& S, const DevNet& M, const Blk& B, int n) {
  const double v = B.vol[n];
  B.hEff[n] = v > kEmptyVolume ? B.level[n] : kNoHead;
  B.aEff[n] = v >= S.vmax[n] ? S.aov[n] : jmax(B.area[n], S.amin[n]);
  const int s = M.incStart[n], e = s + M.incCount[n];
  int d = 0;
  for (int k = s; k < e; ++k) {
    const int c = M.incConn[k];
    if (!B.cen[c]) continue;
    const int a = M.ca[c], b = M.cb[c];
    if (a >= 0 && head_of(B, a) < kNoHeadTest && head_of(B, b) < kNoHeadTest) continue;
    ++d;
  }
  B.deg[n] = d;
}

// flows.hpp:234-274: the volume moved through connection c this step (a -> b positive), after the limiter
__device__ inline double connection_flow(const DevNet& M, const Frame& F, double t, const Blk& B, int c) {
  if (!B.cen[c] || B.ctOn[c] > t) { B.q[c] = 0; return 0; }
  if (M.skipGroup[c] >= 0 && B.merged[M.skipGroup[c]]) { B.q[c] = 0; return 0; }
  const int a = M.ca[c], b = M.cb[c];
  const double zo = F.R20 * M.cx[c] + F.R21 * M.cy[c] + F.R22 * M.cz[c] + F.tz;
  const double ha = a < 0 ? 0.0 : B.hEff[a], hbv = B.hEff[b];
  const double Ha = ha > zo ? ha - zo : 0.0, Hb = hbv > zo ? hbv - zo : 0.0;
  double q = 0;
  if (M.law[c] == 0) {
    const double d = Ha - Hb;
    if (d != 0) q = M.coef[c] * B.carea[c] * M.sq2g * d / __builtin_sqrt(jmax(__builtin_fabs(d), kOrificeHeadFloor));
  } else {
    const double H1 = Ha > Hb ? Ha : Hb;
    if (H1 > 0) {
      const double H2 = Ha > Hb ? Hb : Ha;
      double f = 1;
      if (H2 > 0) {
        const double r = H2 / H1;
        f = fdl::pow(jmax(0.0, 1 - fdl::pow(r, kWeirExponent)), kVillemonteExponent);
      }
      q = M.coef[c] * B.carea[c] * kWeirCoefficient * M.sq2g * H1 * __builtin_sqrt(H1) * f * (Ha >= Hb ? 1.0 : -1.0);
    }
  }
  if (q == 0) { B.q[c] = 0; return 0; }
  double dV = q * M.dt;
  const int src = dV > 0 ? a : b, dst = dV > 0 ? b : a;
  const double hs = src < 0 ? 0.0 : B.hEff[src];
  const double hd = dst < 0 ? 0.0 : B.hEff[dst];
  const double inv = (src < 0 ? 0.0 : 1 / B.aEff[src]) + (dst < 0 ? 0.0 : 1 / B.aEff[dst]);
  const int degRef = src < 0 ? B.deg[dst] : (dst < 0 ? B.deg[src] : (B.deg[src] > B.deg[dst] ? B.deg[src] : B.deg[dst]));
  const double share = kLimiterShare / static_cast<double>(degRef > 1 ? degRef : 1);
  double lim = inv > 0 ? share * (hs - jmax(hd, zo)) / inv : __builtin_fabs(dV);
  if (lim < 0) lim = 0;
  if (__builtin_fabs(dV) > lim) dV = dV > 0 ? lim : -lim;
  B.q[c] = dV / M.dt;
  return dV;
}

// flows.hpp:277-307 for node n (the gather in increasing connection order and the capped volume update), then
// the node's terms of the loads (body.hpp:21-34), which only depend on this node once its volume is final.
__device__ inline void node_update(const DevShip& S, const DevNet& M, const Frame& F, double zO, const Blk& B, int n) {
  const int s = M.incStart[n], e = s + M.incCount[n];
  double acc = 0;
  for (int k = s; k < e; ++k) {
    const double d = B.dV[M.incConn[k]];
    acc = M.incSide[k] ? acc + d : acc - d;
  }
  double v = B.vol[n] + acc;
  if (v < 0) v = 0;
  const double vm = S.vmax[n];
  double excess = 0;
  if (v > vm) {
    const double head = jmax(0.0, -B.zmax[n]);
    const double cap = vm + S.aov[n] * head;
    if (v > cap) { excess = v - cap; v = cap; }
  }
  B.excess[n] = excess;
  B.vol[n] = v;
  if (v > 0) {
    const double W = M.rho * M.g;
    const double x = B.cx[n], y = B.cy[n], z = B.cz[n];
    const double wx = F.R00 * x + F.R01 * y + F.R02 * z + F.tx - M.XO;
    const double wy = F.R10 * x + F.R11 * y + F.R12 * z + F.ty - M.YO;
    const double wz = F.R20 * x + F.R21 * y + F.R22 * z + F.tz - zO;
    const double f = -W * v;
    const double m = M.rho * v;
    B.tX[n] = wy * f;
    B.tY[n] = wx * f;
    B.tF[n] = f;
    B.tM[n] = m;
    B.tIp[n] = m * (wx * wx + wz * wz);
    B.tIr[n] = m * (wy * wy + wz * wz);
  }
}

// ------------------------------------------------------------------------------------------- readouts

// readouts.hpp:175-191 (7 values: t trim list water(t) inflow(t/min) draftF draftA), one thread
__device__ inline void readouts7(const DevNet& M, const Blk& B, int nn, double* r) {
  const double t = B.scal[S_T], zO = B.scal[S_ZO], pitch = B.scal[S_PITCH], roll = B.scal[S_ROLL];
  const Frame F = make_frame(M, zO, fdl::cos(pitch), fdl::sin(pitch), fdl::cos(roll), fdl::sin(roll));
  double water = 0;
  for (int n = 0; n < nn; ++n) water += B.vol[n];
  r[0] = t;
  r[1] = pitch * 180 / kPi;
  r[2] = roll * 180 / kPi;
  r[3] = water * M.rho / 1000;
  r[4] = B.scal[O_INFLOW] * M.rho / 1000 * 60;
  r[5] = -world_z(F, M.xFP, 0, 0);
  r[6] = -world_z(F, M.xAP, 0, 0);
}

__device__ inline void push_event(const DevNet& M, const Blk& B, double t, int id, int kind) {
  const int k = static_cast<int>(B.scal[S_EVCOUNT]);
  if (k < M.evCap) {
    B.evT[k] = t; B.evId[k] = id; B.evKind[k] = static_cast<uint8_t>(kind);
    B.scal[S_EVCOUNT] = k + 1;
  }
}

// ------------------------------------------------------------------------------------------- the step

// Pose and buoyancy partials: after this every thread holds the frame, and the lead lane of each 32-lane group
// has written its group's tree result to warpHyd. Ends without a barrier after the partials.
template <bool Mixed>
__device__ inline void pose_and_buoyancy(const DevShip& S, const DevNet& M, const Blk& B, int tid, Frame& F, FrameF& f) {
  if (tid < 4) B.trig[tid] = pose_trig(tid, B.scal[S_PITCH], B.scal[S_ROLL]);
  __syncthreads();
  F = make_frame(M, B.scal[S_ZO], B.trig[0], B.trig[1], B.trig[2], B.trig[3]);
  f = frame_f(F);
  const int lane = tid & (kLanes - 1), unit = tid / kLanes;
  if constexpr (Mixed) {
    HP<float> p{0.0f, 0.0f, 0.0f, 0.0f, -1e9f};
    for (int i = tid; i < S.ncol; i += kThreads) column(S, f, i, p);
    tree32(p, lane);
    if (lane == 0) {
      double* w = B.warpHyd + 5 * unit;
      w[0] = p.V; w[1] = p.Mx; w[2] = p.My; w[3] = p.Mz; w[4] = p.top;
    }
  } else {
    HP<double> p{0.0, 0.0, 0.0, 0.0, -1e9};
    for (int i = tid; i < S.ncol; i += kThreads) column(S, F, i, p);
    tree32(p, lane);
    if (lane == 0) {
      double* w = B.warpHyd + 5 * unit;
      w[0] = p.V; w[1] = p.Mx; w[2] = p.My; w[3] = p.Mz; w[4] = p.top;
    }
  }
}

// the 8 group results combined in order, then finish_hydro (hydro.hpp:231-239, 255-257)
__device__ inline void finish_buoyancy(const Blk& B, double& V, double& bx, double& by, double& bz, double& top) {
  double tV = B.warpHyd[0], tMx = B.warpHyd[1], tMy = B.warpHyd[2], tMz = B.warpHyd[3], tTop = B.warpHyd[4];
  for (int w = 1; w < kUnits; ++w) {
    const double* q = B.warpHyd + 5 * w;
    tV += q[0]; tMx += q[1]; tMy += q[2]; tMz += q[3];
    if (q[4] > tTop) tTop = q[4];
  }
  V = tV;
  bx = tV > 0 ? tMx / tV : 0;
  by = tV > 0 ? tMy / tV : 0;
  bz = tV > 0 ? tMz / tV : 0;
  top = tTop;
}

template <bool Mixed>
__device__ void device_step(const DevShip& S, const DevNet& M, const Blk& B, int tid) {
  const int lane = tid & (kLanes - 1), unit = tid / kLanes;
  const int NN = S.nn, NC = M.nc, NG = M.ng;
  Frame F;
  FrameF f;
  pose_and_buoyancy<Mixed>(S, M, B, tid, F, f);

  // (b) free surfaces: logical group `unit` solves groups unit, unit + 8, ...
  {
    const UnitPass<Mixed> pass{S, F, f, lane};
    double* P = B.passes + 7 * M.maxGroup * unit;
    for (int g = unit; g < NG; g += kUnits) solve_group_entry(S, M, F, pass, g, B, P, lane == 0);
  }
  __syncthreads();

  // (c) heads and degrees per node
  for (int n = tid; n < NN; n += kThreads) node_heads(S, M, B, n);
  __syncthreads();

  // (c) flows per connection
  const double t = B.scal[S_T];
  for (int c = tid; c < NC; c += kThreads) B.dV[c] = connection_flow(M, F, t, B, c);
  __syncthreads();

  // (c) gathers, volume updates and load terms per node; monitor sums (events.hpp:73-75, per monitor in
  // increasing connection order); the sea inflow (flows.hpp:288-292) on the last thread
  const double zO = B.scal[S_ZO];
  for (int n = tid; n < NN; n += kThreads) node_update(S, M, F, zO, B, n);
  for (int m = tid; m < M.nmon; m += kThreads) {
    double o = 0;
    const int s = M.monStart[m], e = s + M.monCount[m];
    for (int k = s; k < e; ++k) {
      const double qc = B.q[M.monConn[k]];
      if (qc > 0) o += qc;
    }
    B.over[m] = o;
  }
  if (tid == kThreads - 1) {
    double s = 0;
    for (int k = 0; k < M.seaCount; ++k) s += B.dV[M.sea[k]];
    B.scal[S_SEA] = s;
  }
  __syncthreads();

  // the serial tail (flows.hpp:310-314, step.hpp:147-152, body.hpp, events.hpp) on one lane
  if (tid == 0) {
    double seaIn = B.scal[S_SEA];
    for (int n = 0; n < NN; ++n)
      if (B.excess[n] > 0) seaIn -= B.excess[n];
    B.scal[O_INFLOW] = seaIn / M.dt;
    double hV, hx, hy, hz, htop;
    finish_buoyancy(B, hV, hx, hy, hz, htop);
    // loads (body.hpp:11-40)
    const double W = M.rho * M.g;
    const double Fb = W * hV;
    const double bxw = F.R00 * hx + F.R01 * hy + F.R02 * hz + F.tx;
    const double byw = F.R10 * hx + F.R11 * hy + F.R12 * hz + F.ty;
    double tauX = (byw - M.YO) * Fb, tauY = -(bxw - M.XO) * Fb;
    double Fz = Fb - M.g * M.ms;
    double mw = 0, Iw_p = 0, Iw_r = 0;
    for (int n = 0; n < NN; ++n) {
      if (B.vol[n] <= 0) continue;
      tauX += B.tX[n]; tauY -= B.tY[n]; Fz += B.tF[n];
      mw += B.tM[n];
      Iw_p += B.tIp[n];
      Iw_r += B.tIr[n];
    }
    const double Mth = tauY, Mph = tauX * F.cth;
    // integrate (body.hpp:42-57)
    double vz = B.scal[S_VZ], wth = B.scal[S_WTH], wph = B.scal[S_WPH];
    double zOn = B.scal[S_ZO], pitch = B.scal[S_PITCH], roll = B.scal[S_ROLL];
    const double dt = M.dt;
    const double mh = M.mh0 + mw;
    const double Ith = M.Ith0 + Iw_p, Iph = M.Iph0 + Iw_r;
    const double az = (Fz - M.cDz * vz - M.qz * __builtin_fabs(vz) * vz) / mh;
    const double ath = (Mth - M.cDth * wth - M.qth * __builtin_fabs(wth) * wth) / Ith;
    const double aph = (Mph - M.cDph * wph - M.qph * __builtin_fabs(wph) * wph) / Iph;
    vz += az * dt; wth += ath * dt; wph += aph * dt;
    zOn += vz * dt; pitch += wth * dt; roll += wph * dt;
    if (pitch > kPitchClamp) { pitch = kPitchClamp; wth = 0; }
    if (pitch < -kPitchClamp) { pitch = -kPitchClamp; wth = 0; }
    if (roll > kRollClamp) { roll = kRollClamp; wph = 0; }
    if (roll < -kRollClamp) { roll = -kRollClamp; wph = 0; }
    const double tn = t + dt;
    B.scal[S_T] = tn; B.scal[S_ZO] = zOn; B.scal[S_PITCH] = pitch; B.scal[S_ROLL] = roll;
    B.scal[S_VZ] = vz; B.scal[S_WTH] = wth; B.scal[S_WPH] = wph;
    B.scal[O_VB] = hV; B.scal[O_TOP] = htop; B.scal[O_MW] = mw; B.scal[O_BX] = hx; B.scal[O_BY] = hy; B.scal[O_BZ] = hz;
    // events (events.hpp:11-38), with the pose of the step and the time after it
    for (int m = 0; m < M.nmon; ++m) {
      if (B.over[m] > M.monThreshold && B.overT[m] < 0) {
        B.overT[m] = tn;
        push_event(M, B, tn, m, 0);
      }
    }
    for (int i = 0; i < M.nmark; ++i) {
      if (B.markT[i] >= 0) continue;
      const double wz = world_z(F, M.mkx[i], M.mky[i], M.mkz[i]);
      const bool under = M.mkWhen[i] == 0;
      if ((under && wz < 0) || (!under && wz > 0)) {
        B.markT[i] = tn;
        push_event(M, B, tn, i, 1);
      }
    }
    if (B.scal[S_FOUNDERED] == 0 &&
        (htop < M.hullTopBelow || __builtin_fabs(pitch) > M.pitchAbsAbove || (hV < M.buoyancyBelow && tn > M.minT))) {
      B.scal[S_FOUNDERED] = 1;
      B.scal[S_FOUNDERT] = tn;
      push_event(M, B, tn, -1, 2);
    }
  }
  __syncthreads();
}

}  // namespace ssa
