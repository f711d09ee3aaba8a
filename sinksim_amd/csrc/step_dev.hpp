// The flooding step on gfx950: one 256-thread workgroup (four wave64s) per simulation, the simulation's state,
// scratch and connection overrides in LDS, the shared hull geometry streamed from L2.
//
// Bits: the step reproduces the reference's portable numerics (double) and portable32 numerics (mixed) exactly
// (reference: engine/include/sinksim/kernel/step.hpp:140-153 with Order::Tree / Order::TreeMixed). Three sums
// have a canonical order and are reproduced as such:
//   - buoyancy: 256 logical partials, partial t holding columns t, t+256, ...; each 32-partial group reduced by
//     the halving tree (l absorbs l+16, l+8, ... l+1); the 8 group results combined 0..7 in order
//     (reduce.hpp:12-107, hydro.hpp:242-258). A logical 32-lane group is one half of a wave64.
//   - node pass: 32 lane partials over the node's segments with stride 32, then the same tree (hydro.hpp:306-316).
//   - every other sum (sea inflow, node gathers, loads, monitors) is serial in the reference's index order, one
//     chain per lane where several independent chains exist.
// All other arithmetic is element-wise and written operation for operation as in the reference; the library is
// compiled with -ffp-contract=off so that nothing is fused.
//
// Work layout of a step (| = workgroup barrier):
//   buoyancy partials + the level solve, one logical 32-lane group per free-surface group | per node: effective
//   heads and limiter degrees | per connection: flows | per node: gather, volume update, load terms; per monitor:
//   flow sums; one lane: sea inflow | wave 0: the tail (loads as seven parallel chains, integration, events, and
//   the pose of the next step) |
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fdlibm_dev.hpp"

#define SSG __attribute__((address_space(1)))  // global memory
#define SSC __attribute__((address_space(4)))  // constant memory (scalar loads)
#define SSL __attribute__((address_space(3)))  // LDS

namespace ssa {

constexpr int kThreads = 256;  // logical block of the canonical buoyancy order
constexpr int kLanes = 32;     // logical warp of the canonical trees
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

// four 32-bit integers or four doubles loaded as one vector (the packed per-connection and incidence records)
typedef int32_t i4v __attribute__((ext_vector_type(4)));
typedef double d4v __attribute__((ext_vector_type(4)));

using gcd = const SSG double;
using gcf = const SSG float;
using gci = const SSG int32_t;
using gcb = const SSG uint8_t;

// ------------------------------------------------------------------------------------------- device views

// The hull, shared by every simulation. Segment arrays are pre-gathered per segment (x, y and footprint of the
// segment's column) so that a node pass reads five contiguous streams; the f32 copies are the same values
// rounded once on the host, which is what the portable32 arithmetic converts them to anyway.
struct DevShip {
  int32_t ncol, nseg, nn;
  gcd *x, *y, *dxdy, *zlo, *zhi;
  gcf *xf, *yf, *dxdyf, *zlof, *zhif;
  gci *segStart, *segCount;
  gcd *sx, *sy, *sA, *sa, *se;
  gcf *sxf, *syf, *sAf, *saf, *sef;
  gcd *mu, *vmax, *amin, *aov;
};

// The scenario's immutable network and constants.
struct DevNet {
  double rho, g, sq2g, dt;
  int32_t ng;
  gci *gStart, *gCount, *gNodes;
  gcb* gMode;
  gcd *gpx, *gpy, *gpz;
  double gMargin;
  int32_t nc;
  gci *incStart, *incCount;
  const SSG i4v* inc;   // per incidence entry (node order, increasing connection): connection, its a, its b, side (1 = b)
  const SSG i4v* conI;  // per connection: a, b, skipGroup, law
  const SSG d4v* conD;  // per connection: x, y, z, coef
  int32_t seaCount;
  gci* sea;
  int32_t nmon;
  double monThreshold;
  gci *monStart, *monCount, *monConn;  // per monitor, its connections in increasing order
  int32_t nmark;
  gcd *mkx, *mky, *mkz;
  gcb* mkWhen;
  double hullTopBelow, pitchAbsAbove, buoyancyBelow, minT;
  double ms, g0x, g0y, g0z, XO, YO, mh0, Ith0, Iph0, cDz, cDth, cDph, qz, qth, qph;
  double xFP, xAP;
  int32_t maxGroup, evCap;
};

// Per-instance global arrays of the batch.
struct DevBatch {
  int32_t count;
  gcb* en;
  gcd *area, *tOn;                      // [count * nc]
  SSG double* st;                       // [count * 7]: t zO pitch roll vz wth wph
  SSG double* out;                      // [count * 7]: inflow Vb hullTop mw bx by bz
  SSG double *vol, *level, *cx, *cy, *cz;  // [count * nn]
  SSG double* evT;
  SSG int32_t* evId;
  SSG uint8_t* evKind;
  SSG int32_t* evCount;
  SSG double *overT, *markT;
  SSG uint8_t* foundered;
  SSG double* founderT;
  SSG long long* steps;
  SSG double* final7;                   // [count * 7] readouts at the end of the last launch
  int32_t traced, denseSteps, maxSamples, recSize;
  double sampleEvery;
  SSG double *denseRec, *sampleRec;
  SSG int32_t *denseCount, *sampleCount;
  SSG double* nextSample;
  double curveEvery;
  int32_t curveMax;
  SSG double* curveRec;
  SSG int32_t* curveCount;
  SSG double* nextCurve;
};

// LDS layout of one simulation, as offsets in doubles (the host computes it: lds_layout in batch.hip).
struct LdsLayout {
  int32_t vol, level, cx, cy, cz, area, zmin, zmax, hEff, aEff, excess;
  int32_t tX, tY, tF, tM, tIp, tIr;
  int32_t dV, q, carea, ctOn, passes, warpHyd, frame, scal, over, overT, markT, evT;
  int32_t deg, evId;            // int32 arrays, offsets in doubles
  int32_t merged, cen, evKind;  // byte arrays, offsets in doubles
  int32_t total;                // doubles
};

// Everything a launch needs, in constant memory; device code reads fields on demand with scalar loads.
struct Params {
  DevShip S;
  DevNet M;
  DevBatch Bt;
  LdsLayout L;
};
using PC = const SSC Params;

// frame slots: R00 R01 R02 R10 R11 R12 R20 R21 R22 tx ty tz cth iR, then the f32 frame R20 R21 R22 tz iR
enum : int { F_R00 = 0, F_R01, F_R02, F_R10, F_R11, F_R12, F_R20, F_R21, F_R22, F_TX, F_TY, F_TZ, F_CTH, F_IR, F_F32, F_COUNT = F_F32 + 3 };

// scalar slots of the `scal` block
enum : int { S_T = 0, S_ZO, S_PITCH, S_ROLL, S_VZ, S_WTH, S_WPH, O_INFLOW, O_VB, O_TOP, O_MW, O_BX, O_BY, O_BZ, S_SEA, S_FOUNDERT, S_EVCOUNT,
             S_FOUNDERED, S_NEXTCURVE, S_CURVECOUNT, S_COUNT };

struct Blk {
  SSL double *vol, *level, *cx, *cy, *cz, *area, *zmin, *zmax, *hEff, *aEff, *excess;
  SSL double *tX, *tY, *tF, *tM, *tIp, *tIr;
  SSL double *dV, *q, *carea, *ctOn, *passes, *warpHyd, *frame, *scal, *over, *overT, *markT, *evT;
  SSL int32_t *deg, *evId;
  SSL uint8_t *merged, *cen, *evKind;
};

__device__ inline Blk carve(SSL double* base, PC& P) {
  const SSC LdsLayout& L = P.L;
  Blk b;
  b.vol = base + L.vol; b.level = base + L.level; b.cx = base + L.cx; b.cy = base + L.cy; b.cz = base + L.cz;
  b.area = base + L.area; b.zmin = base + L.zmin; b.zmax = base + L.zmax; b.hEff = base + L.hEff; b.aEff = base + L.aEff;
  b.excess = base + L.excess;
  b.tX = base + L.tX; b.tY = base + L.tY; b.tF = base + L.tF; b.tM = base + L.tM; b.tIp = base + L.tIp; b.tIr = base + L.tIr;
  b.dV = base + L.dV; b.q = base + L.q; b.carea = base + L.carea; b.ctOn = base + L.ctOn; b.passes = base + L.passes;
  b.warpHyd = base + L.warpHyd; b.frame = base + L.frame; b.scal = base + L.scal; b.over = base + L.over; b.overT = base + L.overT;
  b.markT = base + L.markT; b.evT = base + L.evT;
  b.deg = reinterpret_cast<SSL int32_t*>(base + L.deg);
  b.evId = reinterpret_cast<SSL int32_t*>(base + L.evId);
  b.merged = reinterpret_cast<SSL uint8_t*>(base + L.merged);
  b.cen = reinterpret_cast<SSL uint8_t*>(base + L.cen);
  b.evKind = reinterpret_cast<SSL uint8_t*>(base + L.evKind);
  return b;
}

// ------------------------------------------------------------------------------------------- pose

// Out of line: called four lanes at a time once per step, and for the 30 s readouts.
__device__ __noinline__ double trig(double x, int wantCos) { return fdl::sin_or_cos(x, wantCos); }
__device__ __noinline__ double dpow(double x, double y) { return fdl::pow(x, y); }

// reference: kernel/frame.hpp:10-21, written to the frame slots, with iR = 1 / R22 (hydro.hpp:209,245,264,309)
// and the f32 frame of reduce.hpp:118-126
__device__ inline void write_frame(PC& P, SSL double* fr, double zO, double cth, double sth, double cph, double sph) {
  const double R00 = cth, R01 = sth * sph, R02 = sth * cph;
  const double R10 = 0, R11 = cph, R12 = -sph;
  const double R20 = -sth, R21 = cth * sph, R22 = cth * cph;
  const double tx = P.M.XO - (R00 * P.M.g0x + R01 * P.M.g0y + R02 * P.M.g0z);
  const double ty = P.M.YO - (R10 * P.M.g0x + R11 * P.M.g0y + R12 * P.M.g0z);
  const double tz = zO - (R20 * P.M.g0x + R21 * P.M.g0y + R22 * P.M.g0z);
  fr[F_R00] = R00; fr[F_R01] = R01; fr[F_R02] = R02; fr[F_R10] = R10; fr[F_R11] = R11; fr[F_R12] = R12;
  fr[F_R20] = R20; fr[F_R21] = R21; fr[F_R22] = R22; fr[F_TX] = tx; fr[F_TY] = ty; fr[F_TZ] = tz; fr[F_CTH] = cth;
  fr[F_IR] = 1 / R22;
  const float f20 = static_cast<float>(R20), f21 = static_cast<float>(R21), f22 = static_cast<float>(R22), ftz = static_cast<float>(tz);
  SSL float* ff = reinterpret_cast<SSL float*>(fr + F_F32);
  ff[0] = f20; ff[1] = f21; ff[2] = f22; ff[3] = ftz; ff[4] = 1.0f / f22;
}

// The pose (zO, pitch, roll: the same value in every lane) into the frame slots: lanes 0..3 of wave 0 evaluate
// cos/sin of pitch and roll, lane 0 writes the frame. Call from wave 0 only; visible to the workgroup after the
// next barrier.
__device__ inline void pose_wave0(PC& P, const Blk& B, int lane, double zO, double pitch, double roll) {
  double v = 0;
  if (lane < 4) v = trig(lane < 2 ? pitch : roll, (lane & 1) == 0 ? 1 : 0);
  const double cth = __builtin_bit_cast(double, ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(__builtin_bit_cast(uint64_t, v) >> 32), 0) << 32) |
                                                    (uint32_t)__builtin_amdgcn_readlane((int)__builtin_bit_cast(uint64_t, v), 0));
  const double sth = __builtin_bit_cast(double, ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(__builtin_bit_cast(uint64_t, v) >> 32), 1) << 32) |
                                                    (uint32_t)__builtin_amdgcn_readlane((int)__builtin_bit_cast(uint64_t, v), 1));
  const double cph = __builtin_bit_cast(double, ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(__builtin_bit_cast(uint64_t, v) >> 32), 2) << 32) |
                                                    (uint32_t)__builtin_amdgcn_readlane((int)__builtin_bit_cast(uint64_t, v), 2));
  const double sph = __builtin_bit_cast(double, ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(__builtin_bit_cast(uint64_t, v) >> 32), 3) << 32) |
                                                    (uint32_t)__builtin_amdgcn_readlane((int)__builtin_bit_cast(uint64_t, v), 3));
  if (lane == 0) write_frame(P, B.frame, zO, cth, sth, cph, sph);
}

__device__ __forceinline__ double world_z(const SSL double* fr, double x, double y, double z) {
  return fr[F_R20] * x + fr[F_R21] * y + fr[F_R22] * z + fr[F_TZ];
}

// ------------------------------------------------------------------------------------------- cross-lane

// Lane l of the logical 32-lane group receives lane l + off (off = 8, 4, 2, 1 by DPP row shifts inside a row of
// 16; off = 16 by an LDS-unit swizzle that swaps the two rows of each group). Only lanes l < off use the result.
template <int OFF>
__device__ __forceinline__ uint32_t down32(uint32_t v) {
  if constexpr (OFF == 16) return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(static_cast<int>(v), 0x401F));
  else return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(v), static_cast<int>(v), 0x100 + OFF, 0xF, 0xF, false));
}
template <int OFF>
__device__ __forceinline__ double down(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = down32<OFF>(static_cast<uint32_t>(b)), hi = down32<OFF>(static_cast<uint32_t>(b >> 32));
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}
template <int OFF>
__device__ __forceinline__ float down(float v) {
  return __builtin_bit_cast(float, down32<OFF>(__builtin_bit_cast(uint32_t, v)));
}

// lane 0 of the logical group (wave lane 0 or 32) to every lane of the group
__device__ __forceinline__ uint32_t lead32(uint32_t v, bool upper) {
  const uint32_t a = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 0));
  const uint32_t b = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 32));
  return upper ? b : a;
}
__device__ __forceinline__ double lead(double v, bool upper) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  return __builtin_bit_cast(double, (static_cast<uint64_t>(lead32(static_cast<uint32_t>(b >> 32), upper)) << 32) |
                                        lead32(static_cast<uint32_t>(b), upper));
}
__device__ __forceinline__ float lead(float v, bool upper) { return __builtin_bit_cast(float, lead32(__builtin_bit_cast(uint32_t, v), upper)); }

// ------------------------------------------------------------------------------------------- partials

template <class T>
struct HP {
  T V, Mx, My, Mz, top;
};
template <class T>
struct PP {
  T vol, dv, mx, my, mz, zmin, zmax;
};

template <int OFF, class T>
__device__ __forceinline__ void tree_level(HP<T>& p, int lane) {
  const T V = down<OFF>(p.V), Mx = down<OFF>(p.Mx), My = down<OFF>(p.My), Mz = down<OFF>(p.Mz), top = down<OFF>(p.top);
  if (lane < OFF) {
    p.V += V; p.Mx += Mx; p.My += My; p.Mz += Mz;
    if (top > p.top) p.top = top;
  }
}
template <int OFF, class T>
__device__ __forceinline__ void tree_level(PP<T>& p, int lane) {
  const T vol = down<OFF>(p.vol), dv = down<OFF>(p.dv), mx = down<OFF>(p.mx), my = down<OFF>(p.my), mz = down<OFF>(p.mz);
  const T zmin = down<OFF>(p.zmin), zmax = down<OFF>(p.zmax);
  if (lane < OFF) {
    p.vol += vol; p.dv += dv; p.mx += mx; p.my += my; p.mz += mz;
    if (zmin < p.zmin) p.zmin = zmin;
    if (zmax > p.zmax) p.zmax = zmax;
  }
}

// The halving tree of a logical 32-lane group (reduce.hpp:102-107): l absorbs l + 16, 8, 4, 2, 1; the total is in
// lane 0 of the group.
template <class P_>
__device__ __forceinline__ void tree32(P_& p, int lane) {
  tree_level<16>(p, lane);
  tree_level<8>(p, lane);
  tree_level<4>(p, lane);
  tree_level<2>(p, lane);
  tree_level<1>(p, lane);
}

// one column into a buoyancy partial (reduce.hpp:67-77 / 159-170)
struct Fr64 {
  double R20, R21, R22, tz, iR;
};
struct Fr32 {
  float R20, R21, R22, tz, iR;
};

__device__ __forceinline__ Fr64 fr64(const SSL double* fr) { return Fr64{fr[F_R20], fr[F_R21], fr[F_R22], fr[F_TZ], fr[F_IR]}; }
__device__ __forceinline__ Fr32 fr32(const SSL double* fr) {
  const SSL float* f = reinterpret_cast<const SSL float*>(fr + F_F32);
  return Fr32{f[0], f[1], f[2], f[3], f[4]};
}

// A column's or a segment's geometry, fetched ahead of its evaluation so that a lane's loads of several
// columns or segments are in flight together.
template <class T>
struct Geo {
  T x, y, lo, hi, A;
};
__device__ __forceinline__ Geo<double> column_geo(PC& P, const double*, int i) { return Geo<double>{P.S.x[i], P.S.y[i], P.S.zlo[i], P.S.zhi[i], P.S.dxdy[i]}; }
__device__ __forceinline__ Geo<float> column_geo(PC& P, const float*, int i) { return Geo<float>{P.S.xf[i], P.S.yf[i], P.S.zlof[i], P.S.zhif[i], P.S.dxdyf[i]}; }
__device__ __forceinline__ Geo<double> segment_geo(PC& P, const double*, int j) { return Geo<double>{P.S.sx[j], P.S.sy[j], P.S.sa[j], P.S.se[j], P.S.sA[j]}; }
__device__ __forceinline__ Geo<float> segment_geo(PC& P, const float*, int j) { return Geo<float>{P.S.sxf[j], P.S.syf[j], P.S.saf[j], P.S.sef[j], P.S.sAf[j]}; }

constexpr int kFetch = 4;  // columns or segments a lane fetches before evaluating them in order

__device__ __forceinline__ void column(const Fr64& F, const Geo<double>& g, HP<double>& p) {
  const double x = g.x, y = g.y, lo = g.lo, hi = g.hi;
  const double base = F.R20 * x + F.R21 * y + F.tz;
  const double wt = base + F.R22 * hi;
  if (wt > p.top) p.top = wt;
  const double zp = -base * F.iR;
  if (zp <= lo) return;
  const double s = zp >= hi ? hi - lo : zp - lo;
  const double v = g.A * s;
  p.V += v; p.Mx += v * x; p.My += v * y; p.Mz += v * (lo + 0.5 * s);
}
__device__ __forceinline__ void column(const Fr32& f, const Geo<float>& g, HP<float>& p) {
  const float x = g.x, y = g.y, lo = g.lo, hi = g.hi;
  const float base = f.R20 * x + f.R21 * y + f.tz;
  const float wt = base + f.R22 * hi;
  if (wt > p.top) p.top = wt;
  const float zp = -base * f.iR;
  if (zp <= lo) return;
  const float s = zp >= hi ? hi - lo : zp - lo;
  const float v = g.A * s;
  p.V += v; p.Mx += v * x; p.My += v * y; p.Mz += v * (lo + 0.5f * s);
}

// one segment into a node-pass partial (reduce.hpp:80-99 / 172-192)
__device__ __forceinline__ void segment(const Fr64& F, double h, const Geo<double>& g, PP<double>& p) {
  const double x = g.x, y = g.y, a = g.lo, b = g.hi;
  const double base = F.R20 * x + F.R21 * y + F.tz;
  const double wa = base + F.R22 * a, wb = base + F.R22 * b;
  if (wa < p.zmin) p.zmin = wa;
  if (wb > p.zmax) p.zmax = wb;
  const double zp = (h - base) * F.iR;
  if (zp <= a) return;
  const double A = g.A;
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
__device__ __forceinline__ void segment(const Fr32& f, float h, const Geo<float>& g, PP<float>& p) {
  const float x = g.x, y = g.y, a = g.lo, b = g.hi;
  const float base = f.R20 * x + f.R21 * y + f.tz;
  const float wa = base + f.R22 * a, wb = base + f.R22 * b;
  if (wa < p.zmin) p.zmin = wa;
  if (wb > p.zmax) p.zmax = wb;
  const float zp = (h - base) * f.iR;
  if (zp <= a) return;
  const float A = g.A;
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
  PC& P;
  const SSL double* fr;
  int lane;
  bool upper;
  __device__ __forceinline__ Pass operator()(int n, double h) const {
    const int s0 = P.S.segStart[n], s1 = s0 + P.S.segCount[n];
    double vol, dv, mx, my, mz, zmin, zmax;
    if constexpr (Mixed) {
      const Fr32 f = fr32(fr);
      const float hf = static_cast<float>(h);
      PP<float> p{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1e9f, -1e9f};
      for (int j0 = s0 + lane; j0 < s1; j0 += kFetch * kLanes) {
        Geo<float> g[kFetch];
#pragma unroll
        for (int u = 0; u < kFetch; ++u) g[u] = segment_geo(P, static_cast<const float*>(nullptr), j0 + u * kLanes < s1 ? j0 + u * kLanes : j0);
#pragma unroll
        for (int u = 0; u < kFetch; ++u)
          if (j0 + u * kLanes < s1) segment(f, hf, g[u], p);
      }
      tree32(p, lane);
      vol = static_cast<double>(lead(p.vol, upper)); dv = static_cast<double>(lead(p.dv, upper));
      mx = static_cast<double>(lead(p.mx, upper)); my = static_cast<double>(lead(p.my, upper)); mz = static_cast<double>(lead(p.mz, upper));
      zmin = static_cast<double>(lead(p.zmin, upper)); zmax = static_cast<double>(lead(p.zmax, upper));
    } else {
      const Fr64 F = fr64(fr);
      PP<double> p{0.0, 0.0, 0.0, 0.0, 0.0, 1e9, -1e9};
      for (int j0 = s0 + lane; j0 < s1; j0 += kFetch * kLanes) {
        Geo<double> g[kFetch];
#pragma unroll
        for (int u = 0; u < kFetch; ++u) g[u] = segment_geo(P, static_cast<const double*>(nullptr), j0 + u * kLanes < s1 ? j0 + u * kLanes : j0);
#pragma unroll
        for (int u = 0; u < kFetch; ++u)
          if (j0 + u * kLanes < s1) segment(F, h, g[u], p);
      }
      tree32(p, lane);
      vol = lead(p.vol, upper); dv = lead(p.dv, upper); mx = lead(p.mx, upper); my = lead(p.my, upper); mz = lead(p.mz, upper);
      zmin = lead(p.zmin, upper); zmax = lead(p.zmax, upper);
    }
    const double mu = P.S.mu[n];
    const double iR = fr[F_IR];
    Pass r;
    r.vol = vol * mu;
    r.dv = dv * mu * iR;
    r.mx = mx * mu; r.my = my * mu; r.mz = mz * mu;
    r.zmin = zmin; r.zmax = zmax;
    return r;
  }
};

// ------------------------------------------------------------------------------------------- (b) levels

__device__ __forceinline__ void put_centroid(const Blk& B, int n, const Pass& p) {
  if (p.vol > kCentroidMinVolume) { B.cx[n] = p.mx / p.vol; B.cy[n] = p.my / p.vol; B.cz[n] = p.mz / p.vol; }
}

__device__ __forceinline__ void put_pass(SSL double* Pp, int k, const Pass& p) {
  SSL double* q = Pp + 7 * k;
  q[0] = p.vol; q[1] = p.dv; q[2] = p.mx; q[3] = p.my; q[4] = p.mz; q[5] = p.zmin; q[6] = p.zmax;
}
__device__ __forceinline__ Pass get_pass(const SSL double* Pp, int k) {
  const SSL double* q = Pp + 7 * k;
  return Pass{q[0], q[1], q[2], q[3], q[4], q[5], q[6]};
}

// levels.hpp:13-49 with full = false (the step never asks for the 40-iteration solve)
template <class PassFn>
__device__ void solve_single(PC& P, const PassFn& pass, int n, const Blk& B, bool lead_) {
  const double V = B.vol[n], Vmax = P.S.vmax[n];
  double h = B.level[n];
  Pass p = pass(n, h);
  const double lo = p.zmin, hi = p.zmax;
  if (lead_) { B.zmin[n] = lo; B.zmax[n] = hi; }
  if (V <= kEmptyVolume) {
    if (lead_) { B.level[n] = lo; B.area[n] = 0; }
    return;
  }
  if (V >= Vmax * kFullFraction) {
    p = pass(n, hi);
    if (lead_) {
      B.level[n] = hi + jmax(0.0, V - Vmax) / P.S.aov[n];
      B.area[n] = P.S.aov[n];
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
  if (lead_) {
    B.level[n] = h; B.area[n] = p.dv;
    put_centroid(B, n, p);
  }
}

// levels.hpp:54-162 with full = false. Pp is this unit's pass scratch in LDS (maxGroup entries).
template <class PassFn>
__device__ void solve_group(PC& P, const PassFn& pass, gci* members, int N, const Blk& B, SSL double* Pp, bool lead_) {
  double V = 0, Vmax = 0;
  for (int k = 0; k < N; ++k) { V += B.vol[members[k]]; Vmax += P.S.vmax[members[k]]; }
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
  // Pp is written by every lane of the unit (each lane holds the same pass result) and read back by each lane
  // after its own write, so no lane depends on another lane's LDS store between barriers.
  for (int k = 0; k < N; ++k) {
    const Pass p = pass(members[k], h);
    put_pass(Pp, k, p);
    if (lead_) { B.zmin[members[k]] = p.zmin; B.zmax[members[k]] = p.zmax; }
    if (k == 0) { lo = p.zmin; hi = p.zmax; }
    else { if (p.zmin < lo) lo = p.zmin; if (p.zmax > hi) hi = p.zmax; }
  }
  if (V <= kEmptyVolume) {
    if (lead_) {
      for (int k = 0; k < N; ++k) { const int n = members[k]; B.vol[n] = 0; B.level[n] = Pp[7 * k + 5]; B.area[n] = 0; }
    }
    return;
  }
  if (V >= Vmax * kFullFraction) {
    const double over = jmax(0.0, V - Vmax);
    double aSum = 0;
    for (int k = 0; k < N; ++k) aSum += P.S.aov[members[k]];
    const double dh = over / aSum;
    double assigned = 0;
    for (int k = 0; k < N - 1; ++k) {
      const double v = P.S.vmax[members[k]] + dh * P.S.aov[members[k]];
      if (lead_) B.vol[members[k]] = v;
      assigned += v;
    }
    if (lead_) B.vol[members[N - 1]] = V - assigned;
    for (int k = 0; k < N; ++k) {
      const int n = members[k];
      const Pass p = pass(n, hi);
      if (lead_) {
        B.level[n] = hi + dh; B.area[n] = P.S.aov[n];
        put_centroid(B, n, p);
      }
    }
    return;
  }
  if (!(h >= lo && h <= hi)) {
    h = jmin(hi, jmax(lo, h));
    for (int k = 0; k < N; ++k) {
      const Pass p = pass(members[k], h);
      put_pass(Pp, k, p);
    }
  }
  const double tol = jmax(1e-3 * Vmax, 0.5);
  double a = lo, b = hi;
  for (int it = 0; it < kNewtonFast; ++it) {
    double sum = 0;
    for (int k = 0; k < N; ++k) sum += Pp[7 * k + 0];
    const double err = sum - V;
    if (__builtin_fabs(err) < tol) break;
    if (err > 0) b = h; else a = h;
    double dv = 0;
    for (int k = 0; k < N; ++k) dv += Pp[7 * k + 1];
    double hn = dv > kSlopeMin ? h - err / dv : 0.5 * (a + b);
    if (!(hn > a && hn < b)) hn = 0.5 * (a + b);
    h = hn;
    for (int k = 0; k < N; ++k) {
      const Pass p = pass(members[k], h);
      put_pass(Pp, k, p);
    }
  }
  // share the water in proportion to the geometry at this level (levels.hpp:125-161)
  if (!lead_) return;
  double tot = 0;
  for (int k = 0; k < N; ++k) tot += Pp[7 * k + 0];
  double assigned = 0;
  for (int k = 0; k < N - 1; ++k) {
    double v = tot > 0 ? V * Pp[7 * k + 0] / tot : V / static_cast<double>(N);
    if (v > P.S.vmax[members[k]]) v = P.S.vmax[members[k]];
    B.vol[members[k]] = v;
    assigned += v;
  }
  const double vmaxLast = P.S.vmax[members[N - 1]];
  if (V - assigned > vmaxLast) {
    const double need = V - vmaxLast;
    double capped = 0;
    for (int k = N - 2; k >= 0; --k) {
      double prior = 0;
      for (int j = 0; j < k; ++j) prior += B.vol[members[j]];
      const double v = (need - capped) - prior;
      if (v > P.S.vmax[members[k]] && k > 0) {
        B.vol[members[k]] = P.S.vmax[members[k]];
        capped += P.S.vmax[members[k]];
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
    const Pass p = get_pass(Pp, k);
    B.level[n] = h; B.area[n] = p.dv;
    put_centroid(B, n, p);
  }
}

// levels.hpp:166-188: the merge decision, then the solve, for group g
template <class PassFn>
__device__ void solve_group_entry(PC& P, const PassFn& pass, int g, const Blk& B, SSL double* Pp, bool lead_) {
  const int N = P.M.gCount[g];
  gci* members = P.M.gNodes + P.M.gStart[g];
  const uint8_t mode = P.M.gMode[g];
  bool merged = false;
  if (N >= 2) {
    if (mode == 1) {
      merged = true;
    } else if (mode == 2) {
      const double zMw = world_z(B.frame, P.M.gpx[g], P.M.gpy[g], P.M.gpz[g]);
      merged = true;
      for (int k = 0; k < N; ++k) {
        const int n = members[k];
        if (!(B.vol[n] > kEmptyVolume && B.level[n] > zMw + P.M.gMargin)) { merged = false; break; }
      }
    }
  }
  if (lead_) B.merged[g] = merged ? 1 : 0;
  if (merged) solve_group(P, pass, members, N, B, Pp, lead_);
  else for (int k = 0; k < N; ++k) solve_single(P, pass, members[k], B, lead_);
}

// ------------------------------------------------------------------------------------------- (c) network

__device__ __forceinline__ double head_of(const Blk& B, int n) { return B.vol[n] > kEmptyVolume ? B.level[n] : kNoHead; }

// flows.hpp:211-231: effective head and area of node n, and the number of connections that count toward the
// limiter there (an enabled connection not between two empty spaces). The heads of the other ends are
// recomputed from vol and level, so that one pass over the nodes suffices.
__device__ __forceinline__ void node_heads(PC& P, const Blk& B, int n) {
  const double v = B.vol[n];
  B.hEff[n] = v > kEmptyVolume ? B.level[n] : kNoHead;
  B.aEff[n] = v >= P.S.vmax[n] ? P.S.aov[n] : jmax(B.area[n], P.S.amin[n]);
  const int s = P.M.incStart[n], e = s + P.M.incCount[n];
  int d = 0;
  for (int k0 = s; k0 < e; k0 += kFetch) {
    i4v q[kFetch];
#pragma unroll
    for (int u = 0; u < kFetch; ++u) q[u] = P.M.inc[k0 + u < e ? k0 + u : k0];
#pragma unroll
    for (int u = 0; u < kFetch; ++u) {
      if (k0 + u >= e || !B.cen[q[u].x]) continue;
      const int a = q[u].y, b = q[u].z;
      if (a >= 0 && head_of(B, a) < kNoHeadTest && head_of(B, b) < kNoHeadTest) continue;
      ++d;
    }
  }
  B.deg[n] = d;
}

// flows.hpp:234-274: the volume moved through connection c this step (a -> b positive), after the limiter
__device__ inline double connection_flow(PC& P, const SSL double* fr, double t, const Blk& B, int c) {
  const i4v ci = P.M.conI[c];
  const d4v cd = P.M.conD[c];
  if (!B.cen[c] || B.ctOn[c] > t) { B.q[c] = 0; return 0; }
  if (ci.z >= 0 && B.merged[ci.z]) { B.q[c] = 0; return 0; }
  const int a = ci.x, b = ci.y;
  const double zo = fr[F_R20] * cd.x + fr[F_R21] * cd.y + fr[F_R22] * cd.z + fr[F_TZ];
  const double ha = a < 0 ? 0.0 : B.hEff[a], hbv = B.hEff[b];
  const double Ha = ha > zo ? ha - zo : 0.0, Hb = hbv > zo ? hbv - zo : 0.0;
  const double sq2g = P.M.sq2g;
  double q = 0;
  if (ci.w == 0) {
    const double d = Ha - Hb;
    if (d != 0) q = cd.w * B.carea[c] * sq2g * d / __builtin_sqrt(jmax(__builtin_fabs(d), kOrificeHeadFloor));
  } else {
    const double H1 = Ha > Hb ? Ha : Hb;
    if (H1 > 0) {
      const double H2 = Ha > Hb ? Hb : Ha;
      double f = 1;
      if (H2 > 0) {
        const double r = H2 / H1;
        f = dpow(jmax(0.0, 1 - dpow(r, kWeirExponent)), kVillemonteExponent);
      }
      q = cd.w * B.carea[c] * kWeirCoefficient * sq2g * H1 * __builtin_sqrt(H1) * f * (Ha >= Hb ? 1.0 : -1.0);
    }
  }
  if (q == 0) { B.q[c] = 0; return 0; }
  const double dt = P.M.dt;
  double dV = q * dt;
  const int src = dV > 0 ? a : b, dst = dV > 0 ? b : a;
  const double hs = src < 0 ? 0.0 : B.hEff[src];
  const double hd = dst < 0 ? 0.0 : B.hEff[dst];
  const double inv = (src < 0 ? 0.0 : 1 / B.aEff[src]) + (dst < 0 ? 0.0 : 1 / B.aEff[dst]);
  const int degRef = src < 0 ? B.deg[dst] : (dst < 0 ? B.deg[src] : (B.deg[src] > B.deg[dst] ? B.deg[src] : B.deg[dst]));
  const double share = kLimiterShare / static_cast<double>(degRef > 1 ? degRef : 1);
  double lim = inv > 0 ? share * (hs - jmax(hd, zo)) / inv : __builtin_fabs(dV);
  if (lim < 0) lim = 0;
  if (__builtin_fabs(dV) > lim) dV = dV > 0 ? lim : -lim;
  B.q[c] = dV / dt;
  return dV;
}

// flows.hpp:277-307 for node n (the gather in increasing connection order and the capped volume update), then
// the node's terms of the loads (body.hpp:21-34), which depend on this node only once its volume is final.
__device__ inline void node_update(PC& P, const Blk& B, int n) {
  const int s = P.M.incStart[n], e = s + P.M.incCount[n];
  double acc = 0;
  for (int k0 = s; k0 < e; k0 += kFetch) {
    i4v q[kFetch];
    double d[kFetch];
#pragma unroll
    for (int u = 0; u < kFetch; ++u) q[u] = P.M.inc[k0 + u < e ? k0 + u : k0];
#pragma unroll
    for (int u = 0; u < kFetch; ++u) d[u] = B.dV[q[u].x];
#pragma unroll
    for (int u = 0; u < kFetch; ++u)
      if (k0 + u < e) acc = q[u].w ? acc + d[u] : acc - d[u];
  }
  double v = B.vol[n] + acc;
  if (v < 0) v = 0;
  const double vm = P.S.vmax[n];
  double excess = 0;
  if (v > vm) {
    const double head = jmax(0.0, -B.zmax[n]);
    const double cap = vm + P.S.aov[n] * head;
    if (v > cap) { excess = v - cap; v = cap; }
  }
  B.excess[n] = excess;
  B.vol[n] = v;
  if (v > 0) {
    const SSL double* fr = B.frame;
    const double W = P.M.rho * P.M.g;
    const double x = B.cx[n], y = B.cy[n], z = B.cz[n];
    const double wx = fr[F_R00] * x + fr[F_R01] * y + fr[F_R02] * z + fr[F_TX] - P.M.XO;
    const double wy = fr[F_R10] * x + fr[F_R11] * y + fr[F_R12] * z + fr[F_TY] - P.M.YO;
    const double wz = fr[F_R20] * x + fr[F_R21] * y + fr[F_R22] * z + fr[F_TZ] - B.scal[S_ZO];
    const double f = -W * v;
    const double m = P.M.rho * v;
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
__device__ __noinline__ void readouts7(PC& P, const Blk& B, SSG double* r) {
  const double t = B.scal[S_T], zO = B.scal[S_ZO], pitch = B.scal[S_PITCH], roll = B.scal[S_ROLL];
  const double cth = trig(pitch, 1), sth = trig(pitch, 0), cph = trig(roll, 1), sph = trig(roll, 0);
  const double R20 = -sth, R21 = cth * sph, R22 = cth * cph;
  const double tz = zO - (R20 * P.M.g0x + R21 * P.M.g0y + R22 * P.M.g0z);
  double water = 0;
  for (int n = 0; n < P.S.nn; ++n) water += B.vol[n];
  r[0] = t;
  r[1] = pitch * 180 / kPi;
  r[2] = roll * 180 / kPi;
  r[3] = water * P.M.rho / 1000;
  r[4] = B.scal[O_INFLOW] * P.M.rho / 1000 * 60;
  r[5] = -(R20 * P.M.xFP + R21 * 0.0 + R22 * 0.0 + tz);
  r[6] = -(R20 * P.M.xAP + R21 * 0.0 + R22 * 0.0 + tz);
}

__device__ inline void push_event(PC& P, const Blk& B, double t, int id, int kind) {
  const int k = static_cast<int>(B.scal[S_EVCOUNT]);
  if (k < P.M.evCap) {
    B.evT[k] = t; B.evId[k] = id; B.evKind[k] = static_cast<uint8_t>(kind);
    B.scal[S_EVCOUNT] = k + 1;
  }
}

// ------------------------------------------------------------------------------------------- the step

// Buoyancy partials of this thread's 32-lane group; the lead lane stores the group's tree result.
template <bool Mixed>
__device__ __forceinline__ void buoyancy_partials(PC& P, const Blk& B, int tid) {
  const int lane = tid & (kLanes - 1), unit = tid / kLanes;
  SSL double* w = B.warpHyd + 5 * unit;
  if constexpr (Mixed) {
    const Fr32 f = fr32(B.frame);
    HP<float> p{0.0f, 0.0f, 0.0f, 0.0f, -1e9f};
    for (int i0 = tid; i0 < P.S.ncol; i0 += kFetch * kThreads) {
      Geo<float> g[kFetch];
#pragma unroll
      for (int u = 0; u < kFetch; ++u) g[u] = column_geo(P, static_cast<const float*>(nullptr), i0 + u * kThreads < P.S.ncol ? i0 + u * kThreads : i0);
#pragma unroll
      for (int u = 0; u < kFetch; ++u)
        if (i0 + u * kThreads < P.S.ncol) column(f, g[u], p);
    }
    tree32(p, lane);
    if (lane == 0) { w[0] = p.V; w[1] = p.Mx; w[2] = p.My; w[3] = p.Mz; w[4] = p.top; }
  } else {
    const Fr64 F = fr64(B.frame);
    HP<double> p{0.0, 0.0, 0.0, 0.0, -1e9};
    for (int i0 = tid; i0 < P.S.ncol; i0 += kFetch * kThreads) {
      Geo<double> g[kFetch];
#pragma unroll
      for (int u = 0; u < kFetch; ++u) g[u] = column_geo(P, static_cast<const double*>(nullptr), i0 + u * kThreads < P.S.ncol ? i0 + u * kThreads : i0);
#pragma unroll
      for (int u = 0; u < kFetch; ++u)
        if (i0 + u * kThreads < P.S.ncol) column(F, g[u], p);
    }
    tree32(p, lane);
    if (lane == 0) { w[0] = p.V; w[1] = p.Mx; w[2] = p.My; w[3] = p.Mz; w[4] = p.top; }
  }
}

// the 8 group results combined in order, then finish_hydro (hydro.hpp:231-239, 255-257)
__device__ inline void finish_buoyancy(const Blk& B, double& V, double& bx, double& by, double& bz, double& top) {
  double tV = B.warpHyd[0], tMx = B.warpHyd[1], tMy = B.warpHyd[2], tMz = B.warpHyd[3], tTop = B.warpHyd[4];
  for (int w = 1; w < kUnits; ++w) {
    const SSL double* q = B.warpHyd + 5 * w;
    tV += q[0]; tMx += q[1]; tMy += q[2]; tMz += q[3];
    if (q[4] > tTop) tTop = q[4];
  }
  V = tV;
  bx = tV > 0 ? tMx / tV : 0;
  by = tV > 0 ? tMy / tV : 0;
  bz = tV > 0 ? tMz / tV : 0;
  top = tTop;
}

__device__ __forceinline__ double bcast(double v, int from) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(b)), from));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(b >> 32)), from));
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// The tail of a step on wave 0 (flows.hpp:310-314, step.hpp:147-152, body.hpp, events.hpp). Lanes 0..6 run the
// seven serial node-order chains (tauX, tauY, Fz, mw, Iw_p, Iw_r, and the sea inflow less the excess), eight
// nodes' values fetched at a time; every lane then integrates the same uniform values, the lanes test the monitors
// and marks side by side and lane 0 records the crossings in order; lanes 0..3 evaluate the pose of the next step.
// Values flow between lanes through registers only: LDS written here is read after the step's closing barrier.
__device__ void step_tail(PC& P, const Blk& B, int lane) {
  const SSL double* fr = B.frame;
  double hV, hx, hy, hz, htop;
  finish_buoyancy(B, hV, hx, hy, hz, htop);
  const double W = P.M.rho * P.M.g;
  const double Fb = W * hV;
  const double bxw = fr[F_R00] * hx + fr[F_R01] * hy + fr[F_R02] * hz + fr[F_TX];
  const double byw = fr[F_R10] * hx + fr[F_R11] * hy + fr[F_R12] * hz + fr[F_TY];
  const double t0 = B.scal[S_T];
  double vz = B.scal[S_VZ], wth = B.scal[S_WTH], wph = B.scal[S_WPH];
  double zO = B.scal[S_ZO], pitch = B.scal[S_PITCH], roll = B.scal[S_ROLL];
  const bool wasFoundered = B.scal[S_FOUNDERED] != 0;
  const int NN = P.S.nn;
  double acc = 0;
  if (lane < 7) {
    const SSL double* src = lane == 0 ? B.tX : lane == 1 ? B.tY : lane == 2 ? B.tF : lane == 3 ? B.tM : lane == 4 ? B.tIp : lane == 5 ? B.tIr : B.excess;
    const SSL double* gate = lane < 6 ? B.vol : B.excess;
    const bool subtract = lane == 1 || lane == 6;
    const bool loads = lane < 6;
    acc = lane == 0 ? (byw - P.M.YO) * Fb : lane == 1 ? -(bxw - P.M.XO) * Fb : lane == 2 ? Fb - P.M.g * P.M.ms : lane == 6 ? B.scal[S_SEA] : 0.0;
    for (int n0 = 0; n0 < NN; n0 += 8) {
      double gv[8], sv[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int n = n0 + k < NN ? n0 + k : NN - 1;
        gv[k] = gate[n];
        sv[k] = src[n];
      }
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (n0 + k < NN && (loads ? !(gv[k] <= 0) : gv[k] > 0)) acc = subtract ? acc - sv[k] : acc + sv[k];
      }
    }
  }
  const double tauX = bcast(acc, 0), tauY = bcast(acc, 1), Fz = bcast(acc, 2), mw = bcast(acc, 3);
  const double Iw_p = bcast(acc, 4), Iw_r = bcast(acc, 5), seaIn = bcast(acc, 6);
  // integrate (body.hpp:42-57)
  const double Mth = tauY, Mph = tauX * fr[F_CTH];
  const double dt = P.M.dt;
  const double mh = P.M.mh0 + mw;
  const double Ith = P.M.Ith0 + Iw_p, Iph = P.M.Iph0 + Iw_r;
  const double az = (Fz - P.M.cDz * vz - P.M.qz * __builtin_fabs(vz) * vz) / mh;
  const double ath = (Mth - P.M.cDth * wth - P.M.qth * __builtin_fabs(wth) * wth) / Ith;
  const double aph = (Mph - P.M.cDph * wph - P.M.qph * __builtin_fabs(wph) * wph) / Iph;
  vz += az * dt; wth += ath * dt; wph += aph * dt;
  zO += vz * dt; pitch += wth * dt; roll += wph * dt;
  if (pitch > kPitchClamp) { pitch = kPitchClamp; wth = 0; }
  if (pitch < -kPitchClamp) { pitch = -kPitchClamp; wth = 0; }
  if (roll > kRollClamp) { roll = kRollClamp; wph = 0; }
  if (roll < -kRollClamp) { roll = -kRollClamp; wph = 0; }
  const double tn = t0 + dt;
  if (lane == 0) {
    B.scal[O_INFLOW] = seaIn / dt;
    B.scal[S_T] = tn; B.scal[S_ZO] = zO; B.scal[S_PITCH] = pitch; B.scal[S_ROLL] = roll;
    B.scal[S_VZ] = vz; B.scal[S_WTH] = wth; B.scal[S_WPH] = wph;
    B.scal[O_VB] = hV; B.scal[O_TOP] = htop; B.scal[O_MW] = mw; B.scal[O_BX] = hx; B.scal[O_BY] = hy; B.scal[O_BZ] = hz;
  }
  // events (events.hpp:11-38), with the pose of the step and the time after it: monitors in order, then marks
  const int nmon = P.M.nmon, nmark = P.M.nmark;
  for (int m0 = 0; m0 < nmon; m0 += 64) {
    const int m = m0 + lane;
    const bool fire = m < nmon && B.over[m] > P.M.monThreshold && B.overT[m] < 0;
    uint64_t mask = __ballot(fire);
    if (lane == 0) {
      while (mask) {
        const int k = m0 + __builtin_ctzll(mask);
        mask &= mask - 1;
        B.overT[k] = tn;
        push_event(P, B, tn, k, 0);
      }
    }
  }
  for (int i0 = 0; i0 < nmark; i0 += 64) {
    const int i = i0 + lane;
    bool fire = false;
    if (i < nmark && !(B.markT[i] >= 0)) {
      const double wz = world_z(fr, P.M.mkx[i], P.M.mky[i], P.M.mkz[i]);
      const bool under = P.M.mkWhen[i] == 0;
      fire = (under && wz < 0) || (!under && wz > 0);
    }
    uint64_t mask = __ballot(fire);
    if (lane == 0) {
      while (mask) {
        const int k = i0 + __builtin_ctzll(mask);
        mask &= mask - 1;
        B.markT[k] = tn;
        push_event(P, B, tn, k, 1);
      }
    }
  }
  if (lane == 0 && !wasFoundered &&
      (htop < P.M.hullTopBelow || __builtin_fabs(pitch) > P.M.pitchAbsAbove || (hV < P.M.buoyancyBelow && tn > P.M.minT))) {
    B.scal[S_FOUNDERED] = 1;
    B.scal[S_FOUNDERT] = tn;
    push_event(P, B, tn, -1, 2);
  }
  // the pose of the next step (frame.hpp:9-21); the frame slots are free: every reader of this step is done
  pose_wave0(P, B, lane, zO, pitch, roll);
}

// One step. On entry the frame slots hold the pose of the current state (written by the previous step's tail or
// at load time); every thread may read them. Ends with a barrier.
template <bool Mixed>
__device__ void device_step(PC& P, const Blk& B, int tid) {
  const int lane = tid & (kLanes - 1), unit = tid / kLanes;
  const int NN = P.S.nn, NC = P.M.nc, NG = P.M.ng;

  // (a) buoyancy partials, then (b) the free surfaces: logical group `unit` solves groups unit, unit + 8, ...
  buoyancy_partials<Mixed>(P, B, tid);
  {
    const UnitPass<Mixed> pass{P, B.frame, lane, (tid & 32) != 0};
    SSL double* Pp = B.passes + 7 * P.M.maxGroup * unit;
    for (int g = unit; g < NG; g += kUnits) solve_group_entry(P, pass, g, B, Pp, lane == 0);
  }
  __syncthreads();

  // (c) heads and degrees per node
  for (int n = tid; n < NN; n += kThreads) node_heads(P, B, n);
  __syncthreads();

  // (c) flows per connection
  const double t = B.scal[S_T];
  for (int c = tid; c < NC; c += kThreads) B.dV[c] = connection_flow(P, B.frame, t, B, c);
  __syncthreads();

  // (c) gathers, volume updates and load terms per node; monitor sums (events.hpp:73-75, per monitor in increasing
  // connection order); the sea inflow (flows.hpp:288-292) on the last thread
  for (int n = tid; n < NN; n += kThreads) node_update(P, B, n);
  for (int m = tid; m < P.M.nmon; m += kThreads) {
    double o = 0;
    const int s = P.M.monStart[m], e = s + P.M.monCount[m];
    for (int k = s; k < e; ++k) {
      const double qc = B.q[P.M.monConn[k]];
      if (qc > 0) o += qc;
    }
    B.over[m] = o;
  }
  if (tid == kThreads - 1) {
    double s = 0;
    for (int k = 0; k < P.M.seaCount; ++k) s += B.dV[P.M.sea[k]];
    B.scal[S_SEA] = s;
  }
  __syncthreads();

  if (tid < 64) step_tail(P, B, tid);
  __syncthreads();
}

}  // namespace ssa
