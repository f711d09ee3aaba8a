// The flooding step on gfx950: one 256-thread workgroup (four wave64s) per simulation, the simulation's state,
// scratch and connection overrides in LDS, the shared hull geometry streamed from L2.
//
// Bits: the step reproduces the reference's portable numerics (double) and portable32 numerics (mixed) exactly
// (reference: engine/include/sinksim/kernel/step.hpp:41-54 with Order::Tree / Order::TreeMixed). Three sums
// have a canonical order and are reproduced as such:
//   - buoyancy: 256 logical partials, partial t holding columns t, t+256, ...; each 32-partial group reduced by
//     the halving tree (l absorbs l+16, l+8, ... l+1); the 8 group results combined 0..7 in order
//     (reduce.hpp:12-79, hydro.hpp:48-64). A logical 32-lane group is one half of a wave64.
//   - node pass: 32 lane partials over the node's segments with stride 32, then the same tree (hydro.hpp:112-122).
//   - every other sum (sea inflow, node gathers, loads, monitors) is serial in the reference's index order, one
//     chain per lane where several independent chains exist.
// All other arithmetic is element-wise and written operation for operation as in the reference; the library is
// compiled with -ffp-contract=off so that nothing is fused.
//
// Work layout of a step (| = workgroup barrier):
//   buoyancy partials + the level solve, one logical 32-lane group per free-surface group | per node (a quad of
//   lanes): effective heads and limiter degrees | per connection (weirs first): flows | per node: gather, volume
//   update, load terms; per monitor: flow sums; one lane: sea inflow | wave 0: the tail (loads as seven parallel
//   chains, integration, events, and the pose of the next step) |
// Wave 0 issues at a raised priority from the flows to the end of the step, and the waves of the level solve from
// their first round on (higher after five rounds): the critical waves of an instance go ahead of the other
// instances' waves on their SIMDs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fdlibm_dev.hpp"

#define SSG __attribute__((address_space(1)))  // global memory
#define SSC __attribute__((address_space(4)))  // constant memory (scalar loads)
#define SSL __attribute__((address_space(3)))  // LDS

namespace ssa {

constexpr int kThreads = 256;  // logical block of the canonical buoyancy order (the first 256 threads of a workgroup)
constexpr int kLanes = 32;     // logical warp of the canonical trees
constexpr int kUnits = kThreads / kLanes;  // buoyancy's 32-lane groups
constexpr int kCurveStride = 8;             // doubles per device curve record: 7 readouts and a pad

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
constexpr double kCalmStart = 1800, kCalmStop = 1200;  // stopWhenStable: calm counts after 30 min, stops past 20 min (simulation.cpp:179-182)
constexpr double kPitchClamp = 1.35;
constexpr double kRollClamp = 1.4;
constexpr double kPi = 3.141592653589793;

// JavaScript Math.max / Math.min for ordinary numbers (reference: kernel/math.hpp:29-30)
__device__ __forceinline__ double jmax(double a, double b) { return (b > a) ? b : a; }
__device__ __forceinline__ double jmin(double a, double b) { return (b < a) ? b : a; }

// four 32-bit integers or four doubles loaded as one vector (the packed per-connection and incidence records)
typedef int32_t i4v __attribute__((ext_vector_type(4)));
typedef double d4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));

using gcd = const SSG double;
using gcf = const SSG float;
using gci = const SSG int32_t;
using gcb = const SSG uint8_t;

// ------------------------------------------------------------------------------------------- device views

// The hull, shared by every simulation, as records: per column (x, y, zlo, zhi, dxdy) and per segment the same
// five values with the segment's own bottom and top in place of the column's (x, y and footprint pre-gathered from
// the segment's column), as 48-byte (or packed 40-byte) double records and 32-byte float records. The float records
// hold the same values rounded once on the host, which is what the portable32 arithmetic converts them to anyway.
struct DevShip {
  int32_t ncol, nseg, nn;
  int32_t recD;  // bytes per double record: 48 (16-byte aligned) or 40 (packed); the float records are 32 bytes
  const SSG d2v* col;
  const SSG f4v* colF;
  const SSG d2v* seg;
  const SSG f4v* segF;
  gci *segStart, *segCount;
  gcd *mu, *vmax, *amin, *aov;
};

// The scenario's immutable network and constants.
struct DevNet {
  double rho, g, sq2g, dt;
  int32_t ng, nMembers;
  gci *gStart, *gCount, *gNodes;
  gcb* gMode;
  gcd *gpx, *gpy, *gpz;
  double gMargin;
  int32_t nc;
  gci *incStart, *incCount;
  const SSG i4v* inc;   // per incidence entry (node order, increasing connection): connection, its a, its b, side (1 = b)
  const SSG i4v* conI;  // per connection: a (a sea connection's -1 - its place in the sea list), b, skipGroup, law
  const SSG d4v* conD;  // per connection: x, y, z, coef
  gci* flowOrder;       // the connection of each flow slot: weirs, then orifices, each in increasing order
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

// Per-instance global arrays of the batch. The instances of one batch may run different models (hulls and
// scenarios), so an instance's rows start at its own offsets; work items are handed out in `order` through the
// counter, which the host resets before each launch.
struct DevBatch {
  int32_t count;
  const SSG int32_t *instModel, *order;                   // per instance its model; per work item its instance
  const SSG int32_t *nodeOff, *connOff, *evOff, *monOff, *markOff;  // per instance, in elements
  SSG int32_t* counter;
  gcb* en;
  gcd *area, *tOn;                      // rows of the instance's connection count, at connOff
  SSG double* st;                       // [count * 7]: t zO pitch roll vz wth wph
  SSG double* out;                      // [count * 7]: inflow Vb hullTop mw bx by bz
  SSG double *vol, *level, *cx, *cy, *cz;  // rows of the instance's node count, at nodeOff
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
  SSG double* curveRec;  // kCurveStride doubles per record (64-byte aligned)
  SSG int32_t* curveCount;
  SSG double* nextCurve;
  int32_t stopWhenStable;  // stop an instance after 20 min of calm, once past 30 min (simulation.cpp:179-182)
  SSG double* calm;        // [count] the calm time so far (s), kept across launches; > 1200 once stopped
  // Exact stepping (the single-step API): every instance takes exactly its step count in the launch, whether or not it
  // founders or passes tMax (kernel::step does not look at either); stepLimit[inst] is that count (null: the
  // launch's maxSteps for every instance).
  int32_t exact;
  const SSG long long* stepLimit;
};

// LDS layout of one simulation, as offsets in doubles (the host computes it: lds_layout in batch.hip).
struct LdsLayout {
  int32_t vol, level, cx, cy, cz, cvol, area, zmin, zmax, hEff, aEff, excess;
  int32_t tX, tY, tF, tM, tIp, tIr;
  int32_t dV, q, carea, ctOn, seaDV, passes, warpHyd, frame, scal, over, overT, markT, evT, nextGroup;
  int32_t nMu, nVmax, nAov, nAmin, gpx, gpy, gpz;  // copies of the node and group tables
  int32_t deg, evId, nSegStart, nSegCount, gStart, gCount, gNodes;  // int32 arrays, offsets in doubles
  int32_t merged, cen, evKind, gMode;                               // byte arrays, offsets in doubles
  int32_t total;                // doubles
};

// One model (hull, scenario and its LDS layout) in constant memory; device code reads fields on demand with scalar
// loads. A batch holds one record per model.
struct ModelParams {
  DevShip S;
  DevNet M;
  LdsLayout L;
};
using PC = const SSC ModelParams;
using BC = const SSC DevBatch;

// frame slots: R00 R01 R02 R10 R11 R12 R20 R21 R22 tx ty tz cth iR, then the f32 frame R20 R21 R22 tz iR
enum : int { F_R00 = 0, F_R01, F_R02, F_R10, F_R11, F_R12, F_R20, F_R21, F_R22, F_TX, F_TY, F_TZ, F_CTH, F_IR, F_F32, F_COUNT = F_F32 + 3 };

// scalar slots of the `scal` block
enum : int { S_T = 0, S_ZO, S_PITCH, S_ROLL, S_VZ, S_WTH, S_WPH, O_INFLOW, O_VB, O_TOP, O_MW, O_BX, O_BY, O_BZ, S_SEA, S_FOUNDERT, S_EVCOUNT,
             S_FOUNDERED, S_NEXTCURVE, S_CURVECOUNT, S_CALM, S_COUNT };

struct Blk {
  // cx, cy, cz hold a node's centroid between steps; from a node's last pass in the level solve to the heads phase
  // they hold that pass's moments and cvol its volume (0: no new centroid this step), and the heads phase divides
  // (cvol is 0 at every step's start)
  SSL double *vol, *level, *cx, *cy, *cz, *cvol, *area, *zmin, *zmax, *hEff, *aEff, *excess;
  SSL double *tX, *tY, *tF, *tM, *tIp, *tIr;
  SSL double *dV, *q, *carea, *ctOn, *seaDV, *passes, *warpHyd, *frame, *scal, *over, *overT, *markT, *evT;
  SSL int32_t* nextGroup;  // the next free-surface group a solver unit takes (reset to 0 before each step's solve)
  SSL double *nMu, *nVmax, *nAov, *nAmin, *gpx, *gpy, *gpz;
  SSL int32_t *deg, *evId, *nSegStart, *nSegCount, *gStart, *gCount, *gNodes;
  SSL uint8_t *merged, *cen, *evKind, *gMode;
};

__device__ inline Blk carve(SSL double* base, PC& P) {
  const SSC LdsLayout& L = P.L;
  Blk b;
  b.vol = base + L.vol; b.level = base + L.level; b.cx = base + L.cx; b.cy = base + L.cy; b.cz = base + L.cz; b.cvol = base + L.cvol;
  b.area = base + L.area; b.zmin = base + L.zmin; b.zmax = base + L.zmax; b.hEff = base + L.hEff; b.aEff = base + L.aEff;
  b.excess = base + L.excess;
  b.tX = base + L.tX; b.tY = base + L.tY; b.tF = base + L.tF; b.tM = base + L.tM; b.tIp = base + L.tIp; b.tIr = base + L.tIr;
  b.dV = base + L.dV; b.q = base + L.q; b.carea = base + L.carea; b.ctOn = base + L.ctOn; b.seaDV = base + L.seaDV; b.passes = base + L.passes;
  b.warpHyd = base + L.warpHyd; b.frame = base + L.frame; b.scal = base + L.scal; b.over = base + L.over; b.overT = base + L.overT;
  b.markT = base + L.markT; b.evT = base + L.evT;
  b.nextGroup = reinterpret_cast<SSL int32_t*>(base + L.nextGroup);
  b.nMu = base + L.nMu; b.nVmax = base + L.nVmax; b.nAov = base + L.nAov; b.nAmin = base + L.nAmin;
  b.gpx = base + L.gpx; b.gpy = base + L.gpy; b.gpz = base + L.gpz;
  b.nSegStart = reinterpret_cast<SSL int32_t*>(base + L.nSegStart);
  b.nSegCount = reinterpret_cast<SSL int32_t*>(base + L.nSegCount);
  b.gStart = reinterpret_cast<SSL int32_t*>(base + L.gStart);
  b.gCount = reinterpret_cast<SSL int32_t*>(base + L.gCount);
  b.gNodes = reinterpret_cast<SSL int32_t*>(base + L.gNodes);
  b.gMode = reinterpret_cast<SSL uint8_t*>(base + L.gMode);
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

// reference: kernel/frame.hpp:10-21, written to the frame slots, with iR = 1 / R22 (hydro.hpp:15,51,70,115)
// and the f32 frame of reduce.hpp:86-98
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
// 16; off = 16 by an LDS-unit swizzle that swaps the two rows of each group). Only lanes l < off use the result;
// the row shift gives the other lanes of a row whose source lies past the row's end 0 (bound control), so the move
// writes every lane and needs no copy of the destination's old value first.
template <int OFF>
__device__ __forceinline__ uint32_t down32(uint32_t v) {
  if constexpr (OFF == 16) return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(static_cast<int>(v), 0x401F));
  else return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x100 + OFF, 0xF, 0xF, true));
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

// Lane 0 of each logical 32-lane group (wave lanes 0 and 32) to every lane of its group: an LDS-unit swizzle whose
// source lane is ((l & 0) | 0) ^ 0 = 0 within each group of 32.
__device__ __forceinline__ uint32_t lead32(uint32_t v) { return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(static_cast<int>(v), 0x0000)); }
__device__ __forceinline__ double lead(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  return __builtin_bit_cast(double, (static_cast<uint64_t>(lead32(static_cast<uint32_t>(b >> 32))) << 32) | lead32(static_cast<uint32_t>(b)));
}
__device__ __forceinline__ float lead(float v) { return __builtin_bit_cast(float, lead32(__builtin_bit_cast(uint32_t, v))); }

// Lane 16 of each logical 32-lane group (wave lanes 16 and 48) to every lane of its group: source lane
// ((l & 0) | 16) ^ 0 = 16 within each group of 32.
__device__ __forceinline__ uint32_t lead16_32(uint32_t v) { return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(static_cast<int>(v), 0x0200)); }
__device__ __forceinline__ double lead16(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  return __builtin_bit_cast(double, (static_cast<uint64_t>(lead16_32(static_cast<uint32_t>(b >> 32))) << 32) | lead16_32(static_cast<uint32_t>(b)));
}
__device__ __forceinline__ float lead16(float v) { return __builtin_bit_cast(float, lead16_32(__builtin_bit_cast(uint32_t, v))); }

// Lanes 8 and 24 of each logical 32-lane group to every lane of the group: source lane ((l & 0) | 8) and
// ((l & 0) | 24) within each group of 32.
template <int OR>
__device__ __forceinline__ uint32_t lane_of32(uint32_t v) { return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(static_cast<int>(v), OR << 5)); }
template <int OR>
__device__ __forceinline__ double lane_of(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  return __builtin_bit_cast(double, (static_cast<uint64_t>(lane_of32<OR>(static_cast<uint32_t>(b >> 32))) << 32) | lane_of32<OR>(static_cast<uint32_t>(b)));
}
template <int OR>
__device__ __forceinline__ float lane_of(float v) { return __builtin_bit_cast(float, lane_of32<OR>(__builtin_bit_cast(uint32_t, v))); }
template <class T>
__device__ __forceinline__ T lead8(T v) { return lane_of<8>(v); }
template <class T>
__device__ __forceinline__ T lead24(T v) { return lane_of<24>(v); }

// Two values of a logical 32-lane group, a and b, after the first level of the halving tree (l absorbs l + 16)
// packed into one register: lanes 0..15 of the group end with a's sums and lanes 16..31 with b's. The gfx950
// v_permlane16_swap exchanges rows 1 and 3 (16 lanes each) of its first operand with rows 0 and 2 of its second:
// afterwards the first holds a[l] in lanes 0..15 and b[l - 16] in lanes 16..31, the second a[l + 16] and b[l], so
// first + second is the tree's own sum with its own operand order in both halves. The later levels (+8 ... +1) stay
// inside a row of 16 lanes and work on the packed register unchanged; a's total ends in lane 0, b's in lane 16.
__device__ __forceinline__ void swap16(uint32_t& x, uint32_t& y) {
  const auto r = __builtin_amdgcn_permlane16_swap(x, y, false, false);
  x = r[0]; y = r[1];
}
__device__ __forceinline__ void swap16(double& a, double& b) {
  const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
  uint32_t alo = static_cast<uint32_t>(ua), ahi = static_cast<uint32_t>(ua >> 32);
  uint32_t blo = static_cast<uint32_t>(ub), bhi = static_cast<uint32_t>(ub >> 32);
  swap16(alo, blo); swap16(ahi, bhi);
  a = __builtin_bit_cast(double, (static_cast<uint64_t>(ahi) << 32) | alo);
  b = __builtin_bit_cast(double, (static_cast<uint64_t>(bhi) << 32) | blo);
}
__device__ __forceinline__ void swap16(float& a, float& b) {
  uint32_t x = __builtin_bit_cast(uint32_t, a), y = __builtin_bit_cast(uint32_t, b);
  swap16(x, y);
  a = __builtin_bit_cast(float, x); b = __builtin_bit_cast(float, y);
}
// first level of two sum trees, packed
template <class T>
__device__ __forceinline__ T pack_sum16(T a, T b) { swap16(a, b); return a + b; }
// first level of two minimum trees (`if (other < own) own = other`), packed; own is the first after the swap
template <class T>
__device__ __forceinline__ T pack_min16(T a, T b) { swap16(a, b); return b < a ? b : a; }

// ------------------------------------------------------------------------------------------- partials

template <class T>
struct HP {
  T V, Mx, My, Mz, top;
};
template <class T>
struct PP {
  T vol, dv, mx, my, mz, zmin, zmax;
};

// One level of the halving tree (reduce.hpp:74-79: l absorbs l + 16, 8, 4, 2, 1; the total in lane 0 of the
// group). Lanes l < OFF absorb lane l + OFF; the other lanes are never read again, so they add too (their sums are
// discarded) and no lane mask is needed.
template <int OFF, class T>
__device__ __forceinline__ void sum_level(T& v) { v += down<OFF>(v); }
template <int OFF, class T>
__device__ __forceinline__ void min_level(T& v) {
  const T o = down<OFF>(v);
  if (o < v) v = o;
}

// The node pass's halving trees (the reference's sums and comparisons, operand for operand), with (vol, dv) and
// (mx, my) packed in pairs, and with Ext (zmin, -zmax) as a pair of minimum trees: -max(-a, -b) is max(a, b)
// bit for bit, ties and NaN keeping the lane's own value in both. The results stay packed: vol in lane 0, dv in lane
// 16, mx in lane 8 and my in lane 24 of each group (q), mz in lane 0 (z), with Ext zmin in lane 0 and -zmax in lane 16
// (mm).
template <class T>
struct PackedPass {
  T q, z, mm;
};
// a's lanes 0..7 of each row of 16 with b's lanes 0..7 moved into lanes 8..15: a row shift right by 8 that writes
// banks 2 and 3 (lanes 8..15) only, the other lanes keeping a
__device__ __forceinline__ uint32_t merge8_32(uint32_t a, uint32_t b) {
  return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(a), static_cast<int>(b), 0x118, 0xF, 0xC, false));
}
__device__ __forceinline__ double merge8(double a, double b) {
  const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
  const uint32_t lo = merge8_32(static_cast<uint32_t>(ua), static_cast<uint32_t>(ub));
  const uint32_t hi = merge8_32(static_cast<uint32_t>(ua >> 32), static_cast<uint32_t>(ub >> 32));
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}
__device__ __forceinline__ float merge8(float a, float b) {
  return __builtin_bit_cast(float, merge8_32(__builtin_bit_cast(uint32_t, a), __builtin_bit_cast(uint32_t, b)));
}
template <bool Ext, class T>
__device__ __forceinline__ PackedPass<T> tree32_levels(const PP<T>& p) {
  T vd = pack_sum16(p.vol, p.dv), xy = pack_sum16(p.mx, p.my), z = p.mz;
  sum_level<16>(z);
  T mm = 0;
  if constexpr (Ext) mm = pack_min16(p.zmin, -p.zmax);
  sum_level<8>(vd); sum_level<8>(xy); sum_level<8>(z);
  if constexpr (Ext) min_level<8>(mm);
  // after +8 a tree lives in lanes 0..7 of its row: the (mx, my) trees move beside (vol, dv), and the +4, +2 and +1
  // levels serve all four (lanes 0..3 of a row absorb 4..7 and lanes 8..11 absorb 12..15, and so on)
  T q = merge8(vd, xy);
  sum_level<4>(q); sum_level<4>(z);
  if constexpr (Ext) min_level<4>(mm);
  sum_level<2>(q); sum_level<2>(z);
  if constexpr (Ext) min_level<2>(mm);
  sum_level<1>(q); sum_level<1>(z);
  if constexpr (Ext) min_level<1>(mm);
  return PackedPass<T>{q, z, mm};
}
// the trees' totals in every lane of the group
template <bool Ext, class T>
__device__ __forceinline__ PP<T> tree32_all(const PP<T>& p) {
  const PackedPass<T> q = tree32_levels<Ext>(p);
  PP<T> r;
  r.vol = lead(q.q); r.dv = lead16(q.q); r.mx = lead8(q.q); r.my = lead24(q.q); r.mz = lead(q.z);
  r.zmin = 0; r.zmax = 0;
  if constexpr (Ext) { r.zmin = lead(q.mm); r.zmax = -lead16(q.mm); }
  return r;
}
// lane L's value (the same in every lane, in scalar registers)
template <int L>
__device__ __forceinline__ float at_lane(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), L));
}
template <int L>
__device__ __forceinline__ double at_lane(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(b)), L));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(b >> 32)), L));
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}
// the group starting at lane L0's totals, read from the packed trees
template <int L0, bool Ext, class T>
__device__ __forceinline__ PP<T> group_totals(const PackedPass<T>& q) {
  PP<T> r;
  r.vol = at_lane<L0>(q.q); r.dv = at_lane<L0 + 16>(q.q); r.mx = at_lane<L0 + 8>(q.q); r.my = at_lane<L0 + 24>(q.q);
  r.mz = at_lane<L0>(q.z);
  r.zmin = 0; r.zmax = 0;
  if constexpr (Ext) { r.zmin = at_lane<L0>(q.mm); r.zmax = -at_lane<L0 + 16>(q.mm); }
  return r;
}

// The buoyancy's halving trees (sums, and `if (other > own) own = other` for top), with (V, Mx) and (My, Mz)
// packed: V, My and top end in lane 0 of the group, Mx and Mz in lane 16.
template <class T>
__device__ __forceinline__ void tree32_packed(const HP<T>& p, T& vx, T& yz, T& top) {
  vx = pack_sum16(p.V, p.Mx); yz = pack_sum16(p.My, p.Mz); top = p.top;
  T o = down<16>(top);
  if (o > top) top = o;
#define SS_HP_LEVEL(OFF) sum_level<OFF>(vx); sum_level<OFF>(yz); o = down<OFF>(top); if (o > top) top = o;
  SS_HP_LEVEL(8) SS_HP_LEVEL(4) SS_HP_LEVEL(2) SS_HP_LEVEL(1)
#undef SS_HP_LEVEL
}

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

// A column's or a segment's geometry: x, y, bottom, top and footprint (for a segment: those of its column, with
// the segment's own bottom and top). Records are 48 or 40 bytes (double) or 32 bytes (float), addressed by a 32-bit
// byte offset from the table's base.
//
// Loads and their registers: a load writes every register of its destination when it lands, used or not, so a
// register the compiler believes free (the unused rest of a vector) cannot take another value until the load has
// landed, and the compiler waits for the load before reusing it. A double record is therefore loaded as two
// 16-byte loads and one 8-byte load of the footprint (no unused register). A float record keeps two 16-byte loads
// (measured faster than a 16-byte and a 4-byte load), and the three padding lanes of the second are held live
// (`hold`, an empty asm that reads them) until the record's wait in `stream`, after which they are free.
template <class T>
struct GeoPad {};
template <>
struct GeoPad<float> {
  float p1, p2, p3;
};
template <class T>
struct Geo {
  T x, y, lo, hi, A;
  GeoPad<T> pad;
};
__device__ __forceinline__ void hold(const Geo<double>&) {}
__device__ __forceinline__ void hold(const Geo<float>& g) { asm volatile("" ::"v"(g.pad.p1), "v"(g.pad.p2), "v"(g.pad.p3)); }
// record sizes: five doubles padded to 48 bytes or packed in 40 (a batch's choice: the padded records are read
// faster when a hull stays in the caches, the packed ones when hulls stream from memory), five floats padded to 32
template <class T>
__device__ __forceinline__ uint32_t rec_bytes(PC& P) { return sizeof(T) == sizeof(double) ? static_cast<uint32_t>(P.S.recD) : 32u; }
// s_waitcnt immediates (gfx9 encoding: vmcnt in bits 3:0, expcnt 6:4 and lgkmcnt 11:8 at their maxima, i.e. not
// waited for) that wait until at most one record's loads are in flight: three loads per double record, two per
// float record
template <class T>
constexpr int kWaitOneRecord = sizeof(T) == sizeof(double) ? 0x0F73 : 0x0F72;
constexpr int kWaitAll = 0x0F70;  // no vector memory load in flight
// (a packed double record is 8-byte aligned: its pairs are loaded as 16-byte loads of 8-byte alignment)
typedef double d2v8 __attribute__((ext_vector_type(2), aligned(8)));
__device__ __forceinline__ Geo<double> geo_at(const SSG d2v* table, uint32_t off) {
  const SSG d2v8* r = (const SSG d2v8*)((const SSG char*)table + off);
  const d2v8 a = r[0], b = r[1];
  const double A = *(const SSG double*)(r + 2);
  return Geo<double>{a.x, a.y, b.x, b.y, A, {}};
}
__device__ __forceinline__ Geo<float> geo_at(const SSG f4v* table, uint32_t off) {
  const SSG f4v* r = (const SSG f4v*)((const SSG char*)table + off);
  const f4v a = r[0], b = r[1];
  return Geo<float>{a.x, a.y, a.z, a.w, b.x, {b.y, b.z, b.w}};
}

// one column into a buoyancy partial (reduce.hpp:39-49 / 131-142)
__device__ __forceinline__ void column(const Fr64& F, const Geo<double>& g, bool on, HP<double>& p) {
  const double x = g.x, y = g.y, lo = g.lo, hi = g.hi;
  const double base = F.R20 * x + F.R21 * y + F.tz;
  const double wt = base + F.R22 * hi;
  if (on && wt > p.top) p.top = wt;
  const double zp = -base * F.iR;
  if (!on || zp <= lo) return;
  const double s = zp >= hi ? hi - lo : zp - lo;
  const double v = g.A * s;
  p.V += v; p.Mx += v * x; p.My += v * y; p.Mz += v * (lo + 0.5 * s);
}
__device__ __forceinline__ void column(const Fr32& f, const Geo<float>& g, bool on, HP<float>& p) {
  const float x = g.x, y = g.y, lo = g.lo, hi = g.hi;
  const float base = f.R20 * x + f.R21 * y + f.tz;
  const float wt = base + f.R22 * hi;
  if (on && wt > p.top) p.top = wt;
  const float zp = -base * f.iR;
  if (!on || zp <= lo) return;
  const float s = zp >= hi ? hi - lo : zp - lo;
  const float v = g.A * s;
  p.V += v; p.Mx += v * x; p.My += v * y; p.Mz += v * (lo + 0.5f * s);
}

// One segment into a node-pass partial (reduce.hpp:52-71 / 144-164). The world extent of the node (zmin, zmax)
// depends on the pose only, not on the level: it is accumulated (Ext) in the first pass of a node in a step and
// reused by the later ones, which the reference recomputes to the same values.
template <bool Ext>
__device__ __forceinline__ void segment(const Fr64& F, double h, const Geo<double>& g, bool on, PP<double>& p) {
  const double x = g.x, y = g.y, a = g.lo, b = g.hi;
  const double base = F.R20 * x + F.R21 * y + F.tz;
  if constexpr (Ext) {
    const double wa = base + F.R22 * a, wb = base + F.R22 * b;
    if (on && wa < p.zmin) p.zmin = wa;
    if (on && wb > p.zmax) p.zmax = wb;
  }
  const double zp = (h - base) * F.iR;
  if (!on || zp <= a) return;
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
template <bool Ext>
__device__ __forceinline__ void segment(const Fr32& f, float h, const Geo<float>& g, bool on, PP<float>& p) {
  const float x = g.x, y = g.y, a = g.lo, b = g.hi;
  const float base = f.R20 * x + f.R21 * y + f.tz;
  if constexpr (Ext) {
    const float wa = base + f.R22 * a, wb = base + f.R22 * b;
    if (on && wa < p.zmin) p.zmin = wa;
    if (on && wb > p.zmax) p.zmax = wb;
  }
  const float zp = (h - base) * f.iR;
  if (!on || zp <= a) return;
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

// A lane's partial over the records first, first + step, ... below end, in order, each record's load issued while
// the previous record is evaluated. Unrolled by two so that the two records in flight keep their registers (no
// copies between iterations); records are addressed by byte offsets advanced by step records (no multiplication
// per record). The loop is a scalar one with its only exit at the bottom: it runs the records in pairs while at
// least two of the wave's rounds remain (the rounds are the wave's longest lane's record count: within each half of
// the wave `first` ascends with the lane and `end` is common, so the lowest active lane of each half has the most
// records), and an odd last round is evaluated after the loop, from the record already in flight. A lane past its last record
// reloads that record (in bounds), a lane without records reads record 0 (in bounds, since another lane of the wave
// has records), and eval(g, on) leaves the partial unchanged when on is false. Before each evaluation an explicit
// wait for that record's loads (the other record's stay in flight) is made on every path, also where no lane
// evaluates. The compiler's own waits would otherwise come at the loop's head, before the next record's loads are
// issued, on any path that can reach the head with loads in flight (a lane leaving the loop early, a break in the
// middle, an evaluation skipped by a branch): one record in flight instead of two. The last round's second record
// (an odd round count) would be a whole evaluation with no lane on; a break before it is such a path (the compiler
// routes it through the loop's latch), hence the evaluation after the loop.
// The scheduling barriers keep the next record's loads ahead of the current record's arithmetic: without them
// the compiler sinks the loads below the evaluation and waits for them at once (vmcnt(0)).
template <class T, class Tab, class Fn>
__device__ __forceinline__ void stream(const Tab* table, uint32_t rec, int first, int end, int step, Fn&& eval) {
  const int mine = first < end ? (end - first + step - 1) / step : 0;  // this lane's records
  const uint64_t ex = __builtin_amdgcn_read_exec();
  const uint32_t ex0 = static_cast<uint32_t>(ex), ex1 = static_cast<uint32_t>(ex >> 32);
  int rounds = 0;
  if (ex0) rounds = __builtin_amdgcn_readlane(mine, __builtin_ctz(ex0));
  if (ex1) rounds = max(rounds, __builtin_amdgcn_readlane(mine, 32 + __builtin_ctz(ex1)));
  if (rounds == 0) return;
  const uint32_t stride = static_cast<uint32_t>(step) * rec;
  const uint32_t last = mine > 0 ? static_cast<uint32_t>(first + (mine - 1) * step) * rec : 0u;
  uint32_t off = mine > 0 ? static_cast<uint32_t>(first) * rec : 0u;
  Geo<T> g0 = geo_at(table, off);
  int k = 0;
  while (k + 1 < rounds) {
    const uint32_t off1 = min(off + stride, last);
    const Geo<T> g1 = geo_at(table, off1);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(kWaitOneRecord<T>);  // g0 has arrived; g1 stays in flight
    hold(g0);
    eval(g0, k < mine);
    off = min(off1 + stride, last);
    g0 = geo_at(table, off);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(kWaitOneRecord<T>);  // g1 has arrived; the next g0 stays in flight
    hold(g1);
    eval(g1, k + 1 < mine);
    k += 2;
  }
  if (k < rounds) {
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(kWaitAll);  // the last record, the only one in flight
    hold(g0);
    eval(g0, k < mine);
  }
}

struct Pass {
  double vol, dv, mx, my, mz, zmin, zmax;
};

// A lane's partials of the node pass of node n at level h (n < 0: none), in the numerics' type.
template <bool Mixed, bool Ext>
__device__ __forceinline__ auto pass_partials(PC& P, const Blk& B, int lane, int n, double h) {
  const SSL double* fr = B.frame;
  int s0 = 0, s1 = 0;
  if (n >= 0) {
    s0 = B.nSegStart[n];
    s1 = s0 + B.nSegCount[n];
  }
  if constexpr (Mixed) {
    const Fr32 f = fr32(fr);
    const float hf = static_cast<float>(h);
    PP<float> p{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1e9f, -1e9f};
    stream<float>(P.S.segF, rec_bytes<float>(P), s0 + lane, s1, kLanes, [&](const Geo<float>& g, bool on) { segment<Ext>(f, hf, g, on, p); });
    return p;
  } else {
    const Fr64 F = fr64(fr);
    PP<double> p{0.0, 0.0, 0.0, 0.0, 0.0, 1e9, -1e9};
    stream<double>(P.S.seg, rec_bytes<double>(P), s0 + lane, s1, kLanes, [&](const Geo<double>& g, bool on) { segment<Ext>(F, h, g, on, p); });
    return p;
  }
}
// finish_pass (hydro.hpp:101-109) of a node's totals t with its mu and the frame's iR
template <class T>
__device__ __forceinline__ Pass finish_pass(const PP<T>& t, double mu, double iR) {
  Pass r;
  r.vol = static_cast<double>(t.vol) * mu;
  r.dv = static_cast<double>(t.dv) * mu * iR;
  r.mx = static_cast<double>(t.mx) * mu; r.my = static_cast<double>(t.my) * mu; r.mz = static_cast<double>(t.mz) * mu;
  r.zmin = static_cast<double>(t.zmin); r.zmax = static_cast<double>(t.zmax);
  return r;
}

// The node pass of node n at level h by one logical 32-lane group (hydro.hpp:112-122 / 144-155; finish_pass
// hydro.hpp:101-109); the result is in every lane of the group. Both groups of a wave call it together, each with
// its own request; n < 0 is no request (an empty pass whose result is not used). zmin and zmax are computed only
// with Ext.
template <bool Mixed, bool Ext>
__device__ __forceinline__ Pass node_pass(PC& P, const Blk& B, int lane, int n, double h) {
  const auto t = tree32_all<Ext>(pass_partials<Mixed, Ext>(P, B, lane, n, h));
  return finish_pass(t, B.nMu[n >= 0 ? n : 0], B.frame[F_IR]);
}
// The passes of a wave's two groups at one level, group 0 on node n0 and group 1 on node n1 (each group's lane
// gives its own node, n), their totals read from the trees' lanes into scalar registers (no broadcast): both results
// in every lane.
template <bool Mixed, bool Ext>
__device__ __forceinline__ void node_pass_pair(PC& P, const Blk& B, int lane, int n, double h, int n0, int n1, Pass& p0, Pass& p1) {
  const auto q = tree32_levels<Ext>(pass_partials<Mixed, Ext>(P, B, lane, n, h));
  const double iR = B.frame[F_IR];
  p0 = finish_pass(group_totals<0, Ext>(q), B.nMu[n0], iR);
  p1 = finish_pass(group_totals<32, Ext>(q), B.nMu[n1], iR);
}

// ------------------------------------------------------------------------------------------- (b) levels

// A node's new centroid (hydro.hpp:101-109: the pass's moments over its volume, kept when the volume is too small),
// left as the moments and the volume for the heads phase to divide (finish_centroid): the three divisions leave the
// level solve's chain of dependent work.
__device__ __forceinline__ void put_centroid(const Blk& B, int n, const Pass& p) {
  if (p.vol > kCentroidMinVolume) { B.cx[n] = p.mx; B.cy[n] = p.my; B.cz[n] = p.mz; B.cvol[n] = p.vol; }
}
__device__ __forceinline__ void finish_centroid(const Blk& B, int n) {
  const double v = B.cvol[n];
  if (v > 0) {
    B.cx[n] = B.cx[n] / v; B.cy[n] = B.cy[n] / v; B.cz[n] = B.cz[n] / v;
    B.cvol[n] = 0;
  }
}

// Pp (a unit's pass scratch for the members of a merged group) is written by every lane of the unit (each lane
// holds the same pass result) and read back by each lane after its own write.
__device__ __forceinline__ void put_pass(SSL double* Pp, int k, const Pass& p) {
  SSL double* q = Pp + 7 * k;
  q[0] = p.vol; q[1] = p.dv; q[2] = p.mx; q[3] = p.my; q[4] = p.mz; q[5] = p.zmin; q[6] = p.zmax;
}
__device__ __forceinline__ Pass get_pass(const SSL double* Pp, int k) {
  const SSL double* q = Pp + 7 * k;
  return Pass{q[0], q[1], q[2], q[3], q[4], q[5], q[6]};
}

// The level solve (levels.hpp:13-194 with full = false), one free-surface group per wave. The waves of a workgroup
// take groups from an LDS counter. A merged group's state is the same in every lane: its member passes run two per
// round, member k on half k & 1 of the wave at the group's common level, and the Newton update, clamps and the share
// between the members run once per wave under uniform (scalar) branches. An unmerged group's members are solved singly,
// members 0, 2, 4, ... on half 0 and 1, 3, 5, ... on half 1, each half running its own single-node solve; both halves
// issue their passes together. Every pass, comparison and sum is the reference's, in the reference's order.

// a condition every lane of the wave holds alike, as a scalar branch
__device__ __forceinline__ bool uni(bool c) { return __builtin_amdgcn_readfirstlane(c ? 1 : 0) != 0; }

// issue priority only (no value changes): after its first round a solving wave goes ahead of the other instances'
// waves in their other phases on its SIMD, and after five rounds (a straggler, which the step's barrier waits for)
// further ahead
__device__ __forceinline__ void solve_round(int& round) {
  ++round;
  if (round == 1) __builtin_amdgcn_s_setprio(1);
  if (round == 5) __builtin_amdgcn_s_setprio(2);
}

// The passes of the two members of a merged pair at level h in one round, member k on half k: both results, the same in
// every lane (with Ext also each member's extent into zmin / zmax).
template <bool Mixed, bool Ext>
__device__ __forceinline__ void pair_passes(PC& P, const Blk& B, int base, double h, int lane, int half, int& round, Pass& p0, Pass& p1) {
  const int n0 = B.gNodes[base], n1 = B.gNodes[base + 1];
  node_pass_pair<Mixed, Ext>(P, B, lane, half ? n1 : n0, h, n0, n1, p0, p1);
  if constexpr (Ext)
    if ((lane | half) == 0) { B.zmin[n0] = p0.zmin; B.zmax[n0] = p0.zmax; B.zmin[n1] = p1.zmin; B.zmax[n1] = p1.zmax; }
  solve_round(round);
}

// levels.hpp:54-162 for a merged pair (N = 2, the common case), the members' pass results in registers.
template <bool Mixed>
__device__ void solve_pair(PC& P, const Blk& B, int base, int lane, int half, int& round) {
  const bool lead = (lane | half) == 0;
  const int m0 = B.gNodes[base], m1 = B.gNodes[base + 1];
  const double V = (0.0 + B.vol[m0]) + B.vol[m1], Vmax = (0.0 + B.nVmax[m0]) + B.nVmax[m1];
  const bool w0 = B.vol[m0] > 0, w1 = B.vol[m1] > 0;
  double h;
  if (uni(w0 && w1)) h = ((0.0 + B.level[m0]) + B.level[m1]) / 2.0;
  else h = B.level[w0 ? m0 : m1];
  Pass p0, p1;
  pair_passes<Mixed, true>(P, B, base, h, lane, half, round, p0, p1);
  double lo = p0.zmin, hi = p0.zmax;
  if (p1.zmin < lo) lo = p1.zmin;
  if (p1.zmax > hi) hi = p1.zmax;
  if (uni(V <= kEmptyVolume)) {
    if (lead) { B.vol[m0] = 0; B.level[m0] = p0.zmin; B.area[m0] = 0; B.vol[m1] = 0; B.level[m1] = p1.zmin; B.area[m1] = 0; }
    return;
  }
  if (uni(V >= Vmax * kFullFraction)) {
    const double over = jmax(0.0, V - Vmax);
    const double dh = over / ((0.0 + B.nAov[m0]) + B.nAov[m1]);
    if (lead) {
      const double v = B.nVmax[m0] + dh * B.nAov[m0];
      B.vol[m0] = v;
      B.vol[m1] = V - (0.0 + v);
    }
    pair_passes<Mixed, false>(P, B, base, hi, lane, half, round, p0, p1);
    if (lead) {
      B.level[m0] = hi + dh; B.area[m0] = B.nAov[m0]; put_centroid(B, m0, p0);
      B.level[m1] = hi + dh; B.area[m1] = B.nAov[m1]; put_centroid(B, m1, p1);
    }
    return;
  }
  const double tol = jmax(1e-3 * Vmax, 0.5);
  double a = lo, b = hi;
  if (uni(!(h >= lo && h <= hi))) {
    h = jmin(hi, jmax(lo, h));
    pair_passes<Mixed, false>(P, B, base, h, lane, half, round, p0, p1);
  }
  for (int it = 0;; ++it) {
    const double err = ((0.0 + p0.vol) + p1.vol) - V;
    if (uni(it >= kNewtonFast || __builtin_fabs(err) < tol)) break;
    if (err > 0) b = h; else a = h;
    const double dv = (0.0 + p0.dv) + p1.dv;
    double hn = dv > kSlopeMin ? h - err / dv : 0.5 * (a + b);
    if (!(hn > a && hn < b)) hn = 0.5 * (a + b);
    h = hn;
    pair_passes<Mixed, false>(P, B, base, h, lane, half, round, p0, p1);
  }
  if (lead) {
    const double tot = (0.0 + p0.vol) + p1.vol;
    double v = tot > 0 ? V * p0.vol / tot : V / 2.0;
    if (v > B.nVmax[m0]) v = B.nVmax[m0];
    double assigned = 0.0 + v;
    if (V - assigned > B.nVmax[m1]) {
      // the cascade of levels.hpp:136-155 for two members: the first takes V - vmax(last), uncapped
      v = (V - B.nVmax[m1] - 0.0) - 0.0;
      assigned = 0.0 + v;
    }
    B.vol[m0] = v;
    B.vol[m1] = V - assigned;
    B.level[m0] = h; B.area[m0] = p0.dv; put_centroid(B, m0, p0);
    B.level[m1] = h; B.area[m1] = p1.dv; put_centroid(B, m1, p1);
  }
}

// The passes of members 0..N-1 of a merged group at level h, two per round (member k on half k & 1), into the wave's
// pass scratch Pp (with Ext also each member's extent into zmin / zmax). The lanes of each half hold their pass's result;
// its lane 0 stores it, and the wave reads the scratch back after the compiler barrier, in its own order of LDS accesses.
template <bool Mixed, bool Ext>
__device__ __forceinline__ void member_passes(PC& P, const Blk& B, SSL double* Pp, int base, int N, double h, int lane, int half, int& round) {
  for (int k0 = 0; k0 < N; k0 += 2) {
    const int k = k0 + half;
    const int n = k < N ? B.gNodes[base + k] : -1;
    const Pass p = node_pass<Mixed, Ext>(P, B, lane, n, h);
    if (lane == 0 && k < N) {
      put_pass(Pp, k, p);
      if constexpr (Ext) { B.zmin[n] = p.zmin; B.zmax[n] = p.zmax; }
    }
    solve_round(round);
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// levels.hpp:54-162: the members of a merged group under one free surface. Every value here is the same in every lane.
template <bool Mixed>
__device__ void solve_merged(PC& P, const Blk& B, SSL double* Pp, int base, int N, int lane, int half, int& round) {
  const bool lead = (lane | half) == 0;  // lane 0 of the wave writes the results
  double V = 0, Vmax = 0;
  for (int k = 0; k < N; ++k) { V += B.vol[B.gNodes[base + k]]; Vmax += B.nVmax[B.gNodes[base + k]]; }
  // the common starting level: the members' mean level when all are wet, else the first wet member's
  bool allWet = true;
  int firstWet = -1;
  for (int k = 0; k < N; ++k) {
    if (B.vol[B.gNodes[base + k]] > 0) { if (firstWet < 0) firstWet = k; }
    else allWet = false;
  }
  double h;
  if (uni(allWet)) {
    double s = 0;
    for (int k = 0; k < N; ++k) s += B.level[B.gNodes[base + k]];
    h = s / static_cast<double>(N);
  } else {
    h = B.level[B.gNodes[base + (firstWet >= 0 ? firstWet : N - 1)]];
  }
  member_passes<Mixed, true>(P, B, Pp, base, N, h, lane, half, round);
  double lo = Pp[5], hi = Pp[6];
  for (int k = 1; k < N; ++k) {
    const double zl = Pp[7 * k + 5], zh = Pp[7 * k + 6];
    if (zl < lo) lo = zl;
    if (zh > hi) hi = zh;
  }
  if (uni(V <= kEmptyVolume)) {
    if (lead)
      for (int k = 0; k < N; ++k) { const int m = B.gNodes[base + k]; B.vol[m] = 0; B.level[m] = Pp[7 * k + 5]; B.area[m] = 0; }
    return;
  }
  if (uni(V >= Vmax * kFullFraction)) {
    const double over = jmax(0.0, V - Vmax);
    double aSum = 0;
    for (int k = 0; k < N; ++k) aSum += B.nAov[B.gNodes[base + k]];
    const double dh = over / aSum;
    if (lead) {
      double assigned = 0;
      for (int k = 0; k < N - 1; ++k) {
        const int m = B.gNodes[base + k];
        const double v = B.nVmax[m] + dh * B.nAov[m];
        B.vol[m] = v;
        assigned += v;
      }
      B.vol[B.gNodes[base + N - 1]] = V - assigned;
    }
    member_passes<Mixed, false>(P, B, Pp, base, N, hi, lane, half, round);
    if (lead)
      for (int k = 0; k < N; ++k) {
        const int m = B.gNodes[base + k];
        B.level[m] = hi + dh; B.area[m] = B.nAov[m];
        put_centroid(B, m, get_pass(Pp, k));
      }
    return;
  }
  const double tol = jmax(1e-3 * Vmax, 0.5);
  double a = lo, b = hi;
  if (uni(!(h >= lo && h <= hi))) {
    h = jmin(hi, jmax(lo, h));
    member_passes<Mixed, false>(P, B, Pp, base, N, h, lane, half, round);
  }
  for (int it = 0;; ++it) {
    double sum = 0;
    for (int k = 0; k < N; ++k) sum += Pp[7 * k];
    const double err = sum - V;
    if (uni(it >= kNewtonFast || __builtin_fabs(err) < tol)) break;
    if (err > 0) b = h; else a = h;
    double dv = 0;
    for (int k = 0; k < N; ++k) dv += Pp[7 * k + 1];
    double hn = dv > kSlopeMin ? h - err / dv : 0.5 * (a + b);
    if (!(hn > a && hn < b)) hn = 0.5 * (a + b);
    h = hn;
    member_passes<Mixed, false>(P, B, Pp, base, N, h, lane, half, round);
  }
  // levels.hpp:125-161: the water shared in proportion to the geometry at the final level, capped by each member's
  // capacity; when the last member would overflow, the members before it take V - vmax(last), filled from the one
  // nearest the end
  if (lead) {
    double tot = 0;
    for (int k = 0; k < N; ++k) tot += Pp[7 * k];
    double assigned = 0;
    for (int k = 0; k < N - 1; ++k) {
      const int m = B.gNodes[base + k];
      double v = tot > 0 ? V * Pp[7 * k] / tot : V / static_cast<double>(N);
      if (v > B.nVmax[m]) v = B.nVmax[m];
      B.vol[m] = v;
      assigned += v;
    }
    const double vmaxLast = B.nVmax[B.gNodes[base + N - 1]];
    if (V - assigned > vmaxLast) {
      const double need = V - vmaxLast;
      double capped = 0;
      for (int k = N - 2; k >= 0; --k) {
        double prior = 0;
        for (int j = 0; j < k; ++j) prior += B.vol[B.gNodes[base + j]];
        const int m = B.gNodes[base + k];
        const double v = (need - capped) - prior;
        if (v > B.nVmax[m] && k > 0) {
          B.vol[m] = B.nVmax[m];
          capped += B.nVmax[m];
          continue;
        }
        B.vol[m] = v;
        break;
      }
      assigned = 0;
      for (int k = 0; k < N - 1; ++k) assigned += B.vol[B.gNodes[base + k]];
    }
    B.vol[B.gNodes[base + N - 1]] = V - assigned;
    for (int k = 0; k < N; ++k) {
      const int m = B.gNodes[base + k];
      const Pass p = get_pass(Pp, k);
      B.level[m] = h; B.area[m] = p.dv;
      put_centroid(B, m, p);
    }
  }
}

// One single-node solve (levels.hpp:13-49) per half of a wave, as a state machine whose state names the pass it waits
// for: the first pass at the node's level (with its extent), the pass at the top of a full node, or a Newton pass.
enum : int { LV_DONE = 0, LV_FIRST, LV_FULL, LV_NEWTON };

struct Single {
  int st, n, it;
  double V, Vmax, h, a, b, tol, lo, hi;
};

__device__ __forceinline__ void single_begin(const Blk& B, Single& S, int n) {
  S.n = n;
  S.st = n >= 0 ? LV_FIRST : LV_DONE;
  if (n < 0) return;
  S.V = B.vol[n]; S.Vmax = B.nVmax[n]; S.h = B.level[n];
}

// the state after the pass p the state asked for; the half's lane 0 (lead) writes the node's results
__device__ inline void single_advance(const Blk& B, Single& S, const Pass& p, bool lead) {
  const int n = S.n;
  if (S.st == LV_FULL) {
    if (lead) {
      B.level[n] = S.hi + jmax(0.0, S.V - S.Vmax) / B.nAov[n];
      B.area[n] = B.nAov[n];
      put_centroid(B, n, p);
    }
    S.st = LV_DONE;
    return;
  }
  if (S.st == LV_FIRST) {
    S.lo = p.zmin; S.hi = p.zmax;
    if (lead) { B.zmin[n] = S.lo; B.zmax[n] = S.hi; }
    if (S.V <= kEmptyVolume) {
      if (lead) { B.level[n] = S.lo; B.area[n] = 0; }
      S.st = LV_DONE;
      return;
    }
    if (S.V >= S.Vmax * kFullFraction) { S.st = LV_FULL; return; }
    S.tol = jmax(1e-3 * S.Vmax, 0.5);
    S.a = S.lo; S.b = S.hi; S.it = 0;
    if (S.h < S.lo || S.h > S.hi) {  // the pass at the level clamped into the extent comes next
      S.h = jmin(S.hi, jmax(S.lo, S.h));
      S.st = LV_NEWTON;
      return;
    }
  }
  // the Newton iteration after a pass at S.h
  const double err = p.vol - S.V;
  if (S.it >= kNewtonFast || __builtin_fabs(err) < S.tol) {
    if (lead) { B.level[n] = S.h; B.area[n] = p.dv; put_centroid(B, n, p); }
    S.st = LV_DONE;
    return;
  }
  if (err > 0) S.b = S.h; else S.a = S.h;
  double hn = p.dv > kSlopeMin ? S.h - err / p.dv : 0.5 * (S.a + S.b);
  if (!(hn > S.a && hn < S.b)) hn = 0.5 * (S.a + S.b);
  S.h = hn;
  ++S.it;
  S.st = LV_NEWTON;
}

// The members of an unmerged group solved singly, members half, half + 2, ... on each half of the wave.
template <bool Mixed>
__device__ void solve_singles(PC& P, const Blk& B, int base, int N, int lane, int half, int& round) {
  int next = half;
  Single S{};
  single_begin(B, S, next < N ? B.gNodes[base + next] : -1);
  next += 2;
  while (__ballot(S.st != LV_DONE) != 0) {
    const int n = S.st == LV_DONE ? -1 : S.n;
    const double h = S.st == LV_FULL ? S.hi : S.h;
    const Pass p = __ballot(S.st == LV_FIRST) != 0 ? node_pass<Mixed, true>(P, B, lane, n, h) : node_pass<Mixed, false>(P, B, lane, n, h);
    if (S.st != LV_DONE) single_advance(B, S, p, lane == 0);
    if (S.st == LV_DONE && next < N) {
      single_begin(B, S, B.gNodes[base + next]);
      next += 2;
    }
    solve_round(round);
  }
}

// The level solve of every group: each wave takes the next group from the workgroup's counter (its lane 0's LDS atomic,
// read by the whole wave), decides whether the group's members share a surface this step (levels.hpp:166-188), and
// solves it. The decision reads only the group's own members, whose volumes and levels no other group's solve changes.
template <bool Mixed>
__device__ void solve_levels(PC& P, const Blk& B, int tid) {
  const int lane = tid & (kLanes - 1), half = (tid >> 5) & 1, wave = tid >> 6;
  const int NG = P.M.ng;
  SSL double* Pp = B.passes + 7 * P.M.maxGroup * wave;
  int round = 0;
  for (;;) {
    int g = 0;
    if ((tid & 63) == 0) g = atomicAdd((int*)B.nextGroup, 1);
    g = __builtin_amdgcn_readfirstlane(g);
    if (g >= NG) break;
    const int N = __builtin_amdgcn_readfirstlane(B.gCount[g]);
    const int base = __builtin_amdgcn_readfirstlane(B.gStart[g]);
    const int mode = __builtin_amdgcn_readfirstlane(B.gMode[g]);
    bool merged = false;
    if (N >= 2) {
      if (mode == 1) {
        merged = true;
      } else if (mode == 2) {
        const double zMw = world_z(B.frame, B.gpx[g], B.gpy[g], B.gpz[g]);
        merged = true;
        for (int k = 0; k < N; ++k) {
          const int n = B.gNodes[base + k];
          if (uni(!(B.vol[n] > kEmptyVolume && B.level[n] > zMw + P.M.gMargin))) { merged = false; break; }
        }
      }
    }
    if ((tid & 63) == 0) B.merged[g] = merged ? 1 : 0;
    if (merged && N == 2) solve_pair<Mixed>(P, B, base, lane, half, round);
    else if (merged) solve_merged<Mixed>(P, B, Pp, base, N, lane, half, round);
    else solve_singles<Mixed>(P, B, base, N, lane, half, round);
  }
  if (round >= 1) __builtin_amdgcn_s_setprio(0);
}

// ------------------------------------------------------------------------------------------- (c) network

constexpr int kFetch = 4;  // incidence records a lane fetches before using them in order

__device__ __forceinline__ double head_of(const Blk& B, int n) { return B.vol[n] > kEmptyVolume ? B.level[n] : kNoHead; }

// flows.hpp:15-35: effective head and area of node n, and the number of connections that count toward the
// limiter there (an enabled connection not between two empty spaces). The heads of the other ends are
// recomputed from vol and level, so that one pass over the nodes suffices. Four lanes (a quad) per node: lane j
// counts the node's incidence entries j, j + 4, ... (integer counts: their sum is the same in any order).
__device__ __forceinline__ void node_heads(PC& P, const Blk& B, int n, int j) {
  const int s = P.M.incStart[n], e = s + P.M.incCount[n];
  int d = 0;
  for (int k0 = s + j; k0 < e; k0 += 4 * kFetch) {
    i4v q[kFetch];
#pragma unroll
    for (int u = 0; u < kFetch; ++u) q[u] = P.M.inc[k0 + 4 * u < e ? k0 + 4 * u : k0];
#pragma unroll
    for (int u = 0; u < kFetch; ++u) {
      if (k0 + 4 * u >= e || !B.cen[q[u].x]) continue;
      const int a = q[u].y, b = q[u].z;
      if (a >= 0 && head_of(B, a) < kNoHeadTest && head_of(B, b) < kNoHeadTest) continue;
      ++d;
    }
  }
  // the quad's four counts (quad_perm 1,0,3,2 then 2,3,0,1): every lane ends with the node's degree
  d += __builtin_amdgcn_update_dpp(0, d, 0xB1, 0xF, 0xF, false);
  d += __builtin_amdgcn_update_dpp(0, d, 0x4E, 0xF, 0xF, false);
  if (j == 0) {
    const double v = B.vol[n];
    B.hEff[n] = v > kEmptyVolume ? B.level[n] : kNoHead;
    B.aEff[n] = v >= B.nVmax[n] ? B.nAov[n] : jmax(B.area[n], B.nAmin[n]);
    B.deg[n] = d;
  }
  if (j == 1) finish_centroid(B, n);
}

// The Villemonte factor of a submerged weir, (1 - r^1.5)^0.385 with the base clipped at 0 (flows.hpp:59): both
// powers are fdlibm's, inlined into one call with their constant exponents folded in (same arithmetic, fewer
// instructions and one call per weir instead of two).
__device__ __noinline__ double villemonte(double r) {
  return fdl::pow(jmax(0.0, 1 - fdl::pow(r, kWeirExponent)), kVillemonteExponent);
}
// the two powers one at a time (fdlibm's, exponent folded in), for callers that also need r^1.5
__device__ __noinline__ double pow_weir(double r) { return fdl::pow(r, kWeirExponent); }
__device__ __noinline__ double pow_villemonte(double g) { return fdl::pow(g, kVillemonteExponent); }

// flows.hpp:38-78: the volume moved through connection c this step (a -> b positive), after the limiter. carea
// holds the instance's products coef * area, the first multiplication of both flow laws, formed on the host.
__device__ inline double connection_flow(PC& P, const SSL double* fr, double t, const Blk& B, int c) {
  const i4v ci = P.M.conI[c];
  const d4v cd = P.M.conD[c];  // x, y, z (and the coefficient, already in carea)
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
    if (d != 0) q = B.carea[c] * sq2g * d / __builtin_sqrt(jmax(__builtin_fabs(d), kOrificeHeadFloor));
  } else {
    const double H1 = Ha > Hb ? Ha : Hb;
    if (H1 > 0) {
      const double H2 = Ha > Hb ? Hb : Ha;
      double f = 1;
      if (H2 > 0) {
        const double r = H2 / H1;
        f = villemonte(r);
      }
      q = B.carea[c] * kWeirCoefficient * sq2g * H1 * __builtin_sqrt(H1) * f * (Ha >= Hb ? 1.0 : -1.0);
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

// flows.hpp:81-111 for node n (the gather in increasing connection order and the capped volume update), then
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
  const double vm = B.nVmax[n];
  double excess = 0;
  if (v > vm) {
    const double head = jmax(0.0, -B.zmax[n]);
    const double cap = vm + B.nAov[n] * head;
    if (v > cap) { excess = v - cap; v = cap; }
  }
  // The excess is subtracted from the sea inflow in node order (flows.hpp:114-118) and the loads take every node
  // with !(vol <= 0) (body.hpp:21-34). The tail adds these terms in node order without tests: a skipped node
  // contributes -0.0 (x + (-0.0) == x for every x) and a subtracted term is stored negated (x - y == x + (-y)).
  B.excess[n] = -excess;
  B.vol[n] = v;
  double tx = -0.0, ty = -0.0, tf = -0.0, tm = -0.0, tip = -0.0, tir = -0.0;
  if (!(v <= 0)) {
    const SSL double* fr = B.frame;
    const double W = P.M.rho * P.M.g;
    const double x = B.cx[n], y = B.cy[n], z = B.cz[n];
    const double wx = fr[F_R00] * x + fr[F_R01] * y + fr[F_R02] * z + fr[F_TX] - P.M.XO;
    const double wy = fr[F_R10] * x + fr[F_R11] * y + fr[F_R12] * z + fr[F_TY] - P.M.YO;
    const double wz = fr[F_R20] * x + fr[F_R21] * y + fr[F_R22] * z + fr[F_TZ] - B.scal[S_ZO];
    const double f = -W * v;
    const double m = P.M.rho * v;
    tx = wy * f;
    ty = -(wx * f);
    tf = f;
    tm = m;
    tip = m * (wx * wx + wz * wz);
    tir = m * (wy * wy + wz * wz);
  }
  B.tX[n] = tx; B.tY[n] = ty; B.tF[n] = tf; B.tM[n] = tm; B.tIp[n] = tip; B.tIr[n] = tir;
}

// ------------------------------------------------------------------------------------------- readouts

// readouts.hpp:20-36 (7 values: t trim list water(t) inflow(t/min) draftF draftA), one thread. A curve record is padded
// to 8 doubles and written as four 16-byte stores into its own 64-byte line (padded = 1; r 64-byte aligned); the final
// readouts go out as 7 values. It takes the LDS arrays it reads rather than the block's view, which a call would have
// to copy to scratch.
__device__ __noinline__ void readouts7(PC& P, const SSL double* scal, const SSL double* vol, SSG double* r, int padded) {
  const double t = scal[S_T], zO = scal[S_ZO], pitch = scal[S_PITCH], roll = scal[S_ROLL];
  const double cth = trig(pitch, 1), sth = trig(pitch, 0), cph = trig(roll, 1), sph = trig(roll, 0);
  const double R20 = -sth, R21 = cth * sph, R22 = cth * cph;
  const double tz = zO - (R20 * P.M.g0x + R21 * P.M.g0y + R22 * P.M.g0z);
  double water = 0;
  for (int n = 0; n < P.S.nn; ++n) water += vol[n];
  const double v[7] = {t, pitch * 180 / kPi, roll * 180 / kPi, water * P.M.rho / 1000, scal[O_INFLOW] * P.M.rho / 1000 * 60,
                       -(R20 * P.M.xFP + R21 * 0.0 + R22 * 0.0 + tz), -(R20 * P.M.xAP + R21 * 0.0 + R22 * 0.0 + tz)};
  if (padded) {
    SSG d2v* q = reinterpret_cast<SSG d2v*>(r);
    q[0] = d2v{v[0], v[1]}; q[1] = d2v{v[2], v[3]}; q[2] = d2v{v[4], v[5]}; q[3] = d2v{v[6], 0.0};
  } else {
    for (int k = 0; k < 7; ++k) r[k] = v[k];
  }
}

__device__ inline void push_event(PC& P, const Blk& B, double t, int id, int kind) {
  const int k = static_cast<int>(B.scal[S_EVCOUNT]);
  if (k < P.M.evCap) {
    B.evT[k] = t; B.evId[k] = id; B.evKind[k] = static_cast<uint8_t>(kind);
    B.scal[S_EVCOUNT] = k + 1;
  }
}

// ------------------------------------------------------------------------------------------- the step

// Buoyancy partials of this thread's 32-lane group; lanes 0 and 16 store the group's tree results.
template <bool Mixed>
__device__ __forceinline__ void buoyancy_partials(PC& P, const Blk& B, int tid) {
  const int lane = tid & (kLanes - 1), unit = tid / kLanes;
  SSL double* w = B.warpHyd + 5 * unit;
  if constexpr (Mixed) {
    const Fr32 f = fr32(B.frame);
    HP<float> p{0.0f, 0.0f, 0.0f, 0.0f, -1e9f};
    stream<float>(P.S.colF, rec_bytes<float>(P), tid, P.S.ncol, kThreads, [&](const Geo<float>& g, bool on) { column(f, g, on, p); });
    float vx, yz, top;
    tree32_packed(p, vx, yz, top);
    if (lane == 0) { w[0] = vx; w[2] = yz; w[4] = top; }
    if (lane == 16) { w[1] = vx; w[3] = yz; }
  } else {
    const Fr64 F = fr64(B.frame);
    HP<double> p{0.0, 0.0, 0.0, 0.0, -1e9};
    stream<double>(P.S.col, rec_bytes<double>(P), tid, P.S.ncol, kThreads, [&](const Geo<double>& g, bool on) { column(F, g, on, p); });
    double vx, yz, top;
    tree32_packed(p, vx, yz, top);
    if (lane == 0) { w[0] = vx; w[2] = yz; w[4] = top; }
    if (lane == 16) { w[1] = vx; w[3] = yz; }
  }
}

// the 8 group results combined in order, then finish_hydro (hydro.hpp:37-45, 61-63)
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

// The tail of a step on wave 0 (flows.hpp:114-118, step.hpp:48-53, body.hpp, events.hpp). Lanes 0..6 run the
// seven serial node-order chains (tauX, tauY, Fz, mw, Iw_p, Iw_r, and the sea inflow less the excess) over terms
// node_update stored gated and signed, eight nodes' values fetched at a time; every lane then integrates the same uniform values, the lanes test the monitors
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
    // node_update stored each chain's terms gated and signed: the chains only add
    const SSL double* src = lane == 0 ? B.tX : lane == 1 ? B.tY : lane == 2 ? B.tF : lane == 3 ? B.tM : lane == 4 ? B.tIp : lane == 5 ? B.tIr : B.excess;
    acc = lane == 0 ? (byw - P.M.YO) * Fb : lane == 1 ? -(bxw - P.M.XO) * Fb : lane == 2 ? Fb - P.M.g * P.M.ms : lane == 6 ? B.scal[S_SEA] : 0.0;
    // whole batches of eight nodes, each batch's terms read while the previous batch is added (the next batch's
    // start clamped in bounds: the last batch is read twice and added once), then the remaining nodes one by one
    const int full = NN & ~7;
    if (full > 0) {
      double a[8], b[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] = src[k];
      for (int n0 = 0;;) {
        const int nb = min(n0 + 8, full - 8);
#pragma unroll
        for (int k = 0; k < 8; ++k) b[k] = src[nb + k];
#pragma unroll
        for (int k = 0; k < 8; ++k) acc += a[k];
        n0 += 8;
        if (n0 >= full) break;
        const int na = min(n0 + 8, full - 8);
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = src[na + k];
#pragma unroll
        for (int k = 0; k < 8; ++k) acc += b[k];
        n0 += 8;
        if (n0 >= full) break;
      }
    }
    for (int n = full; n < NN; ++n) acc += src[n];
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
    *B.nextGroup = 0;  // for the next step's solve; read after the closing barrier
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

// The model's address made opaque to the compiler: the fields a phase reads are then loaded (scalar loads, from the
// constant cache) within that phase instead of being held in scalar registers across the whole step loop, where
// all of them together exceed the register file and spill.
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(v))));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(v >> 32))));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ PC& model_at(uint64_t& pa) {
  asm volatile("" : "+s"(pa));
  return *(PC*)pa;
}

// One step of the model at address pa. On entry the frame slots hold the pose of the current state (written by the
// previous step's tail or at load time); every thread may read them. Ends with a barrier.
// TPB is the workgroup's size: 256, or 512, where the second 256 threads start the level solve while the first
// 256 compute the buoyancy partials (both depend only on the pose and the state), then join it.
template <bool Mixed, int TPB>
__device__ __forceinline__ void device_step(uint64_t pa, SSL double* lds, int tid) {
  // (a) buoyancy partials (threads 0..255), (b) the free surfaces (every unit, taking groups as it is free)
  {
    PC& P = model_at(pa);
    if (TPB == kThreads || tid < kThreads) buoyancy_partials<Mixed>(P, carve(lds, P), tid);
  }
  {
    PC& P = model_at(pa);
    solve_levels<Mixed>(P, carve(lds, P), tid);
  }
  __syncthreads();

  // (c) heads and degrees per node
  {
    PC& P = model_at(pa);
    const Blk B = carve(lds, P);
    for (int n = tid >> 2; n < P.S.nn; n += TPB >> 2) node_heads(P, B, n, tid & 3);
  }
  __syncthreads();

  // (c) flows per connection
  // wave 0 carries the most work from here to the end of the step (the node gathers, the tail while the other waves
  // wait): it issues at the highest priority, ahead of the other instances' waves on its SIMD (solving waves included)
  if (tid < 64) __builtin_amdgcn_s_setprio(3);
  {
    PC& P = model_at(pa);
    const Blk B = carve(lds, P);
    const double t = B.scal[S_T];
    // slots in flowOrder: a wave's lanes mostly follow one flow law, and the submerged weirs' powers run in the
    // first waves only (each connection's flow depends on itself alone, so the order changes no value). Slots beyond
    // the workgroup's size go to its threads in reverse, the last wave's first: that wave's slots are orifices
    for (int r = 0; r * TPB < P.M.nc; ++r) {
      const int k = r * TPB + ((r & 1) ? TPB - 1 - tid : tid);
      if (k >= P.M.nc) continue;
      const int c = P.M.flowOrder[k];
      const double dv = connection_flow(P, B.frame, t, B, c);
      B.dV[c] = dv;
      const int a = P.M.conI[c].x;
      if (a < 0) B.seaDV[-1 - a] = dv;  // a sea connection's volume also at its place in the sea list
    }
  }
  __syncthreads();

  // (c) gathers, volume updates and load terms per node (the first waves); monitor sums (events.hpp:17-19, per monitor
  // in increasing connection order) on the threads after the nodes' waves; the sea inflow (flows.hpp:92-96) on the
  // last thread
  {
    PC& P = model_at(pa);
    const Blk B = carve(lds, P);
    for (int n = tid; n < P.S.nn; n += TPB) node_update(P, B, n);
    const int mon0 = (P.S.nn + 63) / 64 * 64 % TPB;
    for (int m = (tid - mon0 + TPB) % TPB; m < P.M.nmon; m += TPB) {
      double o = 0;
      const int s = P.M.monStart[m], e = s + P.M.monCount[m];
      for (int k0 = s; k0 < e; k0 += kFetch) {
        int c[kFetch];
        double q[kFetch];
#pragma unroll
        for (int u = 0; u < kFetch; ++u) c[u] = P.M.monConn[k0 + u < e ? k0 + u : k0];
#pragma unroll
        for (int u = 0; u < kFetch; ++u) q[u] = B.q[c[u]];
#pragma unroll
        for (int u = 0; u < kFetch; ++u)
          if (k0 + u < e && q[u] > 0) o += q[u];
      }
      B.over[m] = o;
    }
    if (tid == TPB - 1) {
      double s = 0;
      for (int k = 0; k < P.M.seaCount; ++k) s += B.seaDV[k];  // the sea connections in increasing order
      B.scal[S_SEA] = s;
    }
  }
  __syncthreads();

  {
    PC& P = model_at(pa);  // outside the branch: the address stays uniform
    // wave 0 alone runs the tail, still at the highest priority, then returns to the default one
    if (tid < 64) {
      step_tail(P, carve(lds, P), tid);
      __builtin_amdgcn_s_setprio(0);
    }
  }
  __syncthreads();
}

}  // namespace ssa
