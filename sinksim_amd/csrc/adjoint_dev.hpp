// The reverse-mode adjoint of the flooding step on gfx950 (no counterpart in the reference).
//
// One step is differentiated in two halves on the simulation's workgroup:
//
//   recompute  (jac_step) the step from its stored start state, value for value the forward step (the values take
//              the same branches, so the derivative is that of the same discrete program), and alongside it the
//              step's local Jacobians: the column and segment integrals' first moments over their partly wet
//              columns and segments (the derivative of a clipped integral), each free-surface group's outputs
//              (level, volume share, free-surface area, centroid, extent top) as tangents along five local
//              degrees of freedom -- its volume V, its starting level h0 and the pose (zO, pitch, roll) -- carried
//              one per lane of the group's 32-lane unit through the Newton iteration, and each connection's
//              partial derivatives of its volume moved;
//   reverse    (reverse_step) the vector-Jacobian product: the adjoint of the state after the step, plus the loss's
//              seeds at this step, pulled back through integration, loads, volume updates, flows and the level
//              solve to the adjoint of the state before it and to the parameters' log-derivatives.
//
// Conventions (identical in the CPU dual-number build, oracle/dual_step.hpp, which the tests hold this to):
// sin/cos differentiate to the fdlibm cos/-sin of the same argument; pow(x, y) to y * pow(x, y) / x (0 at x = 0);
// sqrt to 0.5 / sqrt(x) (0 at 0); a selection or clamp to its selected operand. The relaxed level derivative
// replaces a level's tangent by that of the root of vol(h) = V at the final level (one Newton step on the tangent,
// with the slope floored at the space's minimum effective area amin), as if the solve had converged.
#pragma once

#include <type_traits>

#include "step_dev.hpp"

namespace ssa {

constexpr int kDof = 5;     // local degrees of freedom of a level solve: V, h0, zO, pitch, roll
constexpr int kSolveOut = 7;  // per space: level, volume (after group sharing), free-surface area, cx, cy, cz, extent top
constexpr int kPassRec = 23;  // a node pass's reduced record: 7 values, 10 moment sums, 6 extent coordinates
constexpr int kFrameDeriv = 13;  // frame slots F_R00..F_CTH with derivatives along the pose
constexpr int kBw = 18;          // a 32-lane group's buoyancy record: V Mx My Mz top, 10 moment sums, top column x y top
constexpr int kMaxParams = 64;   // parameters of a gradient call at most

// LDS layout of the adjoint's step tape, offsets in doubles from the end of the forward layout (the host computes
// it: adj_layout in adjoint.hip).
struct AdjLayout {
  int32_t sv, iVol, iLvl, iCx, iCy, iCz, fr, dfp, bw, bt;
  int32_t oLvl, oVol, oArea, oCx, oCy, oCz, oZmax, jac, carry, gMer, gH0, mp;
  int32_t hEff, aEff, aSrc, deg, dV, cp, acc, vol1, clip, tot;
  int32_t lamA, lamB, G, gOut, gVol1, gAcc, cg, red, seeds, nseeds, sig;
  int32_t cpa, cpc, cmon, monSeed, sflag, gnext;
  int32_t total;
  int32_t gJac, gCp, gCg, gTotal;  // the global slice (doubles per instance; 0 when everything is in LDS)
};

// The adjoint's per-model record beside ModelParams.
struct AdjModel {
  AdjLayout A;
  int32_t maxGroup;
  int32_t np;
  int32_t mode;  // 0 exact, 1 relaxed
};
// The three values of AdjModel the step's functions read, passed by value (a local AdjModel copy passed by
// reference lived in scratch memory)
struct AdjScalars {
  int32_t maxGroup, np, mode;
};

// ------------------------------------------------------------------------------------------- tangent values

// A value (the same in every lane of a unit) with this lane's tangent component.
struct TV {
  double v, d;
};
__device__ __forceinline__ TV tc(double v) { return TV{v, 0.0}; }
__device__ __forceinline__ TV operator+(TV a, TV b) { return TV{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ TV operator-(TV a, TV b) { return TV{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ TV operator-(TV a) { return TV{-a.v, -a.d}; }
__device__ __forceinline__ TV operator*(TV a, TV b) { return TV{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ TV operator/(TV a, TV b) {
  const double r = a.v / b.v;
  return TV{r, (a.d - r * b.d) / b.v};
}
__device__ __forceinline__ TV operator+(TV a, double b) { return TV{a.v + b, a.d}; }
__device__ __forceinline__ TV operator+(double a, TV b) { return TV{a + b.v, b.d}; }
__device__ __forceinline__ TV operator-(TV a, double b) { return TV{a.v - b, a.d}; }
__device__ __forceinline__ TV operator-(double a, TV b) { return TV{a - b.v, -b.d}; }
__device__ __forceinline__ TV operator*(TV a, double b) { return TV{a.v * b, a.d * b}; }
__device__ __forceinline__ TV operator*(double a, TV b) { return TV{a * b.v, a * b.d}; }
__device__ __forceinline__ TV operator/(TV a, double b) { return TV{a.v / b, a.d / b}; }
__device__ __forceinline__ TV tmaxv(TV a, TV b) { return (b.v > a.v) ? b : a; }
__device__ __forceinline__ TV tminv(TV a, TV b) { return (b.v < a.v) ? b : a; }
__device__ __forceinline__ TV tfabs(TV a) { return a.v < 0 ? -a : a; }

// ------------------------------------------------------------------------------------------- the tape

// The tape's arrays. The three largest -- the level solve's Jacobian rows (35 per space), the connections' partial
// derivatives (6 each) and their gradient terms (5 each) -- are in LDS (GT = false) or, for models whose tape does not
// fit beside the forward layout, in the instance's slice of a global buffer (GT = true); the rest is always in LDS.
template <bool GT>
struct TapeT {
  using XP = std::conditional_t<GT, SSG double*, SSL double*>;
  SSL double *sv, *iVol, *iLvl, *iCx, *iCy, *iCz, *fr, *dfp, *bw, *bt;
  SSL double *oLvl, *oVol, *oArea, *oCx, *oCy, *oCz, *oZmax, *carry, *gMer, *gH0, *mp;
  SSL double *hEff, *aEff, *aSrc, *deg, *dV, *acc, *vol1, *clip, *tot;
  SSL double *lamA, *lamB, *G, *gOut, *gVol1, *gAcc, *red, *seeds, *nseeds, *sig;
  SSL double *cpa, *cpc, *cmon, *monSeed, *sflag;
  SSL double* gnext;  // the Jacobian solve's group counter (an int32)
  XP jac, cp, cg;
};

// base: the tape's LDS; gbase: the instance's global slice (GT only; offsets gJac, gCp, gCg)
template <bool GT>
__device__ inline TapeT<GT> carve_tape(SSL double* base, SSG double* gbase, const SSC AdjLayout& L) {
  TapeT<GT> t;
  t.sv = base + L.sv; t.iVol = base + L.iVol; t.iLvl = base + L.iLvl; t.iCx = base + L.iCx; t.iCy = base + L.iCy; t.iCz = base + L.iCz;
  t.fr = base + L.fr; t.dfp = base + L.dfp; t.bw = base + L.bw; t.bt = base + L.bt;
  t.oLvl = base + L.oLvl; t.oVol = base + L.oVol; t.oArea = base + L.oArea; t.oCx = base + L.oCx; t.oCy = base + L.oCy;
  t.oCz = base + L.oCz; t.oZmax = base + L.oZmax; t.carry = base + L.carry; t.gMer = base + L.gMer;
  t.gH0 = base + L.gH0; t.mp = base + L.mp;
  t.hEff = base + L.hEff; t.aEff = base + L.aEff; t.aSrc = base + L.aSrc; t.deg = base + L.deg; t.dV = base + L.dV;
  t.acc = base + L.acc; t.vol1 = base + L.vol1; t.clip = base + L.clip; t.tot = base + L.tot;
  t.lamA = base + L.lamA; t.lamB = base + L.lamB; t.G = base + L.G; t.gOut = base + L.gOut; t.gVol1 = base + L.gVol1;
  t.gAcc = base + L.gAcc; t.red = base + L.red; t.seeds = base + L.seeds; t.nseeds = base + L.nseeds;
  t.sig = base + L.sig;
  t.cpa = base + L.cpa; t.cpc = base + L.cpc; t.cmon = base + L.cmon; t.monSeed = base + L.monSeed; t.sflag = base + L.sflag;
  t.gnext = base + L.gnext;
  if constexpr (GT) {
    t.jac = gbase + L.gJac; t.cp = gbase + L.gCp; t.cg = gbase + L.gCg;
  } else {
    (void)gbase;
    t.jac = base + L.jac; t.cp = base + L.cp; t.cg = base + L.cg;
  }
  return t;
}

// slots of `tot` (one step's scalars)
enum : int { T_V = 0, T_MX, T_MY, T_MZ, T_IR, T_BX, T_BY, T_BZ, T_FB, T_BXW, T_BYW, T_TAUX, T_TAUY, T_FZ, T_MW, T_IWP, T_IWR,
             T_MH, T_ITH, T_IPH, T_AZ, T_ATH, T_APH, T_CLP, T_CLR, T_VZ1, T_WTH1, T_WPH1, T_ZO1, T_P1, T_R1, T_GFZ, T_GTX, T_GTY,
             T_GMW, T_GIWP, T_GIWR, T_SEA, T_COUNT = 48 };
// moment sums of `bt` (buoyancy) after the 4 values: T00 T01 T02 T03 T11 T12 T13 T22 T23 T33
// moment sums are followed by the hull top's column (x, y, top) at bt[14..16]
// seeds: records of 4 doubles (step, kind, index, value), sorted by step, latest first, and consumed in that order by
// the reverse sweep (sflag[0]: the next record; sflag[1]: whether this step has monitor seeds, in monSeed); kinds:
// 0 pitch after the step, 1 a monitor's flow sum, 2 a mark's world z, 3 roll after the step, 4 the step's hull top,
// 5 the step's buoyancy volume. cpa, cpc, cmon: per connection its parameter indices and its monitor (-1: none).

// ------------------------------------------------------------------------------------------- recompute

// frame derivatives along the pose (zO, pitch, roll): dfp[slot * 3 + j] for the slots F_R00..F_CTH
__device__ inline void write_frame_deriv(PC& P, SSL double* dfp, double cth, double sth, double cph, double sph) {
  double d[kFrameDeriv][3] = {};
  d[F_R00][1] = -sth;
  d[F_R01][1] = cth * sph; d[F_R01][2] = sth * cph;
  d[F_R02][1] = cth * cph; d[F_R02][2] = -(sth * sph);
  d[F_R11][2] = -sph;
  d[F_R12][2] = -cph;
  d[F_R20][1] = -cth;
  d[F_R21][1] = -(sth * sph); d[F_R21][2] = cth * cph;
  d[F_R22][1] = -(sth * cph); d[F_R22][2] = -(cth * sph);
  d[F_CTH][1] = -sth;
  for (int j = 0; j < 3; ++j) {
    d[F_TX][j] = -(d[F_R00][j] * P.M.g0x + d[F_R01][j] * P.M.g0y + d[F_R02][j] * P.M.g0z);
    d[F_TY][j] = -(d[F_R10][j] * P.M.g0x + d[F_R11][j] * P.M.g0y + d[F_R12][j] * P.M.g0z);
    d[F_TZ][j] = (j == 0 ? 1.0 : 0.0) - (d[F_R20][j] * P.M.g0x + d[F_R21][j] * P.M.g0y + d[F_R22][j] * P.M.g0z);
  }
  for (int e = 0; e < kFrameDeriv; ++e)
    for (int j = 0; j < 3; ++j) dfp[e * 3 + j] = d[e][j];
}

// A node pass's reduced record (values finished as finish_pass does, the moment sums over the partly wet segments
// of 1, x, y, zp against each other, and the coordinates of the segments that give the extent).
struct PassRec {
  double vol, dv, mx, my, mz, zmin, zmax;
  double S[10];  // S00 S01 S02 S03 S11 S12 S13 S22 S23 S33 over partial segments, weighted by the footprint
  double xmn, ymn, amn, xmx, ymx, bmx;
};

// A minimum tree that carries three coordinates with the extreme value: `if (other < own)` the lane takes the
// other's value and coordinates (ties and NaN keep its own).
template <int OFF>
__device__ __forceinline__ void argmin_level(double& z, double& u, double& v, double& w) {
  const double oz = down<OFF>(z), ou = down<OFF>(u), ov = down<OFF>(v), ow = down<OFF>(w);
  if (oz < z) { z = oz; u = ou; v = ov; w = ow; }
}
// The pass of node n at level h by a 32-lane unit (hydro.hpp:112-122 values, in the canonical tree), with its
// moment sums; the record is in every lane of the unit.
__device__ inline PassRec pass_rec(PC& P, const Blk& B, const SSL double* fr, int lane, int n, double h) {
  const double R20 = fr[F_R20], R21 = fr[F_R21], R22 = fr[F_R22], tz = fr[F_TZ], iR = fr[F_IR];
  PassRec p;
  p.vol = 0; p.dv = 0; p.mx = 0; p.my = 0; p.mz = 0; p.zmin = 1e9; p.zmax = -1e9;
  p.xmn = 0; p.ymn = 0; p.amn = 0; p.xmx = 0; p.ymx = 0; p.bmx = 0;
#pragma unroll
  for (int k = 0; k < 10; ++k) p.S[k] = 0;
  const int s0 = B.nSegStart[n], s1 = s0 + B.nSegCount[n];
  stream<double>(P.S.seg, rec_bytes<double>(P), s0 + lane, s1, kLanes, [&](const Geo<double>& g, bool on) {
    const double x = g.x, y = g.y, a = g.lo, b = g.hi;
    const double base = R20 * x + R21 * y + tz;
    const double wa = base + R22 * a, wb = base + R22 * b;
    if (on && wa < p.zmin) { p.zmin = wa; p.xmn = x; p.ymn = y; p.amn = a; }
    if (on && wb > p.zmax) { p.zmax = wb; p.xmx = x; p.ymx = y; p.bmx = b; }
    const double zp = (h - base) * iR;
    if (!on || zp <= a) return;
    const double A = g.A;
    double L;
    if (zp >= b) {
      L = b - a;
    } else {
      L = zp - a;
      p.dv += A;
      const double Ax = A * x, Ay = A * y, Az = A * zp;
      p.S[0] += A; p.S[1] += Ax; p.S[2] += Ay; p.S[3] += Az;
      p.S[4] += Ax * x; p.S[5] += Ax * y; p.S[6] += Ax * zp; p.S[7] += Ay * y; p.S[8] += Ay * zp; p.S[9] += Az * zp;
    }
    const double v = A * L;
    p.vol += v; p.mx += v * x; p.my += v * y; p.mz += v * (a + 0.5 * L);
  });
  // The halving tree (l absorbs l + 16, 8, 4, 2, 1) of the 15 sums and of the extent with its coordinates, two
  // values packed per register (pack_sum16 / pack_min16): the sums are the forward pass's to the bit, and the
  // extent's maximum runs as a minimum over -zmax (-max(-a, -b) is max(a, b), ties and NaN keeping the own lane).
  double q[8];
  q[0] = pack_sum16(p.vol, p.dv); q[1] = pack_sum16(p.mx, p.my); q[2] = pack_sum16(p.mz, p.S[0]);
  q[3] = pack_sum16(p.S[1], p.S[2]); q[4] = pack_sum16(p.S[3], p.S[4]); q[5] = pack_sum16(p.S[5], p.S[6]);
  q[6] = pack_sum16(p.S[7], p.S[8]); q[7] = p.S[9];
  sum_level<16>(q[7]);
  double ez = -p.zmax, ex = p.xmx, ey = p.ymx, ea = p.bmx;
  double mz = p.zmin, mxx = p.xmn, myy = p.ymn, maa = p.amn;
  swap16(mz, ez); swap16(mxx, ex); swap16(myy, ey); swap16(maa, ea);
  if (ez < mz) { mz = ez; mxx = ex; myy = ey; maa = ea; }  // own is the first after the swap
#pragma unroll
  for (int k = 0; k < 8; ++k) sum_level<8>(q[k]);
  argmin_level<8>(mz, mxx, myy, maa);
#pragma unroll
  for (int k = 0; k < 8; ++k) sum_level<4>(q[k]);
  argmin_level<4>(mz, mxx, myy, maa);
#pragma unroll
  for (int k = 0; k < 8; ++k) sum_level<2>(q[k]);
  argmin_level<2>(mz, mxx, myy, maa);
#pragma unroll
  for (int k = 0; k < 8; ++k) sum_level<1>(q[k]);
  argmin_level<1>(mz, mxx, myy, maa);
  PassRec r;
  r.vol = lead(q[0]); r.dv = lead16(q[0]); r.mx = lead(q[1]); r.my = lead16(q[1]); r.mz = lead(q[2]);
  r.S[0] = lead16(q[2]); r.S[1] = lead(q[3]); r.S[2] = lead16(q[3]); r.S[3] = lead(q[4]); r.S[4] = lead16(q[4]);
  r.S[5] = lead(q[5]); r.S[6] = lead16(q[5]); r.S[7] = lead(q[6]); r.S[8] = lead16(q[6]); r.S[9] = lead(q[7]);
  r.zmin = lead(mz); r.xmn = lead(mxx); r.ymn = lead(myy); r.amn = lead(maa);
  r.zmax = -lead16(mz); r.xmx = lead16(mxx); r.ymx = lead16(myy); r.bmx = lead16(maa);
  const double mu = B.nMu[n];
  r.vol = r.vol * mu;
  r.dv = r.dv * mu * iR;
  r.mx = r.mx * mu; r.my = r.my * mu; r.mz = r.mz * mu;
  return r;
}

// this lane's frame tangents (R20, R21, R22, tz) along its degree of freedom (lanes 2, 3, 4: zO, pitch, roll)
struct FrameTan {
  double R20, R21, R22, tz;
};
__device__ __forceinline__ FrameTan frame_tan(const SSL double* dfp, int lane) {
  if (lane < 2 || lane >= kDof) return FrameTan{0, 0, 0, 0};
  const int j = lane - 2;
  return FrameTan{dfp[F_R20 * 3 + j], dfp[F_R21 * 3 + j], dfp[F_R22 * 3 + j], dfp[F_TZ * 3 + j]};
}

struct PassTV {
  TV vol, dv, mx, my, mz, zmin, zmax;
};

// The pass values with this lane's tangents for a level tangent dh: with dzp = iR (dh - dtz - x dR20 - y dR21 -
// zp dR22) on every partly wet segment, d(vol, mx, my, mz) = mu iR S (dh - dtz, -dR20, -dR21, -dR22); the slope
// dv = mu iR sum A changes only through iR; the extents move with their segments.
__device__ __forceinline__ PassTV pass_tv(const PassRec& r, double dh, const FrameTan& f, double mu, double iR) {
  const double k = mu * iR, w = dh - f.tz;
  const double* S = r.S;
  PassTV t;
  t.vol = TV{r.vol, k * (S[0] * w - S[1] * f.R20 - S[2] * f.R21 - S[3] * f.R22)};
  t.mx = TV{r.mx, k * (S[1] * w - S[4] * f.R20 - S[5] * f.R21 - S[6] * f.R22)};
  t.my = TV{r.my, k * (S[2] * w - S[5] * f.R20 - S[7] * f.R21 - S[8] * f.R22)};
  t.mz = TV{r.mz, k * (S[3] * w - S[6] * f.R20 - S[8] * f.R21 - S[9] * f.R22)};
  t.dv = TV{r.dv, -(r.dv * iR * f.R22)};
  t.zmin = TV{r.zmin, r.xmn * f.R20 + r.ymn * f.R21 + r.amn * f.R22 + f.tz};
  t.zmax = TV{r.zmax, r.xmx * f.R20 + r.ymx * f.R21 + r.bmx * f.R22 + f.tz};
  return t;
}

__device__ __forceinline__ void put_rec(SSL double* m, const PassRec& r) {
  m[0] = r.vol; m[1] = r.dv; m[2] = r.mx; m[3] = r.my; m[4] = r.mz; m[5] = r.zmin; m[6] = r.zmax;
  for (int k = 0; k < 10; ++k) m[7 + k] = r.S[k];
  m[17] = r.xmn; m[18] = r.ymn; m[19] = r.amn; m[20] = r.xmx; m[21] = r.ymx; m[22] = r.bmx;
}
__device__ __forceinline__ PassRec get_rec(const SSL double* m) {
  PassRec r;
  r.vol = m[0]; r.dv = m[1]; r.mx = m[2]; r.my = m[3]; r.mz = m[4]; r.zmin = m[5]; r.zmax = m[6];
  for (int k = 0; k < 10; ++k) r.S[k] = m[7 + k];
  r.xmn = m[17]; r.ymn = m[18]; r.amn = m[19]; r.xmx = m[20]; r.ymx = m[21]; r.bmx = m[22];
  return r;
}

// The solve's outputs of space n with this lane's tangent row (lanes 0..4 write their component).
template <bool GT>
__device__ __forceinline__ void put_out(const TapeT<GT>& T, int n, int lane, TV lvl, TV vol, TV area, TV cx, TV cy, TV cz, bool carry, TV zmax) {
  if (lane == 0) {
    T.oLvl[n] = lvl.v; T.oVol[n] = vol.v; T.oArea[n] = area.v; T.oCx[n] = cx.v; T.oCy[n] = cy.v; T.oCz[n] = cz.v; T.oZmax[n] = zmax.v;
    T.carry[n] = carry ? 1.0 : 0.0;
  }
  if (lane < kDof) {
    const auto J = T.jac + n * (kSolveOut * kDof);
    J[0 * kDof + lane] = lvl.d; J[1 * kDof + lane] = vol.d; J[2 * kDof + lane] = area.d;
    J[3 * kDof + lane] = carry ? 0.0 : cx.d; J[4 * kDof + lane] = carry ? 0.0 : cy.d; J[5 * kDof + lane] = carry ? 0.0 : cz.d;
    J[6 * kDof + lane] = zmax.d;
  }
}

// levels.hpp:13-49 (full = false) for space n, with tangents along (V, h0, pose) = (vol[n], level[n], pose).
template <bool GT>
__device__ __forceinline__ void solve_single_tv(PC& P, const Blk& B, const TapeT<GT>& T, int lane, int n, bool relaxed) {
  const SSL double* fr = T.fr;
  const FrameTan f = frame_tan(T.dfp, lane);
  const double mu = B.nMu[n], iR = fr[F_IR];
  const TV V = TV{T.iVol[n], lane == 0 ? 1.0 : 0.0};
  const double Vmax = B.nVmax[n];
  TV h = TV{T.iLvl[n], lane == 1 ? 1.0 : 0.0};
  PassRec r = pass_rec(P, B, fr, lane, n, h.v);
  PassTV p = pass_tv(r, h.d, f, mu, iR);
  const TV lo = p.zmin, hi = p.zmax;
  const TV cx0 = tc(T.iCx[n]), cy0 = tc(T.iCy[n]), cz0 = tc(T.iCz[n]);
  if (V.v <= kEmptyVolume) {
    put_out(T, n, lane, lo, V, tc(0.0), cx0, cy0, cz0, true, hi);
    return;
  }
  if (V.v >= Vmax * kFullFraction) {
    r = pass_rec(P, B, fr, lane, n, hi.v);
    p = pass_tv(r, hi.d, f, mu, iR);
    const TV lvl = hi + tmaxv(tc(0.0), V - Vmax) / B.nAov[n];
    const bool c = !(p.vol.v > kCentroidMinVolume);
    put_out(T, n, lane, lvl, V, tc(B.nAov[n]), c ? cx0 : p.mx / p.vol, c ? cy0 : p.my / p.vol, c ? cz0 : p.mz / p.vol, c, hi);
    return;
  }
  if (h.v < lo.v || h.v > hi.v) {
    h = tminv(hi, tmaxv(lo, h));
    r = pass_rec(P, B, fr, lane, n, h.v);
    p = pass_tv(r, h.d, f, mu, iR);
  }
  const double tol = jmax(1e-3 * Vmax, 0.5);
  TV a = lo, b = hi;
  for (int it = 0; it < kNewtonFast; ++it) {
    const TV err = p.vol - V;
    if (__builtin_fabs(err.v) < tol) break;
    if (err.v > 0) b = h; else a = h;
    TV hn = p.dv.v > kSlopeMin ? h - err / p.dv : 0.5 * (a + b);
    if (!(hn.v > a.v && hn.v < b.v)) hn = 0.5 * (a + b);
    h = hn;
    r = pass_rec(P, B, fr, lane, n, h.v);
    p = pass_tv(r, h.d, f, mu, iR);
  }
  if (relaxed) {
    const double as = jmax(p.dv.v, B.nAmin[n]);
    if (as > kSlopeMin) {
      h.d += (V.d - p.vol.d) / as;
      p = pass_tv(r, h.d, f, mu, iR);
    }
  }
  const bool c = !(p.vol.v > kCentroidMinVolume);
  put_out(T, n, lane, h, V, p.dv, c ? cx0 : p.mx / p.vol, c ? cy0 : p.my / p.vol, c ? cz0 : p.mz / p.vol, c, hi);
}

// levels.hpp:54-162 (full = false) for a merged group, with tangents along (V, h0, pose), V the group's total
// volume and h0 its starting level (the members' mean level when all are wet, else one member's level). The
// members' pass records are kept in the unit's scratch (mp); every member's pass of a round is at the same level.
template <bool GT>
__device__ __forceinline__ void solve_group_tv(PC& P, const Blk& B, const TapeT<GT>& T, SSL double* mp, int lane, int g, int first, int N, bool relaxed) {
  const SSL double* fr = T.fr;
  const FrameTan f = frame_tan(T.dfp, lane);
  const double iR = fr[F_IR];
  const SSL int32_t* members = B.gNodes + first;
  double Vs = 0, Vmax = 0;
  for (int k = 0; k < N; ++k) { Vs += T.iVol[members[k]]; Vmax += B.nVmax[members[k]]; }
  const TV V = TV{Vs, lane == 0 ? 1.0 : 0.0};
  bool allWet = true;
  int firstWet = -1;
  for (int k = 0; k < N; ++k) {
    if (T.iVol[members[k]] > 0) { if (firstWet < 0) firstWet = k; }
    else allWet = false;
  }
  double h0;
  if (allWet) {
    double s = 0;
    for (int k = 0; k < N; ++k) s += T.iLvl[members[k]];
    h0 = s / static_cast<double>(N);
  } else {
    h0 = T.iLvl[members[firstWet >= 0 ? firstWet : N - 1]];
  }
  if (lane == 0) T.gH0[g] = allWet ? -1.0 : static_cast<double>(firstWet >= 0 ? firstWet : N - 1);
  TV h = TV{h0, lane == 1 ? 1.0 : 0.0};
  // member k's pass at the round's level hP
  TV hP = h;
  auto passes = [&](double at) {
    for (int k = 0; k < N; ++k) {
      const PassRec r = pass_rec(P, B, fr, lane, members[k], at);
      put_rec(mp + k * kPassRec, r);  // every lane writes the same record and reads back only after its own write
    }
  };
  auto Pk = [&](int k) -> PassTV { return pass_tv(get_rec(mp + k * kPassRec), hP.d, f, B.nMu[members[k]], iR); };
  passes(h.v);
  // the extents (pose only; the first round's records hold them)
  TV lo = Pk(0).zmin, hi = Pk(0).zmax;
  for (int k = 1; k < N; ++k) {
    const PassTV q = Pk(k);
    if (q.zmin.v < lo.v) lo = q.zmin;
    if (q.zmax.v > hi.v) hi = q.zmax;
  }
  // a member's extent top (zmax) and its tangent do not depend on the level: every round's record holds them, so
  // the outputs read them from the last round's records (no per-member array, which would live in scratch memory)

  if (V.v <= kEmptyVolume) {
    for (int k = 0; k < N; ++k) {
      const int n = members[k];
      const PassTV q = Pk(k);
      put_out(T, n, lane, q.zmin, tc(0.0), tc(0.0), tc(T.iCx[n]), tc(T.iCy[n]), tc(T.iCz[n]), true, q.zmax);
    }
    return;
  }
  if (V.v >= Vmax * kFullFraction) {
    const TV over = tmaxv(tc(0.0), V - Vmax);
    double aSum = 0;
    for (int k = 0; k < N; ++k) aSum += B.nAov[members[k]];
    const TV dh = over / aSum;
    // member k < N - 1 takes its capacity plus its pressed-up share (recomputed at the output), the last the rest
    TV assigned = tc(0.0);
    for (int k = 0; k < N - 1; ++k) assigned = assigned + (B.nVmax[members[k]] + dh * B.nAov[members[k]]);
    hP = hi;
    passes(hi.v);
    for (int k = 0; k < N; ++k) {
      const int n = members[k];
      const PassTV q = Pk(k);
      const bool c = !(q.vol.v > kCentroidMinVolume);
      const TV vk = k < N - 1 ? B.nVmax[n] + dh * B.nAov[n] : V - assigned;
      put_out(T, n, lane, hi + dh, vk, tc(B.nAov[n]), c ? tc(T.iCx[n]) : q.mx / q.vol, c ? tc(T.iCy[n]) : q.my / q.vol,
              c ? tc(T.iCz[n]) : q.mz / q.vol, c, q.zmax);
    }
    return;
  }
  if (!(h.v >= lo.v && h.v <= hi.v)) {
    h = tminv(hi, tmaxv(lo, h));
    hP = h;
    passes(h.v);
  }
  const double tol = jmax(1e-3 * Vmax, 0.5);
  TV a = lo, b = hi;
  for (int it = 0; it < kNewtonFast; ++it) {
    TV sum = tc(0.0);
    for (int k = 0; k < N; ++k) sum = sum + Pk(k).vol;
    const TV err = sum - V;
    if (__builtin_fabs(err.v) < tol) break;
    if (err.v > 0) b = h; else a = h;
    TV dv = tc(0.0);
    for (int k = 0; k < N; ++k) dv = dv + Pk(k).dv;
    TV hn = dv.v > kSlopeMin ? h - err / dv : 0.5 * (a + b);
    if (!(hn.v > a.v && hn.v < b.v)) hn = 0.5 * (a + b);
    h = hn;
    hP = h;
    passes(h.v);
  }
  if (relaxed) {
    TV sum = tc(0.0), dv = tc(0.0);
    double amin = 0;
    for (int k = 0; k < N; ++k) { const PassTV q = Pk(k); sum = sum + q.vol; dv = dv + q.dv; amin += B.nAmin[members[k]]; }
    const double as = jmax(dv.v, amin);
    if (as > kSlopeMin) {
      h.d += (V.d - sum.d) / as;
      hP = h;
    }
  }
  TV tot = tc(0.0);
  for (int k = 0; k < N; ++k) tot = tot + Pk(k).vol;
  // member k's proportional share, capped at its capacity; recomputed (the same operations on the same records)
  // wherever it is needed rather than kept in a per-member array
  auto share = [&](int k) -> TV {
    TV v = tot.v > 0 ? V * Pk(k).vol / tot : V / static_cast<double>(N);
    if (v.v > B.nVmax[members[k]]) v = tc(B.nVmax[members[k]]);
    return v;
  };
  TV assigned = tc(0.0);
  for (int k = 0; k < N - 1; ++k) assigned = assigned + share(k);
  // the cascade: from member N - 2 down, members take their capacity until member kc takes what is left; the
  // members below kc keep their shares (kc = N - 1: no cascade)
  int kc = N - 1;
  TV vc = tc(0.0);
  const double vmaxLast = B.nVmax[members[N - 1]];
  if ((V - assigned).v > vmaxLast) {
    const TV need = V - vmaxLast;
    TV capped = tc(0.0);
    for (int k = N - 2; k >= 0; --k) {
      TV prior = tc(0.0);
      for (int j = 0; j < k; ++j) prior = prior + share(j);
      const TV v = (need - capped) - prior;
      if (v.v > B.nVmax[members[k]] && k > 0) {
        capped = capped + B.nVmax[members[k]];
        continue;
      }
      kc = k;
      vc = v;
      break;
    }
    assigned = tc(0.0);
    for (int k = 0; k < N - 1; ++k) assigned = assigned + (k < kc ? share(k) : k == kc ? vc : tc(B.nVmax[members[k]]));
  }
  for (int k = 0; k < N; ++k) {
    const int n = members[k];
    const PassTV q = Pk(k);
    const bool c = !(q.vol.v > kCentroidMinVolume);
    const TV vk = k == N - 1 ? V - assigned : k < kc ? share(k) : k == kc ? vc : tc(B.nVmax[n]);
    put_out(T, n, lane, h, vk, q.dv, c ? tc(T.iCx[n]) : q.mx / q.vol, c ? tc(T.iCy[n]) : q.my / q.vol, c ? tc(T.iCz[n]) : q.mz / q.vol,
            c, q.zmax);
  }
}

// ------------------------------------------------------------------------------------------- views and calls

__device__ __forceinline__ const SSC AdjModel& adj_model_at(uint64_t ta) {
  asm volatile("" : "+s"(ta));
  return *(const SSC AdjModel*)ta;
}

// The addresses the views of an instance are carved from: its model, its tape layout (AdjModel) and its slice of the
// global tape. Uniform.
struct ViewAddrs {
  uint64_t pa, ta, gs;
};

// The instance's views (the model, the forward layout and the tape: about a hundred LDS addresses), carved from the
// model and tape-layout addresses behind a scalar barrier, so that they are derived where a phase needs them rather
// than held in registers across the kernel or passed through scratch memory to a function kept out of line.
template <bool GT>
struct StepViews {
  PC& P;
  const Blk B;
  const TapeT<GT> T;
  __device__ __forceinline__ StepViews(uint64_t pa, uint64_t ta, SSL double* lds, SSG double* gslice)
      : P(model_at(pa)), B(carve(lds, P)), T(carve_tape<GT>(lds + P.L.total, gslice, adj_model_at(ta).A)) {}
};

// levels.hpp:166-194: unit u of the TPB / 32 units takes groups u, u + TPB / 32, ... (the merge decision is the
// forward's: every group's outputs are the same whichever unit computes them)
template <int TPB, bool GT>
__device__ __forceinline__ void solve_levels_tv(PC& P, const Blk& B, const TapeT<GT>& T, int maxGroup, int tid, bool relaxed) {
  const int lane = tid & (kLanes - 1), unit = tid / kLanes;
  // each unit takes the next group from the counter (its lane 0's LDS atomic, broadcast to the unit), so units that
  // start late (after the buoyancy partials) or draw cheap groups take more
  for (;;) {
    int g = 0;
    if (lane == 0) g = atomicAdd((int*)(SSL int32_t*)T.gnext, 1);
    g = static_cast<int>(lead32(static_cast<uint32_t>(g)));
    if (g >= P.M.ng) break;
    const int N = B.gCount[g], first = B.gStart[g];
    const int mode = B.gMode[g];
    bool merged = false;
    if (N >= 2) {
      if (mode == 1) {
        merged = true;
      } else if (mode == 2) {
        const double zMw = world_z(T.fr, B.gpx[g], B.gpy[g], B.gpz[g]);
        merged = true;
        for (int k = 0; k < N; ++k) {
          const int n = B.gNodes[first + k];
          if (!(T.iVol[n] > kEmptyVolume && T.iLvl[n] > zMw + P.M.gMargin)) { merged = false; break; }
        }
      }
    }
    if (lane == 0) T.gMer[g] = merged ? 1.0 : 0.0;
    if (merged) {
      solve_group_tv(P, B, T, T.mp + unit * maxGroup * kPassRec, lane, g, first, N, relaxed);
    } else {
      if (lane == 0) T.gH0[g] = -2.0;  // members solved singly
      for (int k = 0; k < N; ++k) solve_single_tv(P, B, T, lane, B.gNodes[first + k], relaxed);
    }
  }
}

// buoyancy partials (hydro.hpp:48-64 values) with the moment sums of the partly wet columns and the column that
// gives the hull top (carried through the canonical tree: a later lane replaces only a strictly higher top); each
// 32-lane group's result goes to bw (5 values, 10 sums, the top column's x, y and top)
template <bool GT>
__device__ inline void buoyancy_rec(PC& P, const TapeT<GT>& T, int tid) {
  const int lane = tid & (kLanes - 1), unit = tid / kLanes;
  const SSL double* fr = T.fr;
  const Fr64 F = fr64(fr);
  HP<double> p{0.0, 0.0, 0.0, 0.0, -1e9};
  double S[10] = {};
  double tx = 0, ty = 0, th = 0;
  stream<double>(P.S.col, rec_bytes<double>(P), tid, P.S.ncol, kThreads, [&](const Geo<double>& g, bool on) {
    const double x = g.x, y = g.y, lo = g.lo, hi = g.hi;
    const double base = F.R20 * x + F.R21 * y + F.tz;
    const double wt = base + F.R22 * hi;
    if (on && wt > p.top) { p.top = wt; tx = x; ty = y; th = hi; }
    const double zp = -base * F.iR;
    if (!on || zp <= lo) return;
    double s;
    if (zp >= hi) {
      s = hi - lo;
    } else {
      s = zp - lo;
      const double A = g.A, Ax = A * x, Ay = A * y, Az = A * zp;
      S[0] += A; S[1] += Ax; S[2] += Ay; S[3] += Az;
      S[4] += Ax * x; S[5] += Ax * y; S[6] += Ax * zp; S[7] += Ay * y; S[8] += Ay * zp; S[9] += Az * zp;
    }
    const double v = g.A * s;
    p.V += v; p.Mx += v * x; p.My += v * y; p.Mz += v * (lo + 0.5 * s);
  });
  auto level = [&](auto off) {
    constexpr int OFF = decltype(off)::value;
    const double V = down<OFF>(p.V), Mx = down<OFF>(p.Mx), My = down<OFF>(p.My), Mz = down<OFF>(p.Mz), top = down<OFF>(p.top);
    const double ox = down<OFF>(tx), oy = down<OFF>(ty), oh = down<OFF>(th);
    double o[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) o[k] = down<OFF>(S[k]);
    p.V += V; p.Mx += Mx; p.My += My; p.Mz += Mz;
    if (top > p.top) { p.top = top; tx = ox; ty = oy; th = oh; }
#pragma unroll
    for (int k = 0; k < 10; ++k) S[k] += o[k];
  };
  level(std::integral_constant<int, 16>{}); level(std::integral_constant<int, 8>{}); level(std::integral_constant<int, 4>{});
  level(std::integral_constant<int, 2>{}); level(std::integral_constant<int, 1>{});
  if (lane == 0) {
    SSL double* w = T.bw + unit * kBw;
    w[0] = p.V; w[1] = p.Mx; w[2] = p.My; w[3] = p.Mz; w[4] = p.top;
    for (int k = 0; k < 10; ++k) w[5 + k] = S[k];
    w[15] = tx; w[16] = ty; w[17] = th;
  }
}

// The step's tape from the start state in the tape's input slots (sv, iVol..iCz). Ends with a barrier. The
// values written (oLvl.., vol1, the integrated pose and velocities in tot) equal the forward step's.
template <int TPB, bool GT>
__device__ __forceinline__ void jac_step_views(PC& P, const AdjScalars AM, const Blk& B, const TapeT<GT>& T, int tid) {
  const int NN = P.S.nn, NC = P.M.nc;
  const bool relaxed = AM.mode == 1;
  // pose and its derivatives (frame.hpp:9-21)
  if (tid < 64) {
    const int lane = tid;
    const double pitch = T.sv[2], roll = T.sv[3];
    double v = 0;
    if (lane < 4) v = trig(lane < 2 ? pitch : roll, (lane & 1) == 0 ? 1 : 0);
    const double cth = bcast(v, 0), sth = bcast(v, 1), cph = bcast(v, 2), sph = bcast(v, 3);
    if (lane == 0) {
      write_frame(P, T.fr, T.sv[1], cth, sth, cph, sph);
      write_frame_deriv(P, T.dfp, cth, sth, cph, sph);
      *(SSL int32_t*)T.gnext = 0;
    }
  }
  __syncthreads();
  // (a) buoyancy with its moment sums (the canonical 256 partials: the first 256 threads), (b) the level solve with
  // tangents (every unit; with 512 threads the second 256 start it while the first compute the partials)
  if (TPB == kThreads || tid < kThreads) buoyancy_rec(P, T, tid);
  solve_levels_tv<TPB>(P, B, T, AM.maxGroup, tid, relaxed);
  __syncthreads();
  // (c) heads, effective areas (and which of aov / amin / area gave it) and limiter degrees (flows.hpp:15-35)
  for (int n = tid; n < NN; n += TPB) {
    const double v = T.oVol[n];
    T.hEff[n] = v > kEmptyVolume ? T.oLvl[n] : kNoHead;
    double ae, src;
    if (v >= B.nVmax[n]) { ae = B.nAov[n]; src = 0; }
    else if (B.nAmin[n] > T.oArea[n]) { ae = B.nAmin[n]; src = 1; }
    else { ae = T.oArea[n]; src = 2; }
    T.aEff[n] = ae;
    T.aSrc[n] = src;
  }
  __syncthreads();
  for (int n = tid; n < NN; n += TPB) {
    const int s = P.M.incStart[n], e = s + P.M.incCount[n];
    int d = 0;
    for (int k = s; k < e; ++k) {
      const i4v q = P.M.inc[k];
      if (!B.cen[q.x]) continue;
      const int a = q.y, b = q.z;
      if (a >= 0 && T.hEff[a] < kNoHeadTest && T.hEff[b] < kNoHeadTest) continue;
      ++d;
    }
    T.deg[n] = d;
  }
  __syncthreads();
  // (c) flows with their partial derivatives (flows.hpp:38-78): cp = d dV / d(ha, hb, zo, aEff_a, aEff_b, ln ca)
  const double t = T.sv[0];
  for (int k = tid; k < NC; k += TPB) {
    const int c = P.M.flowOrder[k];  // weirs, then orifices (the forward's flow slots)
    const auto pd = T.cp + 6 * c;
    for (int k = 0; k < 6; ++k) pd[k] = 0;
    T.dV[c] = 0;
    const i4v ci = P.M.conI[c];
    const d4v cd = P.M.conD[c];
    if (!B.cen[c] || B.ctOn[c] > t) continue;
    if (ci.z >= 0 && T.gMer[ci.z] != 0) continue;
    const int a = ci.x, b = ci.y;
    const SSL double* fr = T.fr;
    const double zo = fr[F_R20] * cd.x + fr[F_R21] * cd.y + fr[F_R22] * cd.z + fr[F_TZ];
    const double ha = a < 0 ? 0.0 : T.hEff[a], hbv = T.hEff[b];
    const bool wa = ha > zo, wb = hbv > zo;
    const double Ha = wa ? ha - zo : 0.0, Hb = wb ? hbv - zo : 0.0;
    const double sq2g = P.M.sq2g;
    double q = 0, dqHa = 0, dqHb = 0;
    if (ci.w == 0) {
      const double d = Ha - Hb;
      if (d != 0) {
        const double ad = __builtin_fabs(d);
        const bool floor = kOrificeHeadFloor > ad;
        const double sm = __builtin_sqrt(floor ? kOrificeHeadFloor : ad);
        q = B.carea[c] * sq2g * d / sm;
        const double K = B.carea[c] * sq2g;
        const double dqd = floor ? K / sm : 0.5 * K / sm;
        dqHa = dqd; dqHb = -dqd;
      }
    } else {
      const bool aHi = Ha > Hb;
      const double H1 = aHi ? Ha : Hb;
      if (H1 > 0) {
        const double H2 = aHi ? Hb : Ha;
        double fv = 1, dfr = 0, r = 0;
        if (H2 > 0) {
          r = H2 / H1;
          const double pw = pow_weir(r);
          const double u = 1 - pw;
          const double gg = jmax(0.0, u);
          fv = pow_villemonte(gg);
          const double dpw = kWeirExponent * pw / r;
          const double dfg = gg != 0 ? kVillemonteExponent * fv / gg : 0.0;
          dfr = (u > 0 ? dfg : 0.0) * (-dpw);
        }
        const double sgn = Ha >= Hb ? 1.0 : -1.0;
        const double K = B.carea[c] * kWeirCoefficient * sq2g;
        const double sh = __builtin_sqrt(H1);
        q = K * H1 * sh * fv * sgn;
        const double H15 = H1 * sh;
        double dq1 = K * (sh + H1 * (0.5 / sh)) * fv * sgn;
        double dq2 = 0;
        if (H2 > 0) {
          dq1 += K * H15 * sgn * dfr * (-r / H1);
          dq2 = K * H15 * sgn * dfr / H1;
        }
        if (aHi) { dqHa = dq1; dqHb = dq2; } else { dqHb = dq1; dqHa = dq2; }
      }
    }
    if (q == 0) continue;
    const double dt = P.M.dt;
    double dV = q * dt;
    const int src = dV > 0 ? a : b, dst = dV > 0 ? b : a;
    const double hs = src < 0 ? 0.0 : T.hEff[src];
    const double hd = dst < 0 ? 0.0 : T.hEff[dst];
    const double inv = (src < 0 ? 0.0 : 1 / T.aEff[src]) + (dst < 0 ? 0.0 : 1 / T.aEff[dst]);
    const int ds = src < 0 ? 0 : static_cast<int>(T.deg[src]), dd = dst < 0 ? 0 : static_cast<int>(T.deg[dst]);
    const int degRef = src < 0 ? dd : (dst < 0 ? ds : (ds > dd ? ds : dd));
    const double share = kLimiterShare / static_cast<double>(degRef > 1 ? degRef : 1);
    const bool zoUp = zo > hd;  // jmax(hd, zo) takes zo
    double lim = inv > 0 ? share * (hs - (zoUp ? zo : hd)) / inv : __builtin_fabs(dV);
    bool clamped0 = false;
    if (lim < 0) { lim = 0; clamped0 = true; }
    if (__builtin_fabs(dV) > lim) {
      const double sg = dV > 0 ? 1.0 : -1.0;
      dV = dV > 0 ? lim : -lim;
      if (!clamped0 && inv > 0) {
        // d lim = share (dhs - dmax) / inv - lim dinv / inv; dinv = -daEff / aEff^2 per node end
        const double ks = sg * share / inv;
        double gha = 0, ghb = 0, gzo = 0, gaa = 0, gab = 0;
        if (src == a) gha += ks; else if (src == b) ghb += ks;
        if (zoUp) gzo -= ks;
        else { if (dst == a) gha -= ks; else if (dst == b) ghb -= ks; }
        const double kl = -sg * lim / inv;
        if (src >= 0) { const double de = -1.0 / (T.aEff[src] * T.aEff[src]); if (src == a) gaa += kl * de; else gab += kl * de; }
        if (dst >= 0) { const double de = -1.0 / (T.aEff[dst] * T.aEff[dst]); if (dst == a) gaa += kl * de; else gab += kl * de; }
        if (a < 0) { gha = 0; gaa = 0; }
        pd[0] = gha; pd[1] = ghb; pd[2] = gzo; pd[3] = gaa; pd[4] = gab; pd[5] = 0;
      }
    } else {
      // unlimited: dV = q dt, q proportional to ca
      const double ga = wa ? dqHa * dt : 0.0, gb = wb ? dqHb * dt : 0.0;
      pd[0] = a < 0 ? 0.0 : ga; pd[1] = gb; pd[2] = -(ga + gb); pd[3] = 0; pd[4] = 0; pd[5] = dV;
    }
    T.dV[c] = dV;
  }
  __syncthreads();
  // (c) gathers and volume updates (flows.hpp:81-111)
  for (int n = tid; n < NN; n += TPB) {
    const int s = P.M.incStart[n], e = s + P.M.incCount[n];
    double acc = 0;
    for (int k = s; k < e; ++k) {
      const i4v q = P.M.inc[k];
      acc = q.w ? acc + T.dV[q.x] : acc - T.dV[q.x];
    }
    T.acc[n] = acc;
    double v = T.oVol[n] + acc;
    double clip = 1;
    if (v < 0) { v = 0; clip = 0; }
    const double vm = B.nVmax[n];
    if (v > vm) {
      const double head = jmax(0.0, -T.oZmax[n]);
      const double cap = vm + B.nAov[n] * head;
      if (v > cap) { v = cap; clip = 2; }
    }
    T.vol1[n] = v;
    T.clip[n] = clip;
    // the load terms of the space (body.hpp:21-34), stored gated and signed as the forward's node update stores
    // them (a space with vol <= 0 adds -0.0; tauY's term negated), for the six node-order chains of the tail
    double tx = -0.0, ty = -0.0, tf = -0.0, tm = -0.0, tip = -0.0, tir = -0.0;
    if (!(v <= 0)) {
      const SSL double* fr = T.fr;
      const double W = P.M.rho * P.M.g;
      const double x = T.oCx[n], y = T.oCy[n], z = T.oCz[n];
      const double wx = fr[F_R00] * x + fr[F_R01] * y + fr[F_R02] * z + fr[F_TX] - P.M.XO;
      const double wy = fr[F_R10] * x + fr[F_R11] * y + fr[F_R12] * z + fr[F_TY] - P.M.YO;
      const double wz = fr[F_R20] * x + fr[F_R21] * y + fr[F_R22] * z + fr[F_TZ] - T.sv[1];
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
  __syncthreads();
  // (a) buoyancy totals, (d) loads and integration (body.hpp:11-57) on wave 0: lanes 0..5 run the six node-order
  // chains over the stored terms (the forward's step tail), lane 0 integrates
  if (tid < 64) {
    const int lane = tid;
    const SSL double* fr = T.fr;
    double tV = T.bw[0], tMx = T.bw[1], tMy = T.bw[2], tMz = T.bw[3], tTop = T.bw[4];
    double tc3[3] = {T.bw[15], T.bw[16], T.bw[17]};
    double S[10];
    for (int k = 0; k < 10; ++k) S[k] = T.bw[5 + k];
    for (int w = 1; w < kUnits; ++w) {
      const SSL double* q = T.bw + kBw * w;
      tV += q[0]; tMx += q[1]; tMy += q[2]; tMz += q[3];
      if (q[4] > tTop) { tTop = q[4]; tc3[0] = q[15]; tc3[1] = q[16]; tc3[2] = q[17]; }
      for (int k = 0; k < 10; ++k) S[k] += q[5 + k];
    }
    if (lane == 0) {
      for (int k = 0; k < 10; ++k) T.bt[4 + k] = S[k];
      T.bt[14] = tc3[0]; T.bt[15] = tc3[1]; T.bt[16] = tc3[2];
      T.bt[0] = tV; T.bt[1] = tMx; T.bt[2] = tMy; T.bt[3] = tMz;
    }
    const double V = tV;
    const double bx = tV > 0 ? tMx / tV : 0, by = tV > 0 ? tMy / tV : 0, bz = tV > 0 ? tMz / tV : 0;
    const double W = P.M.rho * P.M.g;
    const double Fb = W * V;
    const double bxw = fr[F_R00] * bx + fr[F_R01] * by + fr[F_R02] * bz + fr[F_TX];
    const double byw = fr[F_R10] * bx + fr[F_R11] * by + fr[F_R12] * bz + fr[F_TY];
    const double zO = T.sv[1];
    double acc = 0;
    if (lane < 6) {
      const SSL double* src = lane == 0 ? B.tX : lane == 1 ? B.tY : lane == 2 ? B.tF : lane == 3 ? B.tM : lane == 4 ? B.tIp : B.tIr;
      acc = lane == 0 ? (byw - P.M.YO) * Fb : lane == 1 ? -(bxw - P.M.XO) * Fb : lane == 2 ? Fb - P.M.g * P.M.ms : 0.0;
      for (int n0 = 0; n0 < NN; n0 += 8) {
        double sv8[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) sv8[k] = src[n0 + k < NN ? n0 + k : NN - 1];
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (n0 + k < NN) acc += sv8[k];
      }
    }
    const double tauX = bcast(acc, 0), tauY = bcast(acc, 1), Fz = bcast(acc, 2), mw = bcast(acc, 3);
    const double Iwp = bcast(acc, 4), Iwr = bcast(acc, 5);
    const double Mth = tauY, Mph = tauX * fr[F_CTH];
    const double dt = P.M.dt;
    double vz = T.sv[4], wth = T.sv[5], wph = T.sv[6];
    const double mh = P.M.mh0 + mw, Ith = P.M.Ith0 + Iwp, Iph = P.M.Iph0 + Iwr;
    const double az = (Fz - P.M.cDz * vz - P.M.qz * __builtin_fabs(vz) * vz) / mh;
    const double ath = (Mth - P.M.cDth * wth - P.M.qth * __builtin_fabs(wth) * wth) / Ith;
    const double aph = (Mph - P.M.cDph * wph - P.M.qph * __builtin_fabs(wph) * wph) / Iph;
    vz += az * dt; wth += ath * dt; wph += aph * dt;
    double zO1 = zO + vz * dt, pitch = T.sv[2] + wth * dt, roll = T.sv[3] + wph * dt;
    double clp = 0, clr = 0;
    if (pitch > kPitchClamp) { pitch = kPitchClamp; wth = 0; clp = 1; }
    if (pitch < -kPitchClamp) { pitch = -kPitchClamp; wth = 0; clp = 1; }
    if (roll > kRollClamp) { roll = kRollClamp; wph = 0; clr = 1; }
    if (roll < -kRollClamp) { roll = -kRollClamp; wph = 0; clr = 1; }
    if (lane == 0) {
      SSL double* o = T.tot;
      o[T_V] = V; o[T_MX] = tMx; o[T_MY] = tMy; o[T_MZ] = tMz; o[T_IR] = fr[F_IR]; o[T_BX] = bx; o[T_BY] = by; o[T_BZ] = bz;
      o[T_FB] = Fb; o[T_BXW] = bxw; o[T_BYW] = byw; o[T_TAUX] = tauX; o[T_TAUY] = tauY; o[T_FZ] = Fz; o[T_MW] = mw;
      o[T_IWP] = Iwp; o[T_IWR] = Iwr; o[T_MH] = mh; o[T_ITH] = Ith; o[T_IPH] = Iph; o[T_AZ] = az; o[T_ATH] = ath; o[T_APH] = aph;
      o[T_CLP] = clp; o[T_CLR] = clr; o[T_VZ1] = vz; o[T_WTH1] = wth; o[T_WPH1] = wph; o[T_ZO1] = zO1; o[T_P1] = pitch; o[T_R1] = roll;
    }
  }
  __syncthreads();
}

// ------------------------------------------------------------------------------------------- reverse

// The instance's per-connection parameter indices (pa, pc: its rows) and every connection's monitor into the tape;
// clears the seed state. Ends with a barrier.
template <int TPB, bool GT>
__device__ void load_adjoint_tables(PC& P, const TapeT<GT>& T, const SSG int32_t* pa, const SSG int32_t* pc, int tid) {
  const int NC = P.M.nc;
  for (int c = tid; c < NC; c += TPB) { T.cpa[c] = pa[c]; T.cpc[c] = pc[c]; T.cmon[c] = -1; }
  for (int m = tid; m < P.M.nmon; m += TPB) T.monSeed[m] = 0;
  if (tid == 0) { T.sflag[0] = 0; T.sflag[1] = 0; }
  __syncthreads();
  for (int m = tid; m < P.M.nmon; m += TPB) {
    const int s = P.M.monStart[m], e = s + P.M.monCount[m];
    for (int j = s; j < e; ++j) T.cmon[P.M.monConn[j]] = m;
  }
  __syncthreads();
}

// Frame-slot gradients of one thread (F_R00..F_CTH), and the direct zO gradient of the loads' wz.
struct FrameGrad {
  double g[kFrameDeriv];
  double zO;
};

__device__ __forceinline__ void add_world_z(FrameGrad& G, double gz, double x, double y, double z) {
  G.g[F_R20] += gz * x; G.g[F_R21] += gz * y; G.g[F_R22] += gz * z; G.g[F_TZ] += gz;
}

// The adjoint of step k: from lamCur (the adjoint of the state after the step, record order zO pitch roll vz wth
// wph, vol, level, cx, cy, cz) and the loss's seeds at step k, the adjoint of the state before it in lamNext, and
// the parameters' log-derivatives added to G.
template <int TPB, bool GT>
__device__ __forceinline__ void reverse_step_views(PC& P, const AdjScalars AM, const Blk& B, const TapeT<GT>& T, int tid, long long k,
                                                   const SSL double* lamCur, SSL double* lamNext) {
  const int NN = P.S.nn, NC = P.M.nc;
  const SSL double* fr = T.fr;
  const SSL double* o = T.tot;
  const double W = P.M.rho * P.M.g, dt = P.M.dt;
  FrameGrad FG;
  for (int e = 0; e < kFrameDeriv; ++e) FG.g[e] = 0;
  FG.zO = 0;
  // R0: zero the array part of lamNext; thread 0: integration, loads totals and buoyancy in reverse, mark seeds
  for (int i = 6 + tid; i < 6 + 5 * NN; i += TPB) lamNext[i] = 0;
  if (tid == 0) {
    // this step's seeds (the records are sorted, latest step first)
    double sPitch = 0, sRoll = 0, sTop = 0, sV = 0;
    bool anyMon = false;
    int pos = static_cast<int>(T.sflag[0]);
    const int ns = static_cast<int>(T.nseeds[0]);
    for (; pos < ns && static_cast<long long>(T.seeds[4 * pos]) == k; ++pos) {
      const SSL double* r = T.seeds + 4 * pos;
      const int kind = static_cast<int>(r[1]), idx = static_cast<int>(r[2]);
      const double v = r[3];
      if (kind == 0) sPitch += v;
      else if (kind == 3) sRoll += v;
      else if (kind == 4) sTop += v;
      else if (kind == 5) sV += v;
      else if (kind == 2) add_world_z(FG, v, P.M.mkx[idx], P.M.mky[idx], P.M.mkz[idx]);
      else { T.monSeed[idx] += v; anyMon = true; }
    }
    T.sflag[0] = pos;
    T.sflag[1] = anyMon ? 1.0 : 0.0;
    const double lz = lamCur[0], lvz = lamCur[3], lwp = lamCur[5];
    const double lp = lamCur[1] + sPitch, lr = lamCur[2] + sRoll, lwt = lamCur[4];
    const bool cp_ = o[T_CLP] != 0, cr = o[T_CLR] != 0;
    const double gpRaw = cp_ ? 0.0 : lp, gwtRaw = (cp_ ? 0.0 : lwt) + gpRaw * dt;
    const double grRaw = cr ? 0.0 : lr, gwpRaw = (cr ? 0.0 : lwp) + grRaw * dt;
    const double gvz1 = lvz + lz * dt;
    const double vz = T.sv[4], wth = T.sv[5], wph = T.sv[6];
    const double gaz = gvz1 * dt, gath = gwtRaw * dt, gaph = gwpRaw * dt;
    const double mh = o[T_MH], Ith = o[T_ITH], Iph = o[T_IPH];
    lamNext[0] = lz;      // zO' = zO + ...
    lamNext[1] = gpRaw;   // pitch' = pitch + ...
    lamNext[2] = grRaw;
    lamNext[3] = gvz1 + gaz * (-P.M.cDz - 2 * P.M.qz * __builtin_fabs(vz)) / mh;
    lamNext[4] = gwtRaw + gath * (-P.M.cDth - 2 * P.M.qth * __builtin_fabs(wth)) / Ith;
    lamNext[5] = gwpRaw + gaph * (-P.M.cDph - 2 * P.M.qph * __builtin_fabs(wph)) / Iph;
    const double gFz = gaz / mh, gMth = gath / Ith, gMph = gaph / Iph;
    const double gmw = -gaz * o[T_AZ] / mh, gIwp = -gath * o[T_ATH] / Ith, gIwr = -gaph * o[T_APH] / Iph;
    const double gTauY = gMth, gTauX = gMph * fr[F_CTH];
    FG.g[F_CTH] += gMph * o[T_TAUX];
    // the buoyancy's part of the loads: Fz = Fb - g ms + ..., tauX = (byw - YO) Fb + ..., tauY = -(bxw - XO) Fb - ...
    const double Fb = o[T_FB];
    const double gFb = gFz + gTauX * (o[T_BYW] - P.M.YO) - gTauY * (o[T_BXW] - P.M.XO);
    const double gbyw = gTauX * Fb, gbxw = -gTauY * Fb;
    const double bx = o[T_BX], by = o[T_BY], bz = o[T_BZ];
    FG.g[F_R00] += gbxw * bx; FG.g[F_R01] += gbxw * by; FG.g[F_R02] += gbxw * bz; FG.g[F_TX] += gbxw;
    FG.g[F_R10] += gbyw * bx; FG.g[F_R11] += gbyw * by; FG.g[F_R12] += gbyw * bz; FG.g[F_TY] += gbyw;
    const double gbx = gbxw * fr[F_R00] + gbyw * fr[F_R10], gby = gbxw * fr[F_R01] + gbyw * fr[F_R11], gbz = gbxw * fr[F_R02] + gbyw * fr[F_R12];
    double gV = W * gFb + sV, gMx = 0, gMy = 0, gMz = 0;
    const double V = o[T_V];
    if (V > 0) {
      gMx = gbx / V; gMy = gby / V; gMz = gbz / V;
      gV -= (gbx * bx + gby * by + gbz * bz) / V;
    }
    // the integrals' moments: d(V, Mx, My, Mz) = -iR Tm (dtz, dR20, dR21, dR22) over the partly wet columns
    const SSL double* S = T.bt + 4;
    const double iR = o[T_IR];
    const double T00 = S[0], T01 = S[1], T02 = S[2], T03 = S[3], T11 = S[4], T12 = S[5], T13 = S[6], T22 = S[7], T23 = S[8], T33 = S[9];
    FG.g[F_TZ] += -iR * (gV * T00 + gMx * T01 + gMy * T02 + gMz * T03);
    FG.g[F_R20] += -iR * (gV * T01 + gMx * T11 + gMy * T12 + gMz * T13);
    FG.g[F_R21] += -iR * (gV * T02 + gMx * T12 + gMy * T22 + gMz * T23);
    FG.g[F_R22] += -iR * (gV * T03 + gMx * T13 + gMy * T23 + gMz * T33);
    // hull-top seeds: the highest column top, base + R22 * top of its column
    if (sTop != 0) add_world_z(FG, sTop, T.bt[14], T.bt[15], T.bt[16]);
    SSL double* g = T.tot + T_GFZ;
    g[0] = gFz; g[1] = gTauX; g[2] = gTauY; g[3] = gmw; g[4] = gIwp; g[5] = gIwr;
  }
  __syncthreads();
  // R1: per space: the loads' node terms and the volume update in reverse
  {
    const SSL double* g = T.tot + T_GFZ;
    const double gFz = g[0], gTauX = g[1], gTauY = g[2], gmw = g[3], gIwp = g[4], gIwr = g[5];
    const double zO = T.sv[1];
    for (int n = tid; n < NN; n += TPB) {
      double gv1 = lamCur[6 + n];
      double gcx = lamCur[6 + 2 * NN + n], gcy = lamCur[6 + 3 * NN + n], gcz = lamCur[6 + 4 * NN + n];
      const double Vn = T.vol1[n];
      if (!(Vn <= 0)) {
        const double x = T.oCx[n], y = T.oCy[n], z = T.oCz[n];
        const double wx = fr[F_R00] * x + fr[F_R01] * y + fr[F_R02] * z + fr[F_TX] - P.M.XO;
        const double wy = fr[F_R10] * x + fr[F_R11] * y + fr[F_R12] * z + fr[F_TY] - P.M.YO;
        const double wz = fr[F_R20] * x + fr[F_R21] * y + fr[F_R22] * z + fr[F_TZ] - zO;
        const double fn = -W * Vn, m = P.M.rho * Vn;
        const double gf = gFz + gTauX * wy - gTauY * wx;
        const double gwy = gTauX * fn + gIwr * m * 2 * wy;
        const double gwx = -gTauY * fn + gIwp * m * 2 * wx;
        const double gwz = (gIwp + gIwr) * m * 2 * wz;
        const double gm = gmw + gIwp * (wx * wx + wz * wz) + gIwr * (wy * wy + wz * wz);
        gv1 += -W * gf + P.M.rho * gm;
        gcx += gwx * fr[F_R00] + gwy * fr[F_R10] + gwz * fr[F_R20];
        gcy += gwx * fr[F_R01] + gwy * fr[F_R11] + gwz * fr[F_R21];
        gcz += gwx * fr[F_R02] + gwy * fr[F_R12] + gwz * fr[F_R22];
        FG.g[F_R00] += gwx * x; FG.g[F_R01] += gwx * y; FG.g[F_R02] += gwx * z; FG.g[F_TX] += gwx;
        FG.g[F_R10] += gwy * x; FG.g[F_R11] += gwy * y; FG.g[F_R12] += gwy * z; FG.g[F_TY] += gwy;
        FG.g[F_R20] += gwz * x; FG.g[F_R21] += gwz * y; FG.g[F_R22] += gwz * z; FG.g[F_TZ] += gwz;
        FG.zO -= gwz;
      }
      SSL double* go = T.gOut + n * kSolveOut;
      go[0] = lamCur[6 + NN + n];  // level
      go[3] = gcx; go[4] = gcy; go[5] = gcz;
      go[2] = 0;  // free-surface area, from the flows' limiter below
      const int clip = static_cast<int>(T.clip[n]);
      double gvm = 0, gacc = 0, gzm = 0;
      if (clip == 1) { gvm = gv1; gacc = gv1; }
      else if (clip == 2 && -T.oZmax[n] > 0) gzm = -B.nAov[n] * gv1;
      go[1] = gvm;
      go[6] = gzm;
      T.gAcc[n] = gacc;
    }
  }
  __syncthreads();
  // R2: per connection: the gathers in reverse, monitor seeds, the flow's partial derivatives
  for (int c = tid; c < NC; c += TPB) {
    const auto cg = T.cg + 5 * c;
    const double dV = T.dV[c];
    const i4v ci = P.M.conI[c];
    double g = 0;
    if (dV != 0) {
      g = T.gAcc[ci.y] - (ci.x >= 0 ? T.gAcc[ci.x] : 0.0);
      // monitor seeds: the step's flow sum of a monitor takes q = dV / dt of its connections with q > 0
      if (dV > 0 && T.sflag[1] != 0) {
        const int m = static_cast<int>(T.cmon[c]);
        if (m >= 0) g += T.monSeed[m] / dt;
      }
    }
    const auto pd = T.cp + 6 * c;
    cg[0] = g * pd[0]; cg[1] = g * pd[1]; cg[2] = g * pd[3]; cg[3] = g * pd[4]; cg[4] = g * pd[5];
    if (g != 0 && pd[2] != 0) {
      const d4v cd = P.M.conD[c];
      add_world_z(FG, g * pd[2], cd.x, cd.y, cd.z);
    }
  }
  __syncthreads();
  // R3: per space: heads and effective areas gathered over its connections (incidence order); the parameters'
  // log-derivatives in connection order
  for (int n = tid; n < NN; n += TPB) {
    const int s = P.M.incStart[n], e = s + P.M.incCount[n];
    double gh = 0, ga = 0;
    // four entries' loads issued together, then added in incidence order
    for (int j0 = s; j0 < e; j0 += 4) {
      i4v q[4];
      double a[4], b[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) q[k] = P.M.inc[j0 + k < e ? j0 + k : e - 1];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const auto cg = T.cg + 5 * q[k].x;
        a[k] = q[k].w ? cg[1] : cg[0];
        b[k] = q[k].w ? cg[3] : cg[2];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (j0 + k < e) { gh += a[k]; ga += b[k]; }
    }
    SSL double* go = T.gOut + n * kSolveOut;
    if (T.oVol[n] > kEmptyVolume) go[0] += gh;
    if (T.aSrc[n] == 2) go[2] += ga;
  }
  {
    // wave w takes parameters w, w + TPB/64, ...: lane l reads connections l, l + 64, ... once and adds each to the
    // sums of the wave's parameters it belongs to (per parameter in connection order), then a fixed butterfly
    constexpr int kWaves = TPB / 64, kPer = (kMaxParams + kWaves - 1) / kWaves;
    const int wave = tid >> 6, lane = tid & 63;
    double sp[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) sp[k] = 0;
    for (int c = lane; c < NC; c += 64) {
      const double ia = T.cpa[c], ic = T.cpc[c];
      const double v = T.cg[5 * c + 4];
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int p = wave + k * kWaves;
        if (p < AM.np && (ia == p || ic == p)) sp[k] += v;
      }
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int p = wave + k * kWaves;
      if (p >= AM.np) break;
      double sv = sp[k];
      for (int off = 32; off > 0; off >>= 1) sv += __shfl_xor(sv, off, 64);
      if (lane == 0) T.G[p] += sv;
    }
  }
  if (tid == 0 && T.sflag[1] != 0) {  // the step's monitor seeds are used
    for (int m = 0; m < P.M.nmon; ++m) T.monSeed[m] = 0;
    T.sflag[1] = 0;
  }
  __syncthreads();
  // R4: per group: the level solve in reverse through its local Jacobians (rows along V, h0, zO, pitch, roll)
  double PG[3] = {0, 0, 0};  // pose gradients the solve gives directly
  for (int g = tid; g < P.M.ng; g += TPB) {
    const int N = B.gCount[g], first = B.gStart[g];
    const bool merged = T.gMer[g] != 0;
    double gd[kDof] = {0, 0, 0, 0, 0};
    for (int kk = 0; kk < N; ++kk) {
      const int n = B.gNodes[first + kk];
      const SSL double* go = T.gOut + n * kSolveOut;
      const auto J = T.jac + n * (kSolveOut * kDof);
      double gn[kDof] = {0, 0, 0, 0, 0};
      for (int oo = 0; oo < kSolveOut; ++oo)
        for (int l = 0; l < kDof; ++l) gn[l] += go[oo] * J[oo * kDof + l];
      if (T.carry[n] != 0) {
        lamNext[6 + 2 * NN + n] += go[3]; lamNext[6 + 3 * NN + n] += go[4]; lamNext[6 + 4 * NN + n] += go[5];
      }
      if (merged) {
        for (int l = 0; l < kDof; ++l) gd[l] += gn[l];
      } else {
        lamNext[6 + n] += gn[0];
        lamNext[6 + NN + n] += gn[1];
        PG[0] += gn[2]; PG[1] += gn[3]; PG[2] += gn[4];
      }
    }
    if (merged) {
      // V is the members' total; h0 their mean level (all wet) or one member's level
      const int h0 = static_cast<int>(T.gH0[g]);
      for (int kk = 0; kk < N; ++kk) {
        const int n = B.gNodes[first + kk];
        lamNext[6 + n] += gd[0];
        if (h0 < 0) lamNext[6 + NN + n] += gd[1] / static_cast<double>(N);
        else if (kk == h0) lamNext[6 + NN + n] += gd[1];
      }
      PG[0] += gd[2]; PG[1] += gd[3]; PG[2] += gd[4];
    }
  }
  // R5: every thread's frame-slot gradients to the pose, reduced over the workgroup (a fixed butterfly in each wave,
  // then the waves in order) onto the pose entries of lamNext
  double gp[3];
  for (int j2 = 0; j2 < 3; ++j2) {
    double v = PG[j2] + (j2 == 0 ? FG.zO : 0.0);
    for (int e = 0; e < kFrameDeriv; ++e) v += FG.g[e] * T.dfp[e * 3 + j2];
    gp[j2] = v;
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int j2 = 0; j2 < 3; ++j2) gp[j2] += __shfl_xor(gp[j2], off, 64);
  const int wave = tid / 64;
  if ((tid & 63) == 0)
    for (int j2 = 0; j2 < 3; ++j2) T.red[wave * 3 + j2] = gp[j2];
  __syncthreads();
  if (tid == 0) {
    for (int j2 = 0; j2 < 3; ++j2) {
      double v = 0;
      for (int w = 0; w < TPB / 64; ++w) v += T.red[w * 3 + j2];
      lamNext[j2] += v;
    }
  }
  __syncthreads();
}

// The Jacobian and reverse steps as calls that take scalars (the addresses the views are carved from, the adjoint's
// settings) and carve the views themselves: the kernel that runs them also runs the forward step, and the three
// inlined together exceed the register file; views passed by reference to a function kept out of line would go
// through scratch memory at every call.
template <int TPB, bool GT>
__device__ __attribute__((noinline)) void jac_step(ViewAddrs va, AdjScalars AM, int tid) {
  extern __shared__ double lds_dyn[];
  const StepViews<GT> v(uniform64(va.pa), uniform64(va.ta), (SSL double*)lds_dyn, (SSG double*)uniform64(va.gs));
  const AdjScalars A{__builtin_amdgcn_readfirstlane(AM.maxGroup), __builtin_amdgcn_readfirstlane(AM.np),
                     __builtin_amdgcn_readfirstlane(AM.mode)};
  jac_step_views<TPB>(v.P, A, v.B, v.T, tid);
}
// inA: the adjoint of the state after the step is in lamA (the state before it goes to lamB), else the other way
template <int TPB, bool GT>
__device__ __attribute__((noinline)) void reverse_step(ViewAddrs va, AdjScalars AM, int tid, long long k, bool inA) {
  extern __shared__ double lds_dyn[];
  const StepViews<GT> v(uniform64(va.pa), uniform64(va.ta), (SSL double*)lds_dyn, (SSG double*)uniform64(va.gs));
  const AdjScalars A{__builtin_amdgcn_readfirstlane(AM.maxGroup), __builtin_amdgcn_readfirstlane(AM.np),
                     __builtin_amdgcn_readfirstlane(AM.mode)};
  const bool a = __builtin_amdgcn_readfirstlane(inA ? 1 : 0) != 0;
  reverse_step_views<TPB>(v.P, A, v.B, v.T, tid, k, a ? v.T.lamA : v.T.lamB, a ? v.T.lamB : v.T.lamA);
}

}  // namespace ssa
