// The gradient of a batch (no counterpart in the reference): every instance's loss and its derivative with respect
// to connection scale factors, by a reverse-mode adjoint of the step with checkpoints.
//
// One workgroup per instance runs three sweeps over its own trajectory:
//   1. forward: the plain step (device_step, the batch kernel's step) from the initial state to founder / tMax /
//      maxSteps, storing the state every K steps (a checkpoint) and the loss's ingredients (trim or list after given
//      steps, the signed event signals at and before each event's first firing step);
//   2. loss: the loss, the signals and the seeds of the reverse sweep (the loss's derivative with respect to the
//      pitch or roll after a readout step, and to a monitor's flow sum, a mark's world z or the founder rule's
//      deciding margin -- hull top, pitch or buoyancy volume -- at an event step and the one before);
//   3. reverse: for each segment of K steps, last first: the segment recomputed from its checkpoint with the plain
//      step into a ring of K + 1 states, then its steps in reverse order, each recomputed with its local Jacobians
//      (jac_step, whose values are compared with the ring's next state) and pulled back (reverse_step).
// The parameters enter each connection's product coef * area (the host forms the scaled products); the gradient
// is the sum of the log-derivatives over the connections of a parameter, divided by the parameter.
//
// A batch of portable32 numerics runs the forward and the segment recomputations with the portable32 step (FP32
// column and segment arithmetic), bit for bit the batch kernel's trajectory, and the reverse in FP64: the local
// Jacobians are the FP64 step's, taken at the portable32 trajectory's states. The recomputation is then checked at
// the segment ends (each recomputed segment's last state against the forward's next checkpoint) instead of step by
// step, since the FP64 recomputation inside jac_step does not reproduce portable32 values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/sinksim_hip.h"
#include "adjoint_dev.hpp"
#include "batch_host.hpp"

namespace ssa {

// The gradient run's per-batch view.
struct DevAdj {
  int32_t K, wantGrad;
  long long maxSteps;
  double tMax;
  gcd* theta;                      // [count * np]
  gci *pa, *pc;                    // per instance connection (connOff)
  gcd* ca;                         // the scaled products coef * area per instance connection
  int32_t nRead, nEv;
  const SSG long long* rdStep;
  gci* rdKind;
  gcd *rdW, *rdAfter;
  gci *evKind, *evIdx;
  gcd *evW, *evMissing;
  gcd* obs;                        // [count * (nRead + nEv)]
  SSG double *loss, *grad, *signals;
  SSG long long* steps;
  SSG int32_t* mismatch;           // per instance: recomputed steps whose values differ from the forward's
  SSG double *ckpt, *ring;
  const SSG long long *ckptOff, *ringOff;  // per instance, in doubles
  SSG double* gtape;                       // the tape's arrays kept in global memory (GT kernels): gStride per instance
  long long gStride;
  // the one-step form
  gcd *stepRec, *stepLam, *stepMon, *stepMark;
  SSG double *stepLamOut, *stepRecOut;
};

using DC = const SSC DevAdj;
using AMC = const SSC AdjModel;

// The state of the forward layout as a record of 7 + 5 NN values (t zO pitch roll vz wth wph, vol, level, cx, cy,
// cz). Each copy runs as one loop over the scalars and one over the spaces with the five arrays side by side (a
// select among the arrays by a computed index would become a table of their addresses in scratch memory).
__device__ inline void save_state(PC& P, const Blk& B, SSG double* dst, int tid, int tpb) {
  const int NN = P.S.nn;
  if (tid < 7) dst[tid] = B.scal[S_T + tid];
  for (int n = tid; n < NN; n += tpb) {
    SSG double* d = dst + 7 + n;
    d[0] = B.vol[n]; d[NN] = B.level[n]; d[2 * NN] = B.cx[n]; d[3 * NN] = B.cy[n]; d[4 * NN] = B.cz[n];
  }
}
__device__ inline void restore_state(PC& P, const Blk& B, const SSG double* src, int tid, int tpb) {
  const int NN = P.S.nn;
  if (tid < 7) B.scal[S_T + tid] = src[tid];
  for (int n = tid; n < NN; n += tpb) {
    const SSG double* s = src + 7 + n;
    B.vol[n] = s[0]; B.level[n] = s[NN]; B.cx[n] = s[2 * NN]; B.cy[n] = s[3 * NN]; B.cz[n] = s[4 * NN];
  }
}
template <bool GT>
__device__ inline void load_tape_state(PC& P, const TapeT<GT>& T, const SSG double* src, int tid, int tpb) {
  const int NN = P.S.nn;
  if (tid < 7) T.sv[tid] = src[tid];
  for (int n = tid; n < NN; n += tpb) {
    const SSG double* s = src + 7 + n;
    T.iVol[n] = s[0]; T.iLvl[n] = s[NN]; T.iCx[n] = s[2 * NN]; T.iCy[n] = s[3 * NN]; T.iCz[n] = s[4 * NN];
  }
}

// the recomputed step's values against the stored next state; counts differences into *bad
template <bool GT>
__device__ inline void verify_step(PC& P, const TapeT<GT>& T, const SSG double* next, int tid, int tpb, SSL int32_t* bad) {
  const int NN = P.S.nn;
  auto differ = [](double a, double b) { return __builtin_bit_cast(uint64_t, a) != __builtin_bit_cast(uint64_t, b); };
  int nbad = 0;
  if (tid < 7) {
    const SSL double* o = T.tot;
    const double v = tid == 0 ? T.sv[0] + P.M.dt : tid == 1 ? o[T_ZO1] : tid == 2 ? o[T_P1] : tid == 3 ? o[T_R1] : tid == 4 ? o[T_VZ1]
                     : tid == 5 ? o[T_WTH1] : o[T_WPH1];
    nbad += differ(v, next[tid]);
  }
  for (int n = tid; n < NN; n += tpb) {
    const SSG double* s = next + 7 + n;
    nbad += differ(T.vol1[n], s[0]) + differ(T.oLvl[n], s[NN]) + differ(T.oCx[n], s[2 * NN]) + differ(T.oCy[n], s[3 * NN]) +
            differ(T.oCz[n], s[4 * NN]);
  }
  if (nbad) atomicAdd((int32_t*)bad, nbad);
}

// Loads instance inst into the forward layout (run_instance's load, with the scaled products as carea).
__device__ inline void load_instance(PC& P, BC& Bt, DC& D, const Blk& B, int inst, int tid, int tpb) {
  const int NN = P.S.nn, NC = P.M.nc;
  const size_t iN = static_cast<size_t>(Bt.nodeOff[inst]), iC = static_cast<size_t>(Bt.connOff[inst]);
  if (tid < 7) { B.scal[S_T + tid] = Bt.st[inst * 7 + tid]; B.scal[O_INFLOW + tid] = 0; }
  if (tid == 0) {
    *B.nextGroup = 0;
    B.scal[S_EVCOUNT] = 0; B.scal[S_FOUNDERED] = 0; B.scal[S_FOUNDERT] = -1; B.scal[S_SEA] = 0;
    B.scal[S_NEXTCURVE] = 0; B.scal[S_CURVECOUNT] = 0;
  }
  for (int n = tid; n < NN; n += tpb) {
    B.vol[n] = Bt.vol[iN + n]; B.level[n] = Bt.level[iN + n];
    B.cx[n] = Bt.cx[iN + n]; B.cy[n] = Bt.cy[iN + n]; B.cz[n] = Bt.cz[iN + n]; B.cvol[n] = 0;
    B.area[n] = 0; B.zmin[n] = 0; B.zmax[n] = 0; B.hEff[n] = 0; B.aEff[n] = 0; B.excess[n] = 0; B.deg[n] = 0;
    B.tX[n] = 0; B.tY[n] = 0; B.tF[n] = 0; B.tM[n] = 0; B.tIp[n] = 0; B.tIr[n] = 0;
    B.nSegStart[n] = P.S.segStart[n]; B.nSegCount[n] = P.S.segCount[n];
    B.nMu[n] = P.S.mu[n]; B.nVmax[n] = P.S.vmax[n]; B.nAov[n] = P.S.aov[n]; B.nAmin[n] = P.S.amin[n];
  }
  for (int c = tid; c < NC; c += tpb) {
    B.cen[c] = Bt.en[iC + c]; B.carea[c] = D.ca[iC + c]; B.ctOn[c] = Bt.tOn[iC + c];
    B.dV[c] = 0; B.q[c] = 0;
  }
  for (int g = tid; g < P.M.ng; g += tpb) {
    B.merged[g] = 0;
    B.gStart[g] = P.M.gStart[g]; B.gCount[g] = P.M.gCount[g]; B.gMode[g] = P.M.gMode[g];
    B.gpx[g] = P.M.gpx[g]; B.gpy[g] = P.M.gpy[g]; B.gpz[g] = P.M.gpz[g];
  }
  for (int k = tid; k < P.M.nMembers; k += tpb) B.gNodes[k] = P.M.gNodes[k];
  for (int e = tid; e < P.M.evCap; e += tpb) { B.evT[e] = 0; B.evId[e] = 0; B.evKind[e] = 0; }
  for (int m = tid; m < P.M.nmon; m += tpb) { B.overT[m] = -1; B.over[m] = 0; }
  for (int m = tid; m < P.M.nmark; m += tpb) B.markT[m] = -1;
}

// the signed event signal of loss event e in the step just taken (mark world z taken before the step, in zpre)
__device__ __forceinline__ double event_signal(PC& P, int kind, int idx, double zpre) {
  if (kind == 0) return 0;  // monitors: read from B.over by the caller
  return P.M.mkWhen[idx] == 0 ? -zpre : zpre;
}

constexpr int kMaxTPB = 512;

// the views of the instance in the kernel's own phases (StepViews)
#define SSA_VIEWS                                   \
  const StepViews<GT> views_(pa, ta, lds, gslice);  \
  PC& P = views_.P;                                 \
  const Blk& B = views_.B;                          \
  const TapeT<GT>& T = views_.T

// The forward-only instantiation (FWD) is held to 128 VGPRs, so that two 512-thread workgroups share a compute unit.
// Mixed: the forward and its recomputations in portable32 numerics.
template <int kTPB, bool GT, bool FWD, bool Mixed>
__global__ __launch_bounds__(kTPB) __attribute__((amdgpu_waves_per_eu(FWD ? 4 : 1, 8))) void adjoint_run(const ModelParams* __restrict__ models, const AdjModel* __restrict__ amodels,
                                                     const DevBatch* __restrict__ batch, const DevAdj* __restrict__ adj) {
  BC& Bt = *(BC*)batch;
  DC& D = *(DC*)adj;
  extern __shared__ double lds_raw[];
  __shared__ int32_t bad;
  SSL double* lds = (SSL double*)lds_raw;
  const int tid = threadIdx.x;
  const int inst = blockIdx.x;
  const int mi = Bt.instModel[inst];
  PC& P0 = *(PC*)((const SSC char*)models + static_cast<size_t>(mi) * sizeof(ModelParams));
  AMC& AMs = *(AMC*)((const SSC char*)amodels + static_cast<size_t>(mi) * sizeof(AdjModel));
  const AdjScalars AM{AMs.maxGroup, AMs.np, AMs.mode};
  const uint64_t pa = uniform64(reinterpret_cast<uint64_t>(&P0));
  const uint64_t ta = uniform64(reinterpret_cast<uint64_t>(&AMs));
  SSG double* const gslice = D.gtape + static_cast<long long>(inst) * D.gStride;
  const int NN = P0.S.nn;
  const int SS = 7 + 5 * NN;
  const int nT = D.nRead, nE = D.nEv, nSig = nT + nE;
  SSG double* ckpt = D.ckpt + D.ckptOff[inst];
  SSG double* ring = D.ring + D.ringOff[inst];
  const long long K = D.K;
  // signals (the tape's sig slots): readout value, readout set; per event (kEvRec): fired, step, s_k, s_{k-1},
  // previous s, time at firing, z before the step, previous margin code, code at firing, code before firing (founder
  // margins: 0 hull top, +1 / -1 pitch by its sign, 2 buoyancy)
  constexpr int kEvRec = 10;

  {
    SSA_VIEWS;
    load_instance(P, Bt, D, B, inst, tid, kTPB);
    for (int i = tid; i < 2 * nT + kEvRec * nE; i += kTPB) T.sig[i] = 0;
    if (tid == 0) bad = 0;
    __syncthreads();
    if (tid < 64) pose_wave0(P, B, tid, B.scal[S_ZO], B.scal[S_PITCH], B.scal[S_ROLL]);
    __syncthreads();
  }

  // 1. forward with checkpoints and the loss's ingredients
  long long done = 0;
  for (;;) {
    {
      SSA_VIEWS;
      // every thread reads the same state (after the step's last barrier): the exit is uniform, and so are the step
      // count and every loop over it below
      const int stop = B.scal[S_FOUNDERED] != 0 || B.scal[S_T] >= D.tMax || done >= D.maxSteps;
      if (__builtin_amdgcn_readfirstlane(stop)) break;
      if (done % K == 0) save_state(P, B, ckpt + (done / K) * SS, tid, kTPB);
      SSL double* ev = T.sig + 2 * nT;
      if (tid == 0)
        for (int e = 0; e < nE; ++e)
          if (D.evKind[e] == 1) { const int i = D.evIdx[e]; ev[kEvRec * e + 6] = world_z(B.frame, P.M.mkx[i], P.M.mky[i], P.M.mkz[i]); }
      __syncthreads();
    }
    device_step<Mixed, kTPB>(pa, lds, tid);
    ++done;
    {
      SSA_VIEWS;  // outside the branch: the addresses stay uniform
      if (tid != 0) goto next;
      SSL double* rdV = T.sig;
      SSL double* rdSet = rdV + nT;
      SSL double* ev = rdSet + nT;
      for (int j = 0; j < nT; ++j)
        if (D.rdStep[j] == done) { rdV[j] = (D.rdKind[j] == 0 ? B.scal[S_PITCH] : B.scal[S_ROLL]) * 180 / kPi; rdSet[j] = 1; }
      for (int e = 0; e < nE; ++e) {
        SSL double* r = ev + kEvRec * e;
        const int kind = D.evKind[e];
        double s, code = 0;
        if (kind == 0) {
          s = B.over[D.evIdx[e]] - P.M.monThreshold;
        } else if (kind == 1) {
          s = P.M.mkWhen[D.evIdx[e]] == 0 ? -r[6] : r[6];
        } else {
          // the founder rule's largest margin (events.hpp:32-37)
          const double pitch = B.scal[S_PITCH];
          s = P.M.hullTopBelow - B.scal[O_TOP];
          const double sp = __builtin_fabs(pitch) - P.M.pitchAbsAbove;
          if (sp > s) { s = sp; code = pitch < 0 ? -1 : 1; }
          if (B.scal[S_T] > P.M.minT) {
            const double sb = P.M.buoyancyBelow - B.scal[O_VB];
            if (sb > s) { s = sb; code = 2; }
          }
        }
        if (r[0] == 0 && s > 0) {
          r[0] = 1; r[1] = static_cast<double>(done); r[2] = s; r[3] = r[4]; r[5] = B.scal[S_T]; r[8] = code; r[9] = r[7];
        }
        r[4] = s;
        r[7] = code;
      }
    }
  next:
    __syncthreads();
  }

  {
    SSA_VIEWS;
    // the state after the last step: K = 1 stores every state; a portable32 gradient keeps the last one as the end
    // of the last segment's check
    if (K == 1) save_state(P, B, ckpt + done * SS, tid, kTPB);
    else if (Mixed && !FWD && D.wantGrad) save_state(P, B, ckpt + ((done + K - 1) / K) * SS, tid, kTPB);
    // 2. loss, signals, seeds
    if (tid == 0) {
      const SSL double* rdV = T.sig;
      const SSL double* rdSet = rdV + nT;
      const SSL double* ev = rdSet + nT;
      const double tEnd = B.scal[S_T], dt = P.M.dt;
      const bool foundered = B.scal[S_FOUNDERED] != 0;
      SSG double* sigOut = D.signals ? D.signals + static_cast<size_t>(inst) * nSig : nullptr;
      gcd* obs = D.obs + static_cast<size_t>(inst) * nSig;
      double L = 0;
      int ns = 0;
      auto seed = [&](long long step, int kind, int idx, double value) {
        SSL double* sd = T.seeds + 4 * ns++;
        sd[0] = static_cast<double>(step); sd[1] = kind; sd[2] = idx; sd[3] = value;
      };
      for (int j = 0; j < nT; ++j) {
        const double v = rdSet[j] != 0 ? rdV[j] : D.rdAfter[j];
        const double r = v - obs[j];
        L += D.rdW[j] * (r * r);
        if (sigOut) sigOut[j] = v;
        if (rdSet[j] != 0) seed(D.rdStep[j], D.rdKind[j] == 0 ? 0 : 3, 0, 2 * D.rdW[j] * r * 180 / kPi);
      }
      // a founder margin's seed: hull top (s = hullTopBelow - top), pitch (s = |pitch| - pitchAbsAbove), buoyancy
      auto margin_seed = [&](long long step, double code, double g) {
        if (code == 0) seed(step, 4, 0, -g);
        else if (code == 2) seed(step, 5, 0, -g);
        else seed(step, 0, 0, g * code);
      };
      for (int e = 0; e < nE; ++e) {
        const SSL double* r = ev + kEvRec * e;
        const bool fired = r[0] != 0;
        const long long k = static_cast<long long>(r[1]);
        double ts;
        if (fired && k > 1) ts = r[5] - dt * r[2] / (r[2] - r[3]);
        else if (fired) ts = r[5];
        else ts = foundered ? tEnd : D.evMissing[e];
        const double res = ts - obs[nT + e];
        L += D.evW[e] * (res * res);
        if (sigOut) sigOut[nT + e] = ts;
        if (fired && k > 1) {
          const double Gt = 2 * D.evW[e] * res, sk = r[2], sp = r[3], del = sk - sp;
          const double gs = Gt * dt * sp / (del * del), gsp = -Gt * dt * sk / (del * del);
          const int kind = D.evKind[e], idx = D.evIdx[e];
          if (kind == 0) { seed(k, 1, idx, gs); seed(k - 1, 1, idx, gsp); }
          else if (kind == 1) {
            const double sign = P.M.mkWhen[idx] == 0 ? -1.0 : 1.0;
            seed(k, 2, idx, sign * gs); seed(k - 1, 2, idx, sign * gsp);
          } else { margin_seed(k, r[8], gs); margin_seed(k - 1, r[9], gsp); }
        }
      }
      // latest step first (insertion sort: a few dozen records), the order the reverse sweep consumes them in
      for (int a = 1; a < ns; ++a)
        for (int b = a; b > 0 && T.seeds[4 * (b - 1)] < T.seeds[4 * b]; --b)
          for (int q = 0; q < 4; ++q) { const double t = T.seeds[4 * b + q]; T.seeds[4 * b + q] = T.seeds[4 * (b - 1) + q]; T.seeds[4 * (b - 1) + q] = t; }
      T.nseeds[0] = ns;
      D.loss[inst] = L;
      D.steps[inst] = done;
    }
    __syncthreads();
  }
  if (FWD || !D.wantGrad) return;

  // 3. reverse, from the last step to the first. K = 1: every start state is stored. K > 1: when the sweep enters a
  // segment of K steps, the segment is recomputed from its checkpoint into the ring of K + 1 states.
  {
    SSA_VIEWS;
    for (int i = tid; i < 6 + 5 * NN; i += kTPB) { T.lamA[i] = 0; T.lamB[i] = 0; }
    for (int p = tid; p < AM.np; p += kTPB) T.G[p] = 0;
    const size_t iC = static_cast<size_t>(Bt.connOff[inst]);
    load_adjoint_tables<kTPB>(P, T, D.pa + iC, D.pc + iC, tid);
  }
  long long segLo = done;  // the first step of the segment in the ring (none yet)
  for (long long k = done; k >= 1; --k) {
    const SSG double* src = ckpt + (k - 1) * SS;  // the step's start state
    const SSG double* next = ckpt + k * SS;       // and the state after it
    if (K > 1) {
      if (k - 1 < segLo) {
        const long long seg = (k - 1) / K;
        segLo = seg * K;
        const long long len = std::min(K, done - segLo);
        {
          SSA_VIEWS;
          restore_state(P, B, ckpt + seg * SS, tid, kTPB);
          if (tid == 0) *B.nextGroup = 0;
          __syncthreads();
          if (tid < 64) pose_wave0(P, B, tid, B.scal[S_ZO], B.scal[S_PITCH], B.scal[S_ROLL]);
          __syncthreads();
        }
        for (long long i = 0; i < len; ++i) {
          {
            SSA_VIEWS;
            save_state(P, B, ring + i * SS, tid, kTPB);
          }
          __syncthreads();
          device_step<Mixed, kTPB>(pa, lds, tid);
        }
        {
          SSA_VIEWS;
          save_state(P, B, ring + len * SS, tid, kTPB);
        }
        __threadfence_block();
        __syncthreads();
        if (Mixed) {
          // the recomputed segment ends where the forward's next checkpoint (or its last state) begins
          const SSG double* want = ckpt + (seg + 1) * SS;
          const SSG double* got = ring + len * SS;
          for (int j = tid; j < SS; j += kTPB)
            if (__builtin_bit_cast(uint64_t, got[j]) != __builtin_bit_cast(uint64_t, want[j])) atomicAdd((int32_t*)&bad, 1);
        }
      }
      src = ring + (k - 1 - segLo) * SS;
      next = src + SS;
    }
    // the adjoint of the state after step k is in lamA when done - k is even
    const bool inA = ((done - k) & 1) == 0;
    const ViewAddrs va{pa, ta, reinterpret_cast<uint64_t>(gslice)};
    {
      SSA_VIEWS;
      load_tape_state(P, T, src, tid, kTPB);
    }
    __syncthreads();
    jac_step<kTPB, GT>(va, AM, tid);
    if (!Mixed) {
      SSA_VIEWS;
      verify_step(P, T, next, tid, kTPB, (SSL int32_t*)&bad);
    }
    reverse_step<kTPB, GT>(va, AM, tid, k, inA);
  }
  __syncthreads();
  {
    SSA_VIEWS;
    if (tid < AM.np) {
      const double th = D.theta[static_cast<size_t>(inst) * AM.np + tid];
      D.grad[static_cast<size_t>(inst) * AM.np + tid] = T.G[tid] / th;
    }
  }
  if (tid == 0) D.mismatch[inst] = bad;
}

// One step of instance 0 from a given state record, with a given adjoint of the state after it and given seeds on
// the step's monitor sums and mark world z: the adjoint of the state before it, the parameters' log-derivatives and
// the recomputed state after the step (the test hook of the step's reverse).
template <int kTPB, bool GT>
__global__ __launch_bounds__(kTPB) void adjoint_one_step(const ModelParams* __restrict__ models, const AdjModel* __restrict__ amodels,
                                                          const DevBatch* __restrict__ batch, const DevAdj* __restrict__ adj) {
  BC& Bt = *(BC*)batch;
  DC& D = *(DC*)adj;
  extern __shared__ double lds_raw[];
  SSL double* lds = (SSL double*)lds_raw;
  const int tid = threadIdx.x;
  const int inst = 0;
  const int mi = Bt.instModel[inst];
  PC& P = *(PC*)((const SSC char*)models + static_cast<size_t>(mi) * sizeof(ModelParams));
  AMC& AMs = *(AMC*)((const SSC char*)amodels + static_cast<size_t>(mi) * sizeof(AdjModel));
  const AdjScalars AM{AMs.maxGroup, AMs.np, AMs.mode};
  const Blk B = carve(lds, P);
  const TapeT<GT> T = carve_tape<GT>(lds + P.L.total, D.gtape, AMs.A);
  const int NN = P.S.nn;
  load_instance(P, Bt, D, B, inst, tid, kTPB);
  // the record (t zO pitch roll vz wth wph inflow Vb hullTop, vol level cx cy cz) into the tape's input slots
  for (int i = tid; i < 7 + 5 * NN; i += kTPB) {
    const double v = D.stepRec[i < 7 ? i : i + 3];
    if (i < 7) T.sv[i] = v;
    else {
      const int j = i - 7, a = j / NN, n = j - a * NN;
      SSL double* d = a == 0 ? T.iVol : a == 1 ? T.iLvl : a == 2 ? T.iCx : a == 3 ? T.iCy : T.iCz;
      d[n] = v;
    }
  }
  for (int i = tid; i < 6 + 5 * NN; i += kTPB) T.lamA[i] = D.stepLam[i < 6 ? i + 1 : i + 4];
  for (int p = tid; p < AM.np; p += kTPB) T.G[p] = 0;
  if (tid == 0) {
    int ns = 0;
    for (int m = 0; m < P.M.nmon; ++m)
      if (D.stepMon[m] != 0) { SSL double* s = T.seeds + 4 * ns++; s[0] = 1; s[1] = 1; s[2] = m; s[3] = D.stepMon[m]; }
    for (int m = 0; m < P.M.nmark; ++m)
      if (D.stepMark[m] != 0) { SSL double* s = T.seeds + 4 * ns++; s[0] = 1; s[1] = 2; s[2] = m; s[3] = D.stepMark[m]; }
    if (D.stepLam[8] != 0) { SSL double* s = T.seeds + 4 * ns++; s[0] = 1; s[1] = 5; s[2] = 0; s[3] = D.stepLam[8]; }
    if (D.stepLam[9] != 0) { SSL double* s = T.seeds + 4 * ns++; s[0] = 1; s[1] = 4; s[2] = 0; s[3] = D.stepLam[9]; }
    T.nseeds[0] = ns;
  }
  __syncthreads();
  load_adjoint_tables<kTPB>(P, T, D.pa, D.pc, tid);
  jac_step_views<kTPB>(P, AM, B, T, tid);
  reverse_step_views<kTPB>(P, AM, B, T, tid, 1, T.lamA, T.lamB);
  for (int i = tid; i < 10 + 5 * NN; i += kTPB) {
    double lam = 0, rec = 0;
    if (i >= 1 && i <= 6) lam = T.lamB[i - 1];
    else if (i >= 10) lam = T.lamB[i - 4];
    const SSL double* o = T.tot;
    if (i == 0) rec = T.sv[0] + P.M.dt;
    else if (i <= 6) rec = i == 1 ? o[T_ZO1] : i == 2 ? o[T_P1] : i == 3 ? o[T_R1] : i == 4 ? o[T_VZ1] : i == 5 ? o[T_WTH1] : o[T_WPH1];
    else if (i == 8) rec = o[T_V];
    else if (i >= 10) {
      const int j = i - 10, a = j / NN, n = j - a * NN;
      const SSL double* s = a == 0 ? T.vol1 : a == 1 ? T.oLvl : a == 2 ? T.oCx : a == 3 ? T.oCy : T.oCz;
      rec = s[n];
    }
    D.stepLamOut[i] = lam;
    D.stepRecOut[i] = rec;
  }
  for (int p = tid; p < AM.np; p += kTPB) D.grad[p] = T.G[p];
}

}  // namespace ssa

namespace {

// LDS layout of the adjoint's tape, in doubles (after the forward layout).
// global: the Jacobian rows and the connections' partials and gradient terms go to a per-instance global slice.
// forwardOnly: only what the forward and the loss use (the signals and the seed records).
AdjLayout adj_layout(int32_t NN, int32_t NC, int32_t NG, int32_t nmon, int32_t maxGroup, int32_t np, int32_t nseeds, int32_t nsig,
                     bool global, bool forwardOnly = false) {
  AdjLayout L{};
  int32_t off = 0;
  auto take = [&](int32_t doubles) { const int32_t at = off; off += std::max<int32_t>(doubles, 1); return at; };
  if (forwardOnly) {
    L.seeds = take(4 * nseeds); L.nseeds = take(1); L.sig = take(nsig);
    L.total = off;
    return L;
  }
  L.sv = take(8); L.iVol = take(NN); L.iLvl = take(NN); L.iCx = take(NN); L.iCy = take(NN); L.iCz = take(NN);
  L.fr = take(F_COUNT); L.dfp = take(kFrameDeriv * 3); L.bw = take(kBw * kUnits); L.bt = take(20);
  L.oLvl = take(NN); L.oVol = take(NN); L.oArea = take(NN); L.oCx = take(NN); L.oCy = take(NN); L.oCz = take(NN); L.oZmax = take(NN);
  L.jac = global ? 0 : take(NN * kSolveOut * kDof); L.carry = take(NN); L.gMer = take(NG); L.gH0 = take(NG);
  L.mp = take((kMaxTPB / kLanes) * maxGroup * kPassRec);
  L.hEff = take(NN); L.aEff = take(NN); L.aSrc = take(NN); L.deg = take(NN); L.dV = take(NC); L.cp = global ? 0 : take(6 * NC);
  L.acc = take(NN); L.vol1 = take(NN); L.clip = take(NN); L.tot = take(T_COUNT);
  L.lamA = take(6 + 5 * NN); L.lamB = take(6 + 5 * NN); L.G = take(np); L.gOut = take(kSolveOut * NN); L.gVol1 = take(NN);
  L.gAcc = take(NN); L.cg = global ? 0 : take(5 * NC); L.red = take(3 * (kMaxTPB / 64)); L.seeds = take(4 * nseeds); L.nseeds = take(1);
  L.sig = take(nsig);
  L.cpa = take(NC); L.cpc = take(NC); L.cmon = take(NC); L.monSeed = take(nmon); L.sflag = take(2);
  L.gnext = take(1);
  L.total = off;
  if (global) {
    L.gJac = 0;
    L.gCp = NN * kSolveOut * kDof;
    L.gCg = L.gCp + 6 * NC;
    L.gTotal = L.gCg + 5 * NC;
  }
  return L;
}

}  // namespace

namespace ssa {
namespace {

// The gradient run's device buffers and views, for one call.
struct AdjRun {
  Dev<AdjModel> dAdjModels;
  Dev<DevAdj> dAdj;
  Dev<double> theta, ca, rdW, rdAfter, evW, evMissing, obs, loss, grad, signals, ckpt, ring, stepRec, stepLam, stepMon, stepMark, stepLamOut,
      stepRecOut;
  Dev<int32_t> pa, pc, rdKind, evKind, evIdx, mismatch;
  Dev<long long> rdStep, steps, ckptOff, ringOff;
  Dev<double> gtape;
  size_t lds = 0;
  int tpb = kThreads;
  bool global = false;  // the tape's three largest arrays in global memory
  bool forwardOnly = false;  // no reverse sweep: the forward-only kernel and layout
};

struct Setup {
  int32_t np;
  const int32_t *pa, *pc;
  const double* theta;
  int32_t relaxed;
};

// Checks, uploads the batch and the adjoint's tables; returns the launch's LDS size.
// wantGrad false: the forward alone (one checkpoint per instance, no ring).
void prepare(ss_batch* b, AdjRun& R, const Setup& s, const ss_loss* loss, const double* obs, int32_t& K, double tMax, long long maxSteps,
             bool oneStep, bool wantGrad) {
  const size_t count = static_cast<size_t>(b->opt.count);
  need(s.np >= 1 && s.np <= kMaxParams, "ss_batch_gradient: the parameter count must be 1..64");
  need(s.pa != nullptr && s.pc != nullptr && s.theta != nullptr, "ss_batch_gradient: null parameter table");

  need(!(oneStep && b->mixed), "ss_batch_adjoint_step: the one-step form is FP64 (portable numerics)");
  const int32_t nT = loss ? loss->nRead : 0, nE = loss ? loss->nEvent : 0;
  need(nT >= 0 && nE >= 0 && nT + nE <= 256, "ss_batch_gradient: at most 256 loss terms");
  const size_t nconn = b->hArea.size();
  for (size_t k = 0; k < nconn; ++k) need(s.pa[k] >= -1 && s.pa[k] < s.np && s.pc[k] >= -1 && s.pc[k] < s.np, "ss_batch_gradient: parameter index out of range");
  for (size_t k = 0; k < count * static_cast<size_t>(s.np); ++k) need(s.theta[k] > 0 && std::isfinite(s.theta[k]), "ss_batch_gradient: parameters must be positive");
  for (int32_t j = 0; j < nT; ++j) need(loss->rdKind[j] == 0 || loss->rdKind[j] == 1, "ss_batch_gradient: readout kind must be 0 (trim) or 1 (list)");
  for (int32_t e = 0; e < nE; ++e) {
    need(loss->evKind[e] >= 0 && loss->evKind[e] <= 2, "ss_batch_gradient: event kind must be 0 (monitor), 1 (mark) or 2 (founder)");
    if (loss->evKind[e] == 2) continue;
    for (size_t i = 0; i < count; ++i) {
      const Model& m = b->model(static_cast<int32_t>(i));
      need(loss->evIdx[e] >= 0 && loss->evIdx[e] < (loss->evKind[e] == 0 ? m.nmon : m.nmark), "ss_batch_gradient: event index out of range");
    }
  }
  // the scaled products coef * theta[pc] * area * theta[pa], formed as the sweep forms area * scale and the flow law
  // multiplies coef by area
  std::vector<double> ca(nconn);
  for (size_t i = 0; i < count; ++i) {
    const Model& m = b->model(static_cast<int32_t>(i));
    const size_t o = static_cast<size_t>(b->connOff[i]);
    const double* th = s.theta + i * static_cast<size_t>(s.np);
    for (int32_t c = 0; c < m.NC; ++c) {
      const size_t k = o + static_cast<size_t>(c);
      const double area = s.pa[k] >= 0 ? b->hArea[k] * th[s.pa[k]] : b->hArea[k];
      const double coef = s.pc[k] >= 0 ? m.coef0[static_cast<size_t>(c)] * th[s.pc[k]] : m.coef0[static_cast<size_t>(c)];
      ca[k] = coef * area;
    }
  }
  // the launch shape, the first whose LDS fits every model: 512 threads (buoyancy partials beside the level solve, 16
  // solver units) with the whole tape in LDS, 256 threads with the whole tape in LDS, then the same two with the
  // tape's three largest arrays in global memory; the models are uploaded with that size's forward layout
  const int32_t nseeds = nT + 2 * nE + 64 + 64 + 2;  // the one-step form seeds every monitor and mark, V and the top
  auto need_lds = [&](int tpb, bool global) {
    size_t l = 0;
    for (const auto& m : b->models) {
      const AdjLayout A = adj_layout(m->NN, m->NC, m->NG, m->nmon, m->maxGroup, s.np, nseeds, 2 * nT + 10 * nE, global);
      l = std::max(l, (static_cast<size_t>(m->layout(tpb).total) + static_cast<size_t>(A.total)) * sizeof(double));
    }
    return l;
  };
  if (wantGrad) {
    const int shapes[4] = {kMaxTPB, kThreads, kMaxTPB, kThreads};
    int pick = -1;
    for (int k = 0; k < 4 && pick < 0; ++k)
      if (need_lds(shapes[k], k >= 2) <= lds_limit()) pick = k;
    if (pick < 0)
      fail(SS_RUNTIME, "ss_batch_gradient: " + std::to_string(need_lds(kThreads, true)) + " bytes of LDS per simulation exceed the device's limit");
    R.tpb = shapes[pick];
    R.global = pick >= 2;
  } else {
    // the forward alone, with only the loss's records beside the forward layout: 512 threads while the batch leaves
    // compute units without an instance (each instance's latency is the run's), else 256 (more resident instances)
    int cus = 0;
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0), "hipDeviceGetAttribute");
    size_t wideLds = 0;
    for (const auto& m : b->models) {
      const AdjLayout A = adj_layout(m->NN, m->NC, m->NG, m->nmon, m->maxGroup, s.np, nseeds, 2 * nT + 10 * nE, false, true);
      wideLds = std::max(wideLds, (static_cast<size_t>(m->layout(kMaxTPB).total) + static_cast<size_t>(A.total)) * sizeof(double));
    }
    R.tpb = count <= static_cast<size_t>(cus) && wideLds <= lds_limit() ? kMaxTPB : kThreads;
    R.global = false;
    R.forwardOnly = true;
  }
  const int32_t keepTpb = b->tpb;
  b->tpb = R.tpb;
  try {
    upload(b);
  } catch (...) {
    b->tpb = keepTpb;
    throw;
  }
  b->tpb = keepTpb;
  b->ldsBytes = forward_lds_bytes(*b);  // the forward kernel's layout, as the accessor reports it
  b->stepped = false;                   // the device now holds the staged instances again
  R.ca.upload(ca);
  R.theta.upload(s.theta, count * static_cast<size_t>(s.np));
  R.pa.upload(s.pa, nconn);
  R.pc.upload(s.pc, nconn);
  // per model: the tape layout
  std::vector<AdjModel> am;
  size_t lds = 0;
  long long gStride = 0;
  for (const auto& m : b->models) {
    need(m->maxGroup <= 8, "ss_batch_gradient: free-surface groups of more than 8 spaces are not supported");
    need(m->nmon <= 64 && m->nmark <= 64, "ss_batch_gradient: at most 64 monitors and 64 marks");
    AdjModel a{};
    a.A = adj_layout(m->NN, m->NC, m->NG, m->nmon, m->maxGroup, s.np, nseeds, 2 * nT + 10 * nE, R.global, R.forwardOnly);
    gStride = std::max<long long>(gStride, a.A.gTotal);
    a.maxGroup = m->maxGroup;
    a.np = s.np;
    a.mode = s.relaxed ? 1 : 0;
    am.push_back(a);
    lds = std::max(lds, (static_cast<size_t>(m->layout(R.tpb).total) + static_cast<size_t>(a.A.total)) * sizeof(double));
  }
  if (lds > lds_limit())
    fail(SS_RUNTIME, "ss_batch_gradient: " + std::to_string(lds) + " bytes of LDS per simulation exceed the device's limit");
  R.lds = lds;
  b->adjTpb = R.tpb;
  b->adjGlobal = R.global ? 1 : 0;
  b->adjLds = lds;
  R.dAdjModels.upload(am);
  // each instance's step bound and state size; K = 0 chooses 1 (every state stored, no recomputation) when all of them
  // fit in a third of the free device memory, else 256
  std::vector<long long> caps(count), sizes(count);
  double every = 0;
  for (size_t i = 0; i < count; ++i) {
    const Model& m = b->model(static_cast<int32_t>(i));
    sizes[i] = 7 + 5 * static_cast<long long>(m.NN);
    const double span = std::max(0.0, tMax - b->hSt[i * 7]);
    const double bound = span / m.net.dt + 2;
    long long cap = maxSteps > 0 ? maxSteps : LLONG_MAX;
    if (bound < 1e9) cap = std::min(cap, static_cast<long long>(std::ceil(bound)));
    need(cap <= 100000000LL, "ss_batch_gradient: the run is unbounded (set tMax or maxSteps to at most 1e8 steps)");
    caps[i] = cap;
    every += static_cast<double>(cap + 1) * static_cast<double>(sizes[i]) * sizeof(double);
  }
  if (!wantGrad) {
    K = INT32_MAX;
  } else if (K == 0) {
    size_t freeB = 0, totalB = 0;
    hip_check(hipMemGetInfo(&freeB, &totalB), "hipMemGetInfo");
    K = every <= static_cast<double>(freeB) / 3 ? 1 : 256;
  }
  std::vector<long long> ckOff(count), rgOff(count);
  long long ck = 0, rg = 0;
  for (size_t i = 0; i < count; ++i) {
    ckOff[i] = ck;
    rgOff[i] = rg;
    if (!oneStep) {
      // a portable32 gradient also keeps the last state (the last segment's check)
      ck += (caps[i] / K + (b->mixed && wantGrad && K > 1 ? 2 : 1)) * sizes[i];
      if (K > 1 && wantGrad) rg += (static_cast<long long>(K) + 1) * sizes[i];
    }
  }
  R.ckpt.alloc(static_cast<size_t>(std::max(ck, 1LL)));
  R.ring.alloc(static_cast<size_t>(std::max(rg, 1LL)));
  R.ckptOff.upload(ckOff);
  R.ringOff.upload(rgOff);
  R.gtape.alloc(R.global ? static_cast<size_t>(gStride) * (oneStep ? 1 : count) : 1);
  if (nT > 0) {
    R.rdStep.upload(reinterpret_cast<const long long*>(loss->rdStep), static_cast<size_t>(nT));
    R.rdKind.upload(loss->rdKind, static_cast<size_t>(nT));
    R.rdW.upload(loss->rdW, static_cast<size_t>(nT));
    R.rdAfter.upload(loss->rdAfter, static_cast<size_t>(nT));
  }
  if (nE > 0) {
    R.evKind.upload(loss->evKind, static_cast<size_t>(nE)); R.evIdx.upload(loss->evIdx, static_cast<size_t>(nE));
    R.evW.upload(loss->evW, static_cast<size_t>(nE)); R.evMissing.upload(loss->evMissing, static_cast<size_t>(nE));
  }
  const size_t nSig = static_cast<size_t>(nT + nE);
  if (nSig > 0) {
    if (obs) R.obs.upload(obs, count * nSig);
    else R.obs.fill(0.0, count * nSig);
  }
  R.loss.fill(0.0, count);
  R.grad.fill(0.0, std::max<size_t>(count * static_cast<size_t>(s.np), static_cast<size_t>(s.np)));
  R.signals.fill(0.0, std::max<size_t>(count * nSig, 1));
  R.steps.fill(0, count);
  R.mismatch.fill(0, count);
  DevAdj d{};
  d.K = K; d.wantGrad = wantGrad ? 1 : 0; d.maxSteps = maxSteps > 0 ? maxSteps : LLONG_MAX; d.tMax = tMax;
  d.theta = G(R.theta.p); d.pa = G(R.pa.p); d.pc = G(R.pc.p); d.ca = G(R.ca.p);
  d.nRead = nT; d.nEv = nE;
  d.rdStep = (const SSG long long*)R.rdStep.p; d.rdKind = G(R.rdKind.p); d.rdW = G(R.rdW.p); d.rdAfter = G(R.rdAfter.p);
  d.evKind = G(R.evKind.p); d.evIdx = G(R.evIdx.p); d.evW = G(R.evW.p); d.evMissing = G(R.evMissing.p); d.obs = G(R.obs.p);
  d.loss = W(R.loss.p); d.grad = W(R.grad.p); d.signals = W(R.signals.p); d.steps = W(R.steps.p); d.mismatch = W(R.mismatch.p);
  d.ckpt = W(R.ckpt.p); d.ring = W(R.ring.p);
  d.ckptOff = (const SSG long long*)R.ckptOff.p; d.ringOff = (const SSG long long*)R.ringOff.p;
  d.gtape = W(R.gtape.p);
  d.gStride = R.global ? gStride : 0;
  R.dAdj.upload(&d, 1);
}

using AdjKernel = void (*)(const ModelParams*, const AdjModel*, const DevBatch*, const DevAdj*);

void launch(bool oneStep, ss_batch* b, AdjRun& R, int grid, double* ms) {
  const bool wide = R.tpb == kMaxTPB;
  auto run = [&](auto mixed) -> AdjKernel {
    constexpr bool M = decltype(mixed)::value;
    return R.forwardOnly ? (wide ? adjoint_run<kMaxTPB, false, true, M> : adjoint_run<kThreads, false, true, M>)
           : (R.global ? (wide ? adjoint_run<kMaxTPB, true, false, M> : adjoint_run<kThreads, true, false, M>)
                       : (wide ? adjoint_run<kMaxTPB, false, false, M> : adjoint_run<kThreads, false, false, M>));
  };
  const AdjKernel kernel =
      oneStep ? (R.global ? (wide ? adjoint_one_step<kMaxTPB, true> : adjoint_one_step<kThreads, true>)
                          : (wide ? adjoint_one_step<kMaxTPB, false> : adjoint_one_step<kThreads, false>))
              : b->mixed ? run(std::true_type{}) : run(std::false_type{});
  hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(R.lds)),
            "hipFuncSetAttribute");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  hip_check(hipEventRecord(e0, nullptr), "hipEventRecord");
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(R.tpb), R.lds, nullptr, static_cast<const ModelParams*>(b->dModels.p),
                     static_cast<const AdjModel*>(R.dAdjModels.p), static_cast<const DevBatch*>(b->dBatch.p),
                     static_cast<const DevAdj*>(R.dAdj.p));
  hip_check(hipGetLastError(), "adjoint launch");
  hip_check(hipEventRecord(e1, nullptr), "hipEventRecord");
  hip_check(hipEventSynchronize(e1), "adjoint kernel");
  float t = 0;
  hip_check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = t;
}

}  // namespace
}  // namespace ssa

extern "C" {

ss_status ss_batch_gradient(ss_batch* b, int32_t nparams, const int32_t* pa, const int32_t* pc, const double* theta, const ss_loss* loss,
                            const double* observations, const ss_gradient_options* opt, double* lossOut, double* grad, double* signals,
                            int64_t* steps, int32_t* mismatches, double* kernelMs) {
  return guarded([&] {
    need(b != nullptr && opt != nullptr, "ss_batch_gradient: null argument");
    need(lossOut != nullptr, "ss_batch_gradient: null loss output");
    need(opt->checkpointEvery >= 0 && opt->checkpointEvery <= 65536, "ss_batch_gradient: checkpointEvery must be 0..65536");
    need(opt->relaxed == 0 || opt->relaxed == 1, "ss_batch_gradient: relaxed must be 0 or 1");
    const double tMax = opt->tMax > 0 ? opt->tMax : b->opt.tMax;
    AdjRun R;
    int32_t K = opt->checkpointEvery;
    prepare(b, R, Setup{nparams, pa, pc, theta, opt->relaxed}, loss, observations, K, tMax, opt->maxSteps, false, grad != nullptr);
    double ms = 0;
    launch(false, b, R, b->opt.count, &ms);
    const size_t count = static_cast<size_t>(b->opt.count);
    R.loss.download(lossOut, count);
    if (grad) R.grad.download(grad, count * static_cast<size_t>(nparams));
    const size_t nSig = static_cast<size_t>(loss ? loss->nRead + loss->nEvent : 0);
    if (signals && nSig) R.signals.download(signals, count * nSig);
    if (steps) {
      std::vector<long long> s(count);
      R.steps.download(s.data(), count);
      for (size_t i = 0; i < count; ++i) steps[i] = s[i];
    }
    if (mismatches) R.mismatch.download(mismatches, count);
    if (kernelMs) *kernelMs = ms;
  });
}

ss_status ss_batch_adjoint_step(ss_batch* b, int32_t nparams, const int32_t* pa, const int32_t* pc, const double* theta, int32_t relaxed,
                                const double* record, const double* lambda, const double* monitorSeed, const double* markSeed, double* lambdaOut,
                                double* gradLog, double* recordOut) {
  return guarded([&] {
    need(b != nullptr && record != nullptr && lambda != nullptr && lambdaOut != nullptr && gradLog != nullptr && recordOut != nullptr,
         "ss_batch_adjoint_step: null argument");
    AdjRun R;
    int32_t K = 1;
    prepare(b, R, Setup{nparams, pa, pc, theta, relaxed}, nullptr, nullptr, K, b->opt.tMax, 1, true, true);
    const Model& m = b->model(0);
    const size_t rs = 10 + 5 * static_cast<size_t>(m.NN);
    R.stepRec.upload(record, rs);
    R.stepLam.upload(lambda, rs);
    std::vector<double> mon(static_cast<size_t>(std::max(m.nmon, 1)), 0.0), mark(static_cast<size_t>(std::max(m.nmark, 1)), 0.0);
    if (monitorSeed) std::copy(monitorSeed, monitorSeed + m.nmon, mon.begin());
    if (markSeed) std::copy(markSeed, markSeed + m.nmark, mark.begin());
    R.stepMon.upload(mon);
    R.stepMark.upload(mark);
    R.stepLamOut.fill(0.0, rs);
    R.stepRecOut.fill(0.0, rs);
    DevAdj d{};
    R.dAdj.download(&d, 1);
    d.stepRec = G(R.stepRec.p); d.stepLam = G(R.stepLam.p); d.stepMon = G(R.stepMon.p); d.stepMark = G(R.stepMark.p);
    d.stepLamOut = W(R.stepLamOut.p); d.stepRecOut = W(R.stepRecOut.p);
    R.dAdj.upload(&d, 1);
    launch(true, b, R, 1, nullptr);
    R.stepLamOut.download(lambdaOut, rs);
    R.stepRecOut.download(recordOut, rs);
    R.grad.download(gradLog, static_cast<size_t>(nparams));
  });
}

ss_status ss_batch_gradient_launch(const ss_batch* b, int32_t* threads, int32_t* globalTape, int64_t* ldsBytes) {
  return guarded([&] {
    need(b != nullptr, "ss_batch_gradient_launch: null batch");
    if (b->adjTpb == 0) fail(SS_LOGIC, "ss_batch_gradient_launch: no gradient call yet");
    if (threads) *threads = b->adjTpb;
    if (globalTape) *globalTape = b->adjGlobal;
    if (ldsBytes) *ldsBytes = static_cast<int64_t>(b->adjLds);
  });
}

}  // extern "C"
