// libsinksim_hip.so: the batch kernel and the C ABI of include/sinksim_hip.h.
//
// Reference counterpart: engine/cuda/batch.cu (host orchestration) and engine/cuda/device_step.cuh:323-434
// (run_chunk). The host keeps the reference's contract (stage per-instance changes, run to founder or tMax,
// read results, traces and curves) on a design of its own: one persistent launch per chunk with the hull
// prepared once per batch (segment streams pre-gathered, f32 copies for portable32), all per-instance
// results read back in one download, and the closing readouts computed on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sinksim_hip.h"
#include "step_dev.hpp"

namespace ssa {

constexpr int kCurveValues = 7;

// ------------------------------------------------------------------------------------------- the kernel

// Writes the complete state record (the trace layout, 10 + 5 * nodes values), one thread.
__device__ __noinline__ void write_record(SSG double* rec, const Blk& B, int nn) {
  for (int k = 0; k < 7; ++k) rec[k] = B.scal[S_T + k];
  rec[7] = B.scal[O_INFLOW]; rec[8] = B.scal[O_VB]; rec[9] = B.scal[O_TOP];
  SSG double* p = rec + 10;
  for (int n = 0; n < nn; ++n) p[n] = B.vol[n];
  p += nn;
  for (int n = 0; n < nn; ++n) p[n] = B.level[n];
  p += nn;
  for (int n = 0; n < nn; ++n) p[n] = B.cx[n];
  p += nn;
  for (int n = 0; n < nn; ++n) p[n] = B.cy[n];
  p += nn;
  for (int n = 0; n < nn; ++n) p[n] = B.cz[n];
}

// One workgroup per instance: load the instance into LDS, step it until it founders, reaches tMax or has taken
// maxSteps steps in this launch, store it back. Traces and curves are recorded as in the reference
// (device_step.cuh:379-416); the readouts of the state at the end of the launch go to final7.
template <bool Mixed>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4, 8))) void run_chunk(const Params* __restrict__ params, long long maxSteps, double tMax, int firstLaunch) {
  PC& P = *(PC*)params;
  extern __shared__ double lds_raw[];
  SSL double* lds = (SSL double*)lds_raw;
  const int inst = blockIdx.x;
  const int tid = threadIdx.x;
  const int NN = P.S.nn, NC = P.M.nc, evCap = P.M.evCap, nmon = P.M.nmon, nmark = P.M.nmark;
  const size_t iN = static_cast<size_t>(inst) * NN, iC = static_cast<size_t>(inst) * NC;
  const Blk B = carve(lds, P);

  // load the instance
  if (tid < 7) {
    B.scal[S_T + tid] = P.Bt.st[inst * 7 + tid];
    B.scal[O_INFLOW + tid] = P.Bt.out[inst * 7 + tid];
  }
  if (tid == 0) {
    B.scal[S_EVCOUNT] = P.Bt.evCount[inst];
    B.scal[S_FOUNDERED] = P.Bt.foundered[inst];
    B.scal[S_FOUNDERT] = P.Bt.founderT[inst];
    B.scal[S_SEA] = 0;
    if (P.Bt.curveEvery > 0) {
      B.scal[S_NEXTCURVE] = P.Bt.nextCurve[inst];
      B.scal[S_CURVECOUNT] = P.Bt.curveCount[inst];
    }
  }
  for (int n = tid; n < NN; n += kThreads) {
    B.vol[n] = P.Bt.vol[iN + n]; B.level[n] = P.Bt.level[iN + n];
    B.cx[n] = P.Bt.cx[iN + n]; B.cy[n] = P.Bt.cy[iN + n]; B.cz[n] = P.Bt.cz[iN + n];
    B.area[n] = 0; B.zmin[n] = 0; B.zmax[n] = 0; B.hEff[n] = 0; B.aEff[n] = 0; B.excess[n] = 0; B.deg[n] = 0;
    B.tX[n] = 0; B.tY[n] = 0; B.tF[n] = 0; B.tM[n] = 0; B.tIp[n] = 0; B.tIr[n] = 0;
  }
  for (int c = tid; c < NC; c += kThreads) {
    B.cen[c] = P.Bt.en[iC + c]; B.carea[c] = P.Bt.area[iC + c]; B.ctOn[c] = P.Bt.tOn[iC + c];
    B.dV[c] = 0; B.q[c] = 0;
  }
  for (int g = tid; g < P.M.ng; g += kThreads) {
    B.merged[g] = 0;
    B.gStart[g] = P.M.gStart[g]; B.gCount[g] = P.M.gCount[g]; B.gMode[g] = P.M.gMode[g];
    B.gpx[g] = P.M.gpx[g]; B.gpy[g] = P.M.gpy[g]; B.gpz[g] = P.M.gpz[g];
  }
  for (int k = tid; k < P.M.nMembers; k += kThreads) B.gNodes[k] = P.M.gNodes[k];
  for (int n = tid; n < NN; n += kThreads) {
    B.nSegStart[n] = P.S.segStart[n]; B.nSegCount[n] = P.S.segCount[n];
    B.nMu[n] = P.S.mu[n]; B.nVmax[n] = P.S.vmax[n]; B.nAov[n] = P.S.aov[n]; B.nAmin[n] = P.S.amin[n];
  }
  for (int e = tid; e < evCap; e += kThreads) {
    const size_t k = static_cast<size_t>(inst) * evCap + e;
    B.evT[e] = P.Bt.evT[k]; B.evId[e] = P.Bt.evId[k]; B.evKind[e] = P.Bt.evKind[k];
  }
  for (int m = tid; m < nmon; m += kThreads) { B.overT[m] = P.Bt.overT[static_cast<size_t>(inst) * nmon + m]; B.over[m] = 0; }
  for (int m = tid; m < nmark; m += kThreads) B.markT[m] = P.Bt.markT[static_cast<size_t>(inst) * nmark + m];
  __syncthreads();
  if (tid < 64) pose_wave0(P, B, tid, B.scal[S_ZO], B.scal[S_PITCH], B.scal[S_ROLL]);  // the loaded state's pose
  __syncthreads();

  const bool traced = inst < P.Bt.traced;
  long long done = P.Bt.steps[inst];
  if (firstLaunch && traced) {
    // the initial record: buoyancy at the initial pose, no inflow yet (simulation.cpp:137-145, 204-206)
    buoyancy_partials<Mixed>(P, B, tid);
    __syncthreads();
    if (tid == 0) {
      double V, bx, by, bz, top;
      finish_buoyancy(B, V, bx, by, bz, top);
      B.scal[O_INFLOW] = 0; B.scal[O_VB] = V; B.scal[O_TOP] = top; B.scal[O_BX] = bx; B.scal[O_BY] = by; B.scal[O_BZ] = bz;
      const size_t rs = static_cast<size_t>(P.Bt.recSize);
      write_record(P.Bt.denseRec + static_cast<size_t>(inst) * (P.Bt.denseSteps + 1) * rs, B, NN);
      write_record(P.Bt.sampleRec + static_cast<size_t>(inst) * P.Bt.maxSamples * rs, B, NN);
      P.Bt.denseCount[inst] = 1; P.Bt.sampleCount[inst] = 1; P.Bt.nextSample[inst] = P.Bt.sampleEvery;
    }
    __syncthreads();
  }

  const bool curves = P.Bt.curveEvery > 0;
  for (long long s = 0; s < maxSteps; ++s) {
    if (B.scal[S_FOUNDERED] != 0 || B.scal[S_T] >= tMax) break;  // uniform: LDS values after a barrier
    // a readout-curve record when due (simulation.cpp:175-176). Every wave evaluates the condition on the same
    // LDS values; the barrier inside orders all of those reads (and the record's reads of the volumes) before
    // thread 0 advances the bookkeeping, so no wave can see the advanced values and skip the barrier.
    if (curves && B.scal[S_T] >= B.scal[S_NEXTCURVE] && B.scal[S_CURVECOUNT] < P.Bt.curveMax) {
      if (tid == 0) {
        SSG double* rec = P.Bt.curveRec + (static_cast<size_t>(inst) * P.Bt.curveMax + static_cast<size_t>(B.scal[S_CURVECOUNT])) * kCurveValues;
        readouts7(P, B, rec);
      }
      __syncthreads();
      if (tid == 0) {
        B.scal[S_CURVECOUNT] += 1;
        B.scal[S_NEXTCURVE] += P.Bt.curveEvery;
      }
      // read again only at the next iteration, after the step's barriers
    }
    device_step<Mixed>(P, B, tid);
    ++done;
    if (traced) {
      if (tid == 0) {
        const size_t rs = static_cast<size_t>(P.Bt.recSize);
        if (done <= P.Bt.denseSteps) {
          const int k = P.Bt.denseCount[inst]++;
          write_record(P.Bt.denseRec + (static_cast<size_t>(inst) * (P.Bt.denseSteps + 1) + k) * rs, B, NN);
        }
        if (B.scal[S_T] >= P.Bt.nextSample[inst] - 1e-9 && P.Bt.sampleCount[inst] < P.Bt.maxSamples) {
          const int k = P.Bt.sampleCount[inst]++;
          write_record(P.Bt.sampleRec + (static_cast<size_t>(inst) * P.Bt.maxSamples + k) * rs, B, NN);
          P.Bt.nextSample[inst] += P.Bt.sampleEvery;
        }
      }
      __syncthreads();
    }
  }

  // store the instance
  if (tid < 7) {
    P.Bt.st[inst * 7 + tid] = B.scal[S_T + tid];
    P.Bt.out[inst * 7 + tid] = B.scal[O_INFLOW + tid];
  }
  if (tid == 0) {
    P.Bt.evCount[inst] = static_cast<int32_t>(B.scal[S_EVCOUNT]);
    P.Bt.foundered[inst] = static_cast<uint8_t>(B.scal[S_FOUNDERED]);
    P.Bt.founderT[inst] = B.scal[S_FOUNDERT];
    P.Bt.steps[inst] = done;
    if (P.Bt.curveEvery > 0) {
      P.Bt.nextCurve[inst] = B.scal[S_NEXTCURVE];
      P.Bt.curveCount[inst] = static_cast<int32_t>(B.scal[S_CURVECOUNT]);
    }
    readouts7(P, B, P.Bt.final7 + static_cast<size_t>(inst) * 7);
  }
  for (int n = tid; n < NN; n += kThreads) {
    P.Bt.vol[iN + n] = B.vol[n]; P.Bt.level[iN + n] = B.level[n];
    P.Bt.cx[iN + n] = B.cx[n]; P.Bt.cy[iN + n] = B.cy[n]; P.Bt.cz[iN + n] = B.cz[n];
  }
  for (int e = tid; e < evCap; e += kThreads) {
    const size_t k = static_cast<size_t>(inst) * evCap + e;
    P.Bt.evT[k] = B.evT[e]; P.Bt.evId[k] = B.evId[e]; P.Bt.evKind[k] = B.evKind[e];
  }
  for (int m = tid; m < nmon; m += kThreads) P.Bt.overT[static_cast<size_t>(inst) * nmon + m] = B.overT[m];
  for (int m = tid; m < nmark; m += kThreads) P.Bt.markT[static_cast<size_t>(inst) * nmark + m] = B.markT[m];
}

// ------------------------------------------------------------------------------------------- host side

namespace {

thread_local std::string g_error;

// device addresses as the kernel's address-space-qualified pointer types
template <class T>
const SSG T* G(const T* p) { return (const SSG T*)p; }
template <class T>
SSG T* W(T* p) { return (SSG T*)p; }

struct Fail {
  ss_status status;
  std::string what;
};

[[noreturn]] void fail(ss_status s, const std::string& what) { throw Fail{s, what}; }

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(SS_RUNTIME, std::string("HIP: ") + what + ": " + hipGetErrorString(e));
}

template <class F>
ss_status guarded(F&& f) {
  try {
    f();
    return SS_OK;
  } catch (const Fail& e) {
    g_error = e.what;
    return e.status;
  } catch (const std::bad_alloc&) {
    g_error = "out of host memory";
    return SS_RUNTIME;
  } catch (const std::exception& e) {
    g_error = e.what();
    return SS_RUNTIME;
  }
}

// A device allocation owned by the batch.
template <class T>
struct Dev {
  T* p = nullptr;
  size_t n = 0;
  Dev() = default;
  Dev(const Dev&) = delete;
  Dev& operator=(const Dev&) = delete;
  ~Dev() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) hip_check(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)), "hipMalloc");
  }
  void upload(const std::vector<T>& v) {
    alloc(v.size());
    if (!v.empty()) hip_check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy to device");
  }
  void upload(const T* src, size_t count) {
    alloc(count);
    if (count) hip_check(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy to device");
  }
  void fill(const T& v, size_t count) { upload(std::vector<T>(count, v)); }
  void download(T* dst, size_t count, size_t offset = 0) const {
    if (count) hip_check(hipMemcpy(dst, p + offset, count * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy to host");
  }
  std::vector<T> all() const {
    std::vector<T> v(n);
    download(v.data(), n);
    return v;
  }
};

int device_count(std::string* why) {
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    if (why) *why = hipGetErrorString(e);
    return 0;
  }
  if (n == 0 && why) *why = "no HIP device";
  return n;
}

}  // namespace
}  // namespace ssa

using namespace ssa;

struct ss_batch {
  ss_batch_options opt{};
  int32_t NN = 0, NC = 0, NG = 0, nmon = 0, nmark = 0, evCap = 0, maxGroup = 1, ncol = 0, nseg = 0;
  bool mixed = true;
  LdsLayout layout{};
  size_t ldsBytes = 0;
  int32_t recSize = 0, maxSamples = 0, curveMax = 0;
  int64_t totalSteps = 0;
  bool ran = false;
  double kernelMs = 0;
  int launches = 0;

  // ship
  Dev<double> col, seg, mu, vmax, amin, aov;  // col, seg: 6 doubles per record
  Dev<float> colF, segF;                      // 8 floats per record
  Dev<int32_t> segStart, segCount;
  // scenario
  Dev<int32_t> gStart, gCount, gNodes, incStart, incCount, sea, monStart, monCount, monConn;
  Dev<int4> inc, conI;
  Dev<double4> conD;
  Dev<uint8_t> gMode, mkWhen;
  Dev<double> gpx, gpy, gpz, mkx, mky, mkz;
  DevNet net{};
  // per instance
  Dev<uint8_t> en, foundered, evKind;
  Dev<double> area, tOn, st, out, vol, level, ccx, ccy, ccz, evT, overT, markT, founderT, final7, denseRec, sampleRec, nextSample, curveRec,
      nextCurve;
  Dev<int32_t> evId, evCount, denseCount, sampleCount, curveCount;
  Dev<long long> steps;
  Dev<Params> params;
  // host staging of the per-instance tables until run()
  std::vector<uint8_t> hEn;
  std::vector<double> hArea, hTOn, hSt, hVol, hLevel, hCx, hCy, hCz;
  // host copies read back after run()
  std::vector<long long> rSteps;
  std::vector<uint8_t> rFoundered;
  std::vector<double> rFounderT, rSt, rOut, rFinal;
  std::vector<int32_t> rEvCount, rCurveCount;

  DevShip shipView() const {
    DevShip s{};
    s.ncol = ncol; s.nseg = nseg; s.nn = NN;
    s.col = (const SSG d2v*)col.p; s.colF = (const SSG f4v*)colF.p;
    s.seg = (const SSG d2v*)seg.p; s.segF = (const SSG f4v*)segF.p;
    s.segStart = G(segStart.p); s.segCount = G(segCount.p);
    s.mu = G(mu.p); s.vmax = G(vmax.p); s.amin = G(amin.p); s.aov = G(aov.p);
    return s;
  }

  DevBatch batchView() const {
    DevBatch b{};
    b.count = opt.count;
    b.en = G(en.p); b.area = G(area.p); b.tOn = G(tOn.p);
    b.st = W(st.p); b.out = W(out.p);
    b.vol = W(vol.p); b.level = W(level.p); b.cx = W(ccx.p); b.cy = W(ccy.p); b.cz = W(ccz.p);
    b.evT = W(evT.p); b.evId = W(evId.p); b.evKind = W(evKind.p); b.evCount = W(evCount.p);
    b.overT = W(overT.p); b.markT = W(markT.p); b.foundered = W(foundered.p); b.founderT = W(founderT.p); b.steps = W(steps.p);
    b.final7 = W(final7.p);
    b.traced = opt.traced; b.denseSteps = opt.denseSteps; b.maxSamples = maxSamples; b.recSize = recSize; b.sampleEvery = opt.sampleEvery;
    b.denseRec = W(denseRec.p); b.sampleRec = W(sampleRec.p); b.denseCount = W(denseCount.p); b.sampleCount = W(sampleCount.p);
    b.nextSample = W(nextSample.p);
    b.curveEvery = opt.curveEvery; b.curveMax = curveMax; b.curveRec = W(curveRec.p); b.curveCount = W(curveCount.p);
    b.nextCurve = W(nextCurve.p);
    return b;
  }
};

namespace {

// LDS layout of one simulation, in doubles; byte and int arrays are packed into whole doubles.
LdsLayout lds_layout(int32_t NN, int32_t NC, int32_t NG, int32_t nmon, int32_t nmark, int32_t maxGroup, int32_t evCap, int32_t nMembers) {
  LdsLayout L{};
  int32_t off = 0;
  auto take = [&](int32_t doubles) { const int32_t at = off; off += std::max<int32_t>(doubles, 1); return at; };
  auto ints = [](int32_t n) { return (n + 1) / 2; };
  auto bytes = [](int32_t n) { return (n + 7) / 8; };
  L.vol = take(NN); L.level = take(NN); L.cx = take(NN); L.cy = take(NN); L.cz = take(NN); L.area = take(NN); L.zmin = take(NN);
  L.zmax = take(NN); L.hEff = take(NN); L.aEff = take(NN); L.excess = take(NN);
  L.tX = take(NN); L.tY = take(NN); L.tF = take(NN); L.tM = take(NN); L.tIp = take(NN); L.tIr = take(NN);
  L.dV = take(NC); L.q = take(NC); L.carea = take(NC); L.ctOn = take(NC);
  L.passes = take(7 * kUnits * maxGroup);
  L.warpHyd = take(5 * kUnits);
  L.frame = take(F_COUNT);
  L.scal = take(S_COUNT);
  L.over = take(nmon); L.overT = take(nmon); L.markT = take(nmark); L.evT = take(evCap);
  L.nMu = take(NN); L.nVmax = take(NN); L.nAov = take(NN); L.nAmin = take(NN); L.gpx = take(NG); L.gpy = take(NG); L.gpz = take(NG);
  L.deg = take(ints(NN)); L.evId = take(ints(evCap)); L.nSegStart = take(ints(NN)); L.nSegCount = take(ints(NN));
  L.gStart = take(ints(NG)); L.gCount = take(ints(NG)); L.gNodes = take(ints(nMembers));
  L.merged = take(bytes(NG)); L.cen = take(bytes(NC)); L.evKind = take(bytes(evCap)); L.gMode = take(bytes(NG));
  L.total = off;
  return L;
}

void need(bool ok, const std::string& what) {
  if (!ok) fail(SS_INVALID_ARGUMENT, what);
}

template <class T>
void need_ptr(const T* p, const char* what) {
  need(p != nullptr, std::string("ss_batch_create: null ") + what);
}

// Range checks of every index the kernel dereferences, so that a malformed model is an error on the host.
void validate(const ss_ship& S, const ss_sim& M) {
  need(S.ncol >= 1 && S.nseg >= 0 && S.nnode >= 1, "ss_batch_create: empty ship");
  need(M.nnode == S.nnode, "ss_batch_create: ship and simulation disagree on the node count");
  need_ptr(S.x, "ship.x"); need_ptr(S.y, "ship.y"); need_ptr(S.dxdy, "ship.dxdy"); need_ptr(S.zlo, "ship.zlo"); need_ptr(S.zhi, "ship.zhi");
  need_ptr(S.mu, "ship.mu"); need_ptr(S.vmax, "ship.vmax"); need_ptr(S.amin, "ship.amin"); need_ptr(S.aov, "ship.aov");
  need_ptr(S.segStart, "ship.segStart"); need_ptr(S.segCount, "ship.segCount");
  if (S.nseg > 0) { need_ptr(S.segCol, "ship.segCol"); need_ptr(S.segA, "ship.segA"); need_ptr(S.segE, "ship.segE"); }
  for (int32_t j = 0; j < S.nseg; ++j) need(S.segCol[j] >= 0 && S.segCol[j] < S.ncol, "ss_batch_create: segment column out of range");
  for (int32_t n = 0; n < S.nnode; ++n)
    need(S.segStart[n] >= 0 && S.segCount[n] >= 0 && static_cast<int64_t>(S.segStart[n]) + S.segCount[n] <= S.nseg,
         "ss_batch_create: node " + std::to_string(n) + " segment range exceeds the segment table");
  need(M.dt > 0, "ss_batch_create: dt must be positive");
  need(M.ngroup >= 0 && M.nconn >= 0 && M.nmonitor >= 0 && M.nmark >= 0, "ss_batch_create: negative count");
  need_ptr(M.vol0, "sim.vol0"); need_ptr(M.level0, "sim.level0"); need_ptr(M.cx0, "sim.cx0"); need_ptr(M.cy0, "sim.cy0"); need_ptr(M.cz0, "sim.cz0");
  if (M.ngroup > 0) {
    need_ptr(M.gNodeStart, "sim.gNodeStart"); need_ptr(M.gNodeCount, "sim.gNodeCount"); need_ptr(M.gNodes, "sim.gNodes");
    need_ptr(M.gMode, "sim.gMode"); need_ptr(M.gpx, "sim.gpx"); need_ptr(M.gpy, "sim.gpy"); need_ptr(M.gpz, "sim.gpz");
  }
  // every node must be solved by exactly one group, or its level is never set
  std::vector<int> covered(S.nnode, 0);
  for (int32_t g = 0; g < M.ngroup; ++g) {
    need(M.gNodeCount[g] >= 1 && M.gNodeStart[g] >= 0, "ss_batch_create: empty group");
    for (int32_t k = 0; k < M.gNodeCount[g]; ++k) {
      const int32_t n = M.gNodes[M.gNodeStart[g] + k];
      need(n >= 0 && n < S.nnode, "ss_batch_create: group member out of range");
      covered[n]++;
    }
  }
  for (int32_t n = 0; n < S.nnode; ++n) need(covered[n] == 1, "ss_batch_create: node " + std::to_string(n) + " is not in exactly one group");
  if (M.nconn > 0) {
    need_ptr(M.ca, "sim.ca"); need_ptr(M.cb, "sim.cb"); need_ptr(M.law, "sim.law"); need_ptr(M.cx, "sim.cx"); need_ptr(M.cy, "sim.cy");
    need_ptr(M.cz, "sim.cz"); need_ptr(M.area, "sim.area"); need_ptr(M.coef, "sim.coef"); need_ptr(M.en, "sim.en"); need_ptr(M.tOn, "sim.tOn");
    need_ptr(M.skipGroup, "sim.skipGroup"); need_ptr(M.monitor, "sim.monitor");
  }
  for (int32_t c = 0; c < M.nconn; ++c) {
    need(M.ca[c] >= -1 && M.ca[c] < S.nnode && M.cb[c] >= 0 && M.cb[c] < S.nnode, "ss_batch_create: connection end out of range");
    need(M.skipGroup[c] >= -1 && M.skipGroup[c] < M.ngroup, "ss_batch_create: connection skipGroup out of range");
    need(M.monitor[c] >= -1 && M.monitor[c] < M.nmonitor, "ss_batch_create: connection monitor out of range");
  }
  if (M.nmark > 0) { need_ptr(M.mkx, "sim.mkx"); need_ptr(M.mky, "sim.mky"); need_ptr(M.mkz, "sim.mkz"); need_ptr(M.mkWhen, "sim.mkWhen"); }
}

template <class T>
std::vector<T> repeat(const T* v, size_t n, size_t times) {
  std::vector<T> out;
  out.reserve(n * times);
  for (size_t i = 0; i < times; ++i) out.insert(out.end(), v, v + n);
  return out;
}

void check_instance(const ss_batch* b, int32_t i) {
  if (i < 0 || i >= b->opt.count) fail(SS_OUT_OF_RANGE, "instance " + std::to_string(i) + " out of range");
}

// Copies of the device results every accessor reads; one download per array after run().
void read_back(ss_batch* b) {
  const size_t count = static_cast<size_t>(b->opt.count);
  b->rSteps = b->steps.all();
  b->rFoundered = b->foundered.all();
  b->rFounderT = b->founderT.all();
  b->rSt = b->st.all();
  b->rOut = b->out.all();
  b->rFinal = b->final7.all();
  b->rEvCount = b->evCount.all();
  if (b->opt.curveEvery > 0) b->rCurveCount = b->curveCount.all();
  b->totalSteps = 0;
  for (size_t i = 0; i < count; ++i) b->totalSteps += b->rSteps[i];
}

}  // namespace

extern "C" {

const char* ss_last_error(void) { return g_error.c_str(); }

ss_status ss_available(char* why, size_t whyCap) {
  std::string w;
  const int n = device_count(&w);
  if (n > 0) return SS_OK;
  g_error = w;
  if (why && whyCap) {
    std::strncpy(why, w.c_str(), whyCap - 1);
    why[whyCap - 1] = 0;
  }
  return SS_RUNTIME;
}

ss_status ss_device_info(ss_device_properties* out) {
  return guarded([&] {
    need(out != nullptr, "ss_device_info: null output");
    std::string why;
    if (device_count(&why) == 0) fail(SS_RUNTIME, "ss_device_info: " + why);
    hipDeviceProp_t p{};
    hip_check(hipGetDeviceProperties(&p, 0), "hipGetDeviceProperties");
    std::memset(out, 0, sizeof *out);
    std::strncpy(out->name, p.name, sizeof out->name - 1);
    std::strncpy(out->arch, p.gcnArchName, sizeof out->arch - 1);
    out->computeUnits = p.multiProcessorCount;
    out->major = p.major;
    out->minor = p.minor;
    out->globalMemory = p.totalGlobalMem;
    out->sharedPerBlock = p.sharedMemPerBlock;
    int rv = 0, dv = 0;
    (void)hipRuntimeGetVersion(&rv);
    (void)hipDriverGetVersion(&dv);
    out->runtimeVersion = rv;
    out->driverVersion = dv;
  });
}

ss_status ss_batch_create(const ss_ship* ship, const ss_sim* sim, const ss_batch_options* opt, ss_batch** out) {
  return guarded([&] {
    need(out != nullptr, "ss_batch_create: null output handle");
    *out = nullptr;
    need(ship != nullptr, "ss_batch_create: null ship");
    need(sim != nullptr, "ss_batch_create: null simulation");
    need(opt != nullptr, "ss_batch_create: null options");
    need(opt->count >= 1, "ss_batch_create: count must be at least 1");
    need(opt->numerics == SS_NUMERICS_PORTABLE || opt->numerics == SS_NUMERICS_PORTABLE32,
         "ss_batch_create: numerics must be portable (1) or portable32 (2); the oracle and std numerics are CPU-only");
    need(opt->traced >= 0 && opt->denseSteps >= 0, "ss_batch_create: negative trace option");
    need(opt->traced == 0 || opt->sampleEvery > 0, "ss_batch_create: sampleEvery must be positive when tracing");
    need(opt->tMax > 0 && std::isfinite(opt->tMax), "ss_batch_create: tMax must be positive and finite");
    need(opt->curveEvery >= 0, "ss_batch_create: curveEvery must not be negative");
    validate(*ship, *sim);
    std::string why;
    if (device_count(&why) == 0) fail(SS_RUNTIME, "ss_batch_create: " + why);

    std::unique_ptr<ss_batch> B(new ss_batch);
    ss_batch& b = *B;
    b.opt = *opt;
    if (b.opt.traced > b.opt.count) b.opt.traced = b.opt.count;
    b.mixed = opt->numerics == SS_NUMERICS_PORTABLE32;
    const ss_ship& S = *ship;
    const ss_sim& M = *sim;
    b.NN = S.nnode; b.NC = M.nconn; b.NG = M.ngroup; b.nmon = M.nmonitor; b.nmark = M.nmark; b.ncol = S.ncol; b.nseg = S.nseg;
    b.evCap = b.nmon + b.nmark + 1;
    for (int32_t g = 0; g < b.NG; ++g) b.maxGroup = std::max(b.maxGroup, M.gNodeCount[g]);
    size_t members = 0;  // length of the group member list the groups index into
    for (int32_t g = 0; g < b.NG; ++g) members = std::max(members, static_cast<size_t>(M.gNodeStart[g] + M.gNodeCount[g]));
    b.layout = lds_layout(b.NN, b.NC, b.NG, b.nmon, b.nmark, b.maxGroup, b.evCap, static_cast<int32_t>(members));
    b.ldsBytes = static_cast<size_t>(b.layout.total) * sizeof(double);
    hipDeviceProp_t p{};
    hip_check(hipGetDeviceProperties(&p, 0), "hipGetDeviceProperties");
    if (b.ldsBytes > p.sharedMemPerBlock)
      fail(SS_RUNTIME, "ss_batch_create: this model needs " + std::to_string(b.ldsBytes) + " bytes of LDS per simulation, the device allows " +
                           std::to_string(p.sharedMemPerBlock));
    b.recSize = 10 + 5 * b.NN;
    b.maxSamples = b.opt.traced > 0 ? static_cast<int32_t>(std::floor(b.opt.tMax / b.opt.sampleEvery)) + 2 : 0;

    // ship: columns, and per segment the streams of its column
    const size_t nc = static_cast<size_t>(S.ncol), ns = static_cast<size_t>(S.nseg), nn = static_cast<size_t>(S.nnode);
    // records (x, y, bottom, top, footprint): 6 doubles (the last one 0) and 8 floats (the last three 0)
    auto records = [](size_t n, auto&& value, std::vector<double>& d, std::vector<float>& f) {
      d.assign(std::max<size_t>(n, 1) * 6, 0.0);
      f.assign(std::max<size_t>(n, 1) * 8, 0.0f);
      for (size_t i = 0; i < n; ++i)
        for (int k = 0; k < 5; ++k) {
          const double v = value(i, k);
          d[6 * i + k] = v;
          f[8 * i + k] = static_cast<float>(v);
        }
    };
    std::vector<double> recD;
    std::vector<float> recF;
    records(nc, [&](size_t i, int k) { const double* a[5] = {S.x, S.y, S.zlo, S.zhi, S.dxdy}; return a[k][i]; }, recD, recF);
    b.col.upload(recD); b.colF.upload(recF);
    records(ns, [&](size_t j, int k) {
      const size_t c = static_cast<size_t>(S.segCol[j]);
      return k == 0 ? S.x[c] : k == 1 ? S.y[c] : k == 2 ? S.segA[j] : k == 3 ? S.segE[j] : S.dxdy[c];
    }, recD, recF);
    b.seg.upload(recD); b.segF.upload(recF);
    b.segStart.upload(S.segStart, nn); b.segCount.upload(S.segCount, nn);
    b.mu.upload(S.mu, nn); b.vmax.upload(S.vmax, nn); b.amin.upload(S.amin, nn); b.aov.upload(S.aov, nn);

    // scenario: groups, connections, incidence (model.cpp:137-167), monitor lists, marks
    const size_t ncn = static_cast<size_t>(M.nconn), ng = static_cast<size_t>(M.ngroup);
    b.gStart.upload(M.gNodeStart, ng); b.gCount.upload(M.gNodeCount, ng); b.gNodes.upload(M.gNodes, members); b.gMode.upload(M.gMode, ng);
    b.gpx.upload(M.gpx, ng); b.gpy.upload(M.gpy, ng); b.gpz.upload(M.gpz, ng);
    std::vector<int4> conI(std::max<size_t>(ncn, 1));
    std::vector<double4> conD(std::max<size_t>(ncn, 1));
    for (size_t c = 0; c < ncn; ++c) {
      conI[c] = make_int4(M.ca[c], M.cb[c], M.skipGroup[c], M.law[c]);
      conD[c] = make_double4(M.cx[c], M.cy[c], M.cz[c], M.coef[c]);
    }
    b.conI.upload(conI); b.conD.upload(conD);
    std::vector<int32_t> incCount(nn, 0), incStart(nn, 0), fill(nn, 0), sea;
    for (size_t c = 0; c < ncn; ++c) {
      if (M.ca[c] >= 0) incCount[M.ca[c]]++;
      incCount[M.cb[c]]++;
    }
    int32_t total = 0;
    for (size_t n = 0; n < nn; ++n) { incStart[n] = total; total += incCount[n]; }
    std::vector<int4> inc(std::max<int32_t>(total, 1), make_int4(0, 0, 0, 0));
    for (size_t c = 0; c < ncn; ++c) {
      const int32_t ci = static_cast<int32_t>(c);
      if (M.ca[c] >= 0) {
        const int32_t n = M.ca[c];
        inc[incStart[n] + fill[n]++] = make_int4(ci, M.ca[c], M.cb[c], 0);
      } else {
        sea.push_back(ci);
      }
      const int32_t n = M.cb[c];
      inc[incStart[n] + fill[n]++] = make_int4(ci, M.ca[c], M.cb[c], 1);
    }
    b.incStart.upload(incStart); b.incCount.upload(incCount); b.inc.upload(inc); b.sea.upload(sea);
    std::vector<int32_t> monCount(std::max<int32_t>(b.nmon, 0), 0), monStart(std::max<int32_t>(b.nmon, 0), 0), monConn;
    for (int32_t m = 0; m < b.nmon; ++m) {
      monStart[m] = static_cast<int32_t>(monConn.size());
      for (size_t c = 0; c < ncn; ++c)
        if (M.monitor[c] == m) monConn.push_back(static_cast<int32_t>(c));
      monCount[m] = static_cast<int32_t>(monConn.size()) - monStart[m];
    }
    b.monStart.upload(monStart); b.monCount.upload(monCount); b.monConn.upload(monConn);
    b.mkx.upload(M.mkx, M.nmark); b.mky.upload(M.mky, M.nmark); b.mkz.upload(M.mkz, M.nmark); b.mkWhen.upload(M.mkWhen, M.nmark);

    DevNet& d = b.net;
    d.rho = M.rho; d.g = M.g; d.sq2g = std::sqrt(2 * M.g); d.dt = M.dt;
    d.ng = M.ngroup; d.nMembers = static_cast<int32_t>(members); d.gStart = G(b.gStart.p); d.gCount = G(b.gCount.p); d.gNodes = G(b.gNodes.p); d.gMode = G(b.gMode.p);
    d.gpx = G(b.gpx.p); d.gpy = G(b.gpy.p); d.gpz = G(b.gpz.p); d.gMargin = M.gMargin;
    d.nc = M.nconn; d.conI = (const SSG i4v*)b.conI.p; d.conD = (const SSG d4v*)b.conD.p;
    d.incStart = G(b.incStart.p); d.incCount = G(b.incCount.p); d.inc = (const SSG i4v*)b.inc.p;
    d.seaCount = static_cast<int32_t>(sea.size()); d.sea = G(b.sea.p);
    d.nmon = b.nmon; d.monThreshold = M.monitorThreshold; d.monStart = G(b.monStart.p); d.monCount = G(b.monCount.p); d.monConn = G(b.monConn.p);
    d.nmark = M.nmark; d.mkx = G(b.mkx.p); d.mky = G(b.mky.p); d.mkz = G(b.mkz.p); d.mkWhen = G(b.mkWhen.p);
    d.hullTopBelow = M.founder[0]; d.pitchAbsAbove = M.founder[1]; d.buoyancyBelow = M.founder[2]; d.minT = M.founder[3];
    const double* bd = M.body;
    d.ms = bd[0]; d.g0x = bd[2]; d.g0y = bd[3]; d.g0z = bd[4]; d.XO = bd[5]; d.YO = bd[6]; d.mh0 = bd[7]; d.Ith0 = bd[8]; d.Iph0 = bd[9];
    d.cDz = bd[10]; d.cDth = bd[11]; d.cDph = bd[12]; d.qz = bd[13]; d.qth = bd[14]; d.qph = bd[15];
    d.xFP = M.readout[0]; d.xAP = M.readout[1];
    d.maxGroup = b.maxGroup; d.evCap = b.evCap;

    // per-instance tables, staged on the host until run()
    const size_t count = static_cast<size_t>(b.opt.count);
    b.hEn = repeat(M.en, ncn, count); b.hArea = repeat(M.area, ncn, count); b.hTOn = repeat(M.tOn, ncn, count);
    b.hSt = repeat(M.state0, 7, count);
    b.hVol = repeat(M.vol0, nn, count); b.hLevel = repeat(M.level0, nn, count);
    b.hCx = repeat(M.cx0, nn, count); b.hCy = repeat(M.cy0, nn, count); b.hCz = repeat(M.cz0, nn, count);
    *out = B.release();
  });
}

void ss_batch_destroy(ss_batch* b) { delete b; }

ss_status ss_batch_restore(ss_batch* b, int32_t instance, const double* state7, const double* vol, const double* level, const double* cx,
                           const double* cy, const double* cz) {
  return guarded([&] {
    need(b != nullptr, "ss_batch_restore: null batch");
    if (b->ran) fail(SS_LOGIC, "ss_batch_restore: restore after run");
    check_instance(b, instance);
    need(state7 && vol && level && cx && cy && cz, "ss_batch_restore: null array");
    const size_t i = static_cast<size_t>(instance), nn = static_cast<size_t>(b->NN);
    std::copy(state7, state7 + 7, b->hSt.begin() + i * 7);
    std::copy(vol, vol + nn, b->hVol.begin() + i * nn);
    std::copy(level, level + nn, b->hLevel.begin() + i * nn);
    std::copy(cx, cx + nn, b->hCx.begin() + i * nn);
    std::copy(cy, cy + nn, b->hCy.begin() + i * nn);
    std::copy(cz, cz + nn, b->hCz.begin() + i * nn);
  });
}

static ss_status set_one(ss_batch* b, int32_t instance, int32_t conn, const char* what, const std::function<void(size_t)>& f);

ss_status ss_batch_set_connection_enabled(ss_batch* b, int32_t instance, int32_t conn, int32_t enabled) {
  return set_one(b, instance, conn, "ss_batch_set_connection_enabled", [&](size_t k) { b->hEn[k] = enabled ? 1 : 0; });
}
ss_status ss_batch_set_connection_area(ss_batch* b, int32_t instance, int32_t conn, double area) {
  return set_one(b, instance, conn, "ss_batch_set_connection_area", [&](size_t k) { b->hArea[k] = area; });
}
ss_status ss_batch_set_connection_activation(ss_batch* b, int32_t instance, int32_t conn, double tOn) {
  return set_one(b, instance, conn, "ss_batch_set_connection_activation", [&](size_t k) { b->hTOn[k] = tOn; });
}

ss_status ss_batch_set_connections(ss_batch* b, const uint8_t* en, const double* area, const double* tOn) {
  return guarded([&] {
    need(b != nullptr, "ss_batch_set_connections: null batch");
    if (b->ran) fail(SS_LOGIC, "ss_batch_set_connections: set after run");
    const size_t n = b->hEn.size();
    if (en) for (size_t k = 0; k < n; ++k) b->hEn[k] = en[k] ? 1 : 0;
    if (area) std::copy(area, area + n, b->hArea.begin());
    if (tOn) std::copy(tOn, tOn + n, b->hTOn.begin());
  });
}

ss_status ss_batch_run(ss_batch* b, double* seconds) {
  return guarded([&] {
    need(b != nullptr, "ss_batch_run: null batch");
    if (b->ran) fail(SS_LOGIC, "ss_batch_run: run twice");
    b->ran = true;
    const size_t count = static_cast<size_t>(b->opt.count);
    const size_t nn = static_cast<size_t>(b->NN);
    b->en.upload(b->hEn); b->area.upload(b->hArea); b->tOn.upload(b->hTOn);
    b->st.upload(b->hSt);
    b->out.fill(0.0, count * 7);
    b->vol.upload(b->hVol); b->level.upload(b->hLevel); b->ccx.upload(b->hCx); b->ccy.upload(b->hCy); b->ccz.upload(b->hCz);
    const size_t ev = count * static_cast<size_t>(b->evCap);
    b->evT.fill(0.0, ev); b->evId.fill(0, ev); b->evKind.fill(0, ev); b->evCount.fill(0, count);
    b->overT.fill(-1.0, count * static_cast<size_t>(b->nmon));
    b->markT.fill(-1.0, count * static_cast<size_t>(b->nmark));
    b->foundered.fill(0, count); b->founderT.fill(-1.0, count); b->steps.fill(0, count); b->final7.fill(0.0, count * 7);
    const size_t traced = static_cast<size_t>(b->opt.traced), rs = static_cast<size_t>(b->recSize);
    b->denseRec.alloc(traced * static_cast<size_t>(b->opt.denseSteps + 1) * rs);
    b->sampleRec.alloc(traced * static_cast<size_t>(b->maxSamples) * rs);
    b->denseCount.fill(0, traced); b->sampleCount.fill(0, traced); b->nextSample.fill(b->opt.sampleEvery, traced);
    if (b->opt.curveEvery > 0) {
      b->curveMax = static_cast<int32_t>(std::floor(b->opt.tMax / b->opt.curveEvery)) + 2;
      b->curveRec.alloc(count * static_cast<size_t>(b->curveMax) * kCurveValues);
      b->curveCount.fill(0, count);
      b->nextCurve.fill(0.0, count);  // the first record is the initial state, as in the oracle's run
    }
    (void)nn;
    Params prm{};
    prm.S = b->shipView();
    prm.M = b->net;
    prm.Bt = b->batchView();
    prm.L = b->layout;
    b->params.upload(&prm, 1);
    const long long maxSteps = b->opt.stepsPerLaunch > 0 ? static_cast<long long>(b->opt.stepsPerLaunch) : LLONG_MAX;
    auto kernel = b->mixed ? run_chunk<true> : run_chunk<false>;
    if (b->ldsBytes > 64 * 1024)
      hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(b->ldsBytes)),
                "hipFuncSetAttribute");
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const auto t0 = std::chrono::steady_clock::now();
    int first = 1;
    b->kernelMs = 0;
    b->launches = 0;
    std::vector<uint8_t> f(count);
    std::vector<double> st(count * 7);
    for (;;) {
      hip_check(hipEventRecord(e0, nullptr), "hipEventRecord");
      hipLaunchKernelGGL(kernel, dim3(b->opt.count), dim3(kThreads), b->ldsBytes, nullptr, static_cast<const Params*>(b->params.p), maxSteps,
                         b->opt.tMax, first);
      hip_check(hipGetLastError(), "kernel launch");
      hip_check(hipEventRecord(e1, nullptr), "hipEventRecord");
      hip_check(hipEventSynchronize(e1), "kernel execution");
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
      b->kernelMs += ms;
      b->launches++;
      first = 0;
      if (maxSteps == LLONG_MAX) break;
      b->foundered.download(f.data(), count);
      b->st.download(st.data(), count * 7);
      bool active = false;
      for (size_t i = 0; i < count; ++i)
        if (!f[i] && st[i * 7] < b->opt.tMax) { active = true; break; }
      if (!active) break;
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    read_back(b);
    if (seconds) *seconds = secs;
  });
}

ss_status ss_batch_total_steps(const ss_batch* b, int64_t* steps) {
  return guarded([&] {
    need(b != nullptr && steps != nullptr, "ss_batch_total_steps: null argument");
    *steps = b->totalSteps;
  });
}

ss_status ss_batch_shared_memory_bytes(const ss_batch* b, size_t* bytes) {
  return guarded([&] {
    need(b != nullptr && bytes != nullptr, "ss_batch_shared_memory_bytes: null argument");
    *bytes = b->ldsBytes;
  });
}

ss_status ss_batch_result(const ss_batch* b, int32_t instance, ss_instance_result* res, double* vol, double* level, double* cx, double* cy,
                          double* cz, ss_event* events, int32_t evCap) {
  return guarded([&] {
    need(b != nullptr, "ss_batch_result: null batch");
    if (!b->ran) fail(SS_LOGIC, "ss_batch_result: result before run");
    check_instance(b, instance);
    const size_t i = static_cast<size_t>(instance), nn = static_cast<size_t>(b->NN);
    if (res) {
      res->foundered = b->rFoundered[i];
      res->founderT = b->rFounderT[i];
      res->tEnd = b->rSt[i * 7];
      res->steps = b->rSteps[i];
      for (int k = 0; k < 7; ++k) { res->state[k] = b->rSt[i * 7 + k]; res->outputs[k] = b->rOut[i * 7 + k]; }
      res->events = b->rEvCount[i];
    }
    if (vol) b->vol.download(vol, nn, i * nn);
    if (level) b->level.download(level, nn, i * nn);
    if (cx) b->ccx.download(cx, nn, i * nn);
    if (cy) b->ccy.download(cy, nn, i * nn);
    if (cz) b->ccz.download(cz, nn, i * nn);
    if (events && evCap > 0) {
      const int32_t n = std::min(b->rEvCount[i], evCap);
      std::vector<double> t(n);
      std::vector<int32_t> id(n);
      std::vector<uint8_t> kind(n);
      const size_t off = i * static_cast<size_t>(b->evCap);
      b->evT.download(t.data(), n, off);
      b->evId.download(id.data(), n, off);
      b->evKind.download(kind.data(), n, off);
      for (int32_t k = 0; k < n; ++k) events[k] = ss_event{t[k], id[k], kind[k]};
    }
  });
}

ss_status ss_batch_results(const ss_batch* b, int64_t* steps, int32_t* foundered, double* founderT, double* tEnd) {
  return guarded([&] {
    need(b != nullptr, "ss_batch_results: null batch");
    if (!b->ran) fail(SS_LOGIC, "ss_batch_results: results before run");
    for (int32_t i = 0; i < b->opt.count; ++i) {
      if (steps) steps[i] = b->rSteps[i];
      if (foundered) foundered[i] = b->rFoundered[i];
      if (founderT) founderT[i] = b->rFounderT[i];
      if (tEnd) tEnd[i] = b->rSt[static_cast<size_t>(i) * 7];
    }
  });
}

ss_status ss_batch_trace(const ss_batch* b, int32_t instance, double* dense, int32_t denseCap, int32_t* denseCount, double* samples,
                         int32_t samplesCap, int32_t* samplesCount) {
  return guarded([&] {
    need(b != nullptr, "ss_batch_trace: null batch");
    if (!b->ran) fail(SS_LOGIC, "ss_batch_trace: trace before run");
    if (instance < 0 || instance >= b->opt.traced) fail(SS_OUT_OF_RANGE, "ss_batch_trace: instance was not traced");
    const size_t i = static_cast<size_t>(instance), rs = static_cast<size_t>(b->recSize);
    int32_t nd = 0, ns = 0;
    b->denseCount.download(&nd, 1, i);
    b->sampleCount.download(&ns, 1, i);
    if (denseCount) *denseCount = nd;
    if (samplesCount) *samplesCount = ns;
    if (dense) b->denseRec.download(dense, static_cast<size_t>(std::min(nd, denseCap)) * rs, i * static_cast<size_t>(b->opt.denseSteps + 1) * rs);
    if (samples) b->sampleRec.download(samples, static_cast<size_t>(std::min(ns, samplesCap)) * rs, i * static_cast<size_t>(b->maxSamples) * rs);
  });
}

ss_status ss_batch_curve(const ss_batch* b, int32_t instance, double* rec, int32_t cap, int32_t* count) {
  return guarded([&] {
    need(b != nullptr, "ss_batch_curve: null batch");
    if (!b->ran) fail(SS_LOGIC, "ss_batch_curve: curve before run");
    if (b->opt.curveEvery <= 0) fail(SS_LOGIC, "ss_batch_curve: curves were not recorded (curveEvery is 0)");
    check_instance(b, instance);
    const size_t i = static_cast<size_t>(instance);
    const int32_t n = b->rCurveCount[i];
    if (count) *count = n + 1;
    if (rec && cap > 0) {
      const int32_t keep = std::min(n, cap);
      b->curveRec.download(rec, static_cast<size_t>(keep) * kCurveValues, i * static_cast<size_t>(b->curveMax) * kCurveValues);
      if (cap > n) std::copy(b->rFinal.begin() + i * 7, b->rFinal.begin() + i * 7 + 7, rec + static_cast<size_t>(n) * kCurveValues);
    }
  });
}

ss_status ss_batch_final_readouts(const ss_batch* b, int32_t instance, double* rec7) {
  return guarded([&] {
    need(b != nullptr && rec7 != nullptr, "ss_batch_final_readouts: null argument");
    if (!b->ran) fail(SS_LOGIC, "ss_batch_final_readouts: readouts before run");
    check_instance(b, instance);
    std::copy(b->rFinal.begin() + static_cast<size_t>(instance) * 7, b->rFinal.begin() + static_cast<size_t>(instance) * 7 + 7, rec7);
  });
}

// Device time of run(): the sum over launches of the kernel's HIP-event duration, and the launch count.
ss_status ss_batch_kernel_time(const ss_batch* b, double* ms, int32_t* launches) {
  return guarded([&] {
    need(b != nullptr, "ss_batch_kernel_time: null batch");
    if (ms) *ms = b->kernelMs;
    if (launches) *launches = b->launches;
  });
}

}  // extern "C"

static ss_status set_one(ss_batch* b, int32_t instance, int32_t conn, const char* what, const std::function<void(size_t)>& f) {
  return guarded([&] {
    need(b != nullptr, std::string(what) + ": null batch");
    if (b->ran) fail(SS_LOGIC, std::string(what) + ": set after run");
    check_instance(b, instance);
    if (conn < 0 || conn >= b->NC) fail(SS_OUT_OF_RANGE, std::string(what) + ": connection " + std::to_string(conn) + " out of range");
    f(static_cast<size_t>(instance) * b->NC + static_cast<size_t>(conn));
  });
}
