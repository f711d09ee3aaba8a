// Host-side internals of libsinksim_hip.so shared by its translation units (batch.hip: the batch and its C ABI;
// adjoint.hip: the gradient): error reporting, device allocations, the uploaded models and the batch object.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sinksim_hip.h"
#include "step_dev.hpp"

namespace ssa {

constexpr int kCurveValues = 7;

// ------------------------------------------------------------------------------------------- host side

// the last error of a C-ABI call on this thread (ss_last_error)
inline thread_local std::string g_error;

// device addresses as the kernel's address-space-qualified pointer types
template <class T>
inline const SSG T* G(const T* p) { return (const SSG T*)p; }
template <class T>
inline SSG T* W(T* p) { return (SSG T*)p; }

struct Fail {
  ss_status status;
  std::string what;
};

[[noreturn]] inline void fail(ss_status s, const std::string& what) { throw Fail{s, what}; }

inline void hip_check(hipError_t e, const char* what) {
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

inline int device_count(std::string* why) {
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    if (why) *why = hipGetErrorString(e);
    return 0;
  }
  if (n == 0 && why) *why = "no HIP device";
  return n;
}

}  // namespace ssa

using namespace ssa;


// LDS layout of one simulation, in doubles; byte and int arrays are packed into whole doubles.
inline LdsLayout lds_layout(int32_t NN, int32_t NC, int32_t NG, int32_t nmon, int32_t nmark, int32_t maxGroup, int32_t evCap, int32_t nMembers,
                     int32_t units, int32_t nsea) {
  LdsLayout L{};
  int32_t off = 0;
  auto take = [&](int32_t doubles) { const int32_t at = off; off += std::max<int32_t>(doubles, 1); return at; };
  auto ints = [](int32_t n) { return (n + 1) / 2; };
  auto bytes = [](int32_t n) { return (n + 7) / 8; };
  L.vol = take(NN); L.level = take(NN); L.cx = take(NN); L.cy = take(NN); L.cz = take(NN); L.cvol = take(NN); L.area = take(NN); L.zmin = take(NN);
  L.zmax = take(NN); L.hEff = take(NN); L.aEff = take(NN); L.excess = take(NN);
  L.tX = take(NN); L.tY = take(NN); L.tF = take(NN); L.tM = take(NN); L.tIp = take(NN); L.tIr = take(NN);
  L.dV = take(NC); L.q = take(NC); L.carea = take(NC); L.ctOn = take(NC); L.seaDV = take(nsea);
  L.passes = take(7 * units * maxGroup);  // each solver unit's pass scratch
  L.warpHyd = take(5 * kUnits);
  L.frame = take(F_COUNT);
  L.scal = take(S_COUNT);
  L.over = take(nmon); L.overT = take(nmon); L.markT = take(nmark); L.evT = take(evCap); L.nextGroup = take(1);
  L.nMu = take(NN); L.nVmax = take(NN); L.nAov = take(NN); L.nAmin = take(NN); L.gpx = take(NG); L.gpy = take(NG); L.gpz = take(NG);
  L.deg = take(ints(NN)); L.evId = take(ints(evCap)); L.nSegStart = take(ints(NN)); L.nSegCount = take(ints(NN));
  L.gStart = take(ints(NG)); L.gCount = take(ints(NG)); L.gNodes = take(ints(nMembers));
  L.merged = take(bytes(NG)); L.cen = take(bytes(NC)); L.evKind = take(bytes(evCap)); L.gMode = take(bytes(NG));
  L.total = off;
  return L;
}

// One model (a hull and a flooding scenario) on the device: the geometry records, the network tables and their
// views, with the scenario's per-instance defaults kept on the host.
struct Model {
  int32_t NN = 0, NC = 0, NG = 0, nmon = 0, nmark = 0, evCap = 0, maxGroup = 1, ncol = 0, nseg = 0, members = 0, nsea = 0;
  double cost = 0;  // work per step, relative: the order in which a batch hands its instances out
  // ship
  Dev<double> col, seg, mu, vmax, amin, aov;  // col, seg: 6 doubles per record (48 bytes)
  mutable Dev<double> col40, seg40;           // the same records packed in 5 doubles (40 bytes), made on first use
  void pack40() const;                        // makes col40 / seg40 from col / seg on the device
  Dev<float> colF, segF;                      // 8 floats per record
  Dev<int32_t> segStart, segCount;
  // scenario
  Dev<int32_t> gStart, gCount, gNodes, incStart, incCount, sea, monStart, monCount, monConn, flowOrder;
  Dev<int4> inc, conI;
  Dev<double4> conD;
  Dev<uint8_t> gMode, mkWhen;
  Dev<double> gpx, gpy, gpz, mkx, mky, mkz;
  DevNet net{};
  // the scenario's defaults of the per-instance tables and state
  std::vector<uint8_t> en0;
  std::vector<double> area0, tOn0, state0, vol0, level0, cx0, cy0, cz0;
  std::vector<double> coef0;  // the discharge coefficients, which the per-instance products coef * area take

  // the LDS layout of one simulation on a workgroup of tpb threads (tpb / 32 solver units), and its size
  LdsLayout layout(int tpb) const { return lds_layout(NN, NC, NG, nmon, nmark, maxGroup, evCap, members, tpb / kLanes, nsea); }
  size_t lds_bytes(int tpb) const { return static_cast<size_t>(layout(tpb).total) * sizeof(double); }

  // the model's views at a workgroup size and a double-record layout (48 or 40 bytes; 40 needs pack40 first)
  ModelParams params(int tpb, int recBytes) const {
    ModelParams m{};
    DevShip& s = m.S;
    s.ncol = ncol; s.nseg = nseg; s.nn = NN;
    s.recD = recBytes;
    s.col = (const SSG d2v*)(recBytes == 40 ? col40.p : col.p); s.colF = (const SSG f4v*)colF.p;
    s.seg = (const SSG d2v*)(recBytes == 40 ? seg40.p : seg.p); s.segF = (const SSG f4v*)segF.p;
    s.segStart = G(segStart.p); s.segCount = G(segCount.p);
    s.mu = G(mu.p); s.vmax = G(vmax.p); s.amin = G(amin.p); s.aov = G(aov.p);
    m.M = net;
    m.L = layout(tpb);
    return m;
  }
};

struct ss_batch {
  ss_batch_options opt{};
  bool mixed = true;
  std::vector<std::shared_ptr<const Model>> models;
  // per instance: its model and the offsets of its rows; the work order
  std::vector<int32_t> instModel, nodeOff, connOff, evOff, monOff, markOff, order;
  size_t nodeTotal = 0, connTotal = 0, evTotal = 0, monTotal = 0, markTotal = 0;
  size_t ldsBytes = 0;  // the largest model's, at the launch's workgroup size
  int32_t recSize = 0, maxSamples = 0, curveMax = 0;
  int32_t tpb = kThreads;  // threads per simulation (workgroup size): 256 or 512
  int32_t recBytes = 48;   // the double geometry records' layout: 48 bytes (padded) or 40 (packed)
  int32_t blocks = 0;   // workgroups of the persistent launch (0: as many as fit on the device at once)
  bool tailForm = true;  // the automatic shape: a run's last epochs at 512 threads once the live instances fit
  bool refill = false;  // whether the last run used the refilling (persistent) kernel
  int64_t totalSteps = 0;
  bool ran = false;
  double kernelMs = 0;
  int launches = 0;
  bool stopWhenStable = false;  // the oracle's calm rule ends instances that settle (simulation.cpp:179-182)
  Dev<double> calm;             // per instance, with stopWhenStable
  // exact stepping (ss_batch_step): the launch runs every instance for exactly its step count
  bool exact = false;
  bool stepped = false;          // the device holds the instances (ss_batch_step has uploaded them)
  Dev<long long> stepLimit;      // per instance, when the counts differ
  // the last gradient call's launch shape: threads per simulation, the tape's large arrays in global memory, LDS bytes
  int32_t adjTpb = 0, adjGlobal = 0;
  size_t adjLds = 0;

  Dev<ModelParams> dModels;
  Dev<ModelParams> dModelsTail;  // the models at 512 threads, for a run's tail epochs
  Dev<ModelParams> dModelsLast;  // the models at 1,024 threads, for a run's last epochs
  Dev<int32_t> epochPlan;        // the current epoch's live instances (run_epochs)
  Dev<DevBatch> dBatch;
  Dev<int32_t> dInstModel, dOrder, dNodeOff, dConnOff, dEvOff, dMonOff, dMarkOff, counter;
  // per instance
  Dev<uint8_t> en, foundered, evKind;
  Dev<double> area, tOn, st, out, vol, level, ccx, ccy, ccz, evT, overT, markT, founderT, final7, denseRec, sampleRec, nextSample, curveRec,
      nextCurve;
  Dev<int32_t> evId, evCount, denseCount, sampleCount, curveCount;
  Dev<long long> steps;
  // host staging of the per-instance tables until run()
  std::vector<uint8_t> hEn;
  std::vector<double> hArea, hTOn, hSt, hVol, hLevel, hCx, hCy, hCz;
  // host copies read back after run()
  std::vector<long long> rSteps;
  std::vector<uint8_t> rFoundered;
  std::vector<double> rFounderT, rSt, rOut, rFinal;
  std::vector<int32_t> rEvCount, rCurveCount;

  const Model& model(int32_t i) const { return *models[static_cast<size_t>(instModel[static_cast<size_t>(i)])]; }

  DevBatch batchView() const {
    DevBatch b{};
    b.count = opt.count;
    b.instModel = G(dInstModel.p); b.order = G(dOrder.p);
    b.nodeOff = G(dNodeOff.p); b.connOff = G(dConnOff.p); b.evOff = G(dEvOff.p); b.monOff = G(dMonOff.p); b.markOff = G(dMarkOff.p);
    b.counter = W(counter.p);
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
    b.stopWhenStable = stopWhenStable ? 1 : 0; b.calm = W(calm.p);
    b.exact = exact ? 1 : 0; b.stepLimit = exact && stepLimit.p ? G(stepLimit.p) : nullptr;
    return b;
  }
};

namespace ssa {

inline void need(bool ok, const std::string& what) {
  if (!ok) fail(SS_INVALID_ARGUMENT, what);
}

inline void check_instance(const ss_batch* b, int32_t i) {
  if (i < 0 || i >= b->opt.count) fail(SS_OUT_OF_RANGE, "instance " + std::to_string(i) + " out of range");
}

template <class T>
inline void need_ptr(const T* p, const std::string& who, const char* what) {
  need(p != nullptr, who + ": null " + what);
}

// Range checks of every index the kernel dereferences, so that a malformed model is an error on the host.
inline void validate(const ss_ship& S, const ss_sim& M, const std::string& who) {
  need(S.ncol >= 1 && S.nseg >= 0 && S.nnode >= 1, who + ": empty ship");
  need(M.nnode == S.nnode, who + ": ship and simulation disagree on the node count");
  need_ptr(S.x, who, "ship.x"); need_ptr(S.y, who, "ship.y"); need_ptr(S.dxdy, who, "ship.dxdy"); need_ptr(S.zlo, who, "ship.zlo"); need_ptr(S.zhi, who, "ship.zhi");
  need_ptr(S.mu, who, "ship.mu"); need_ptr(S.vmax, who, "ship.vmax"); need_ptr(S.amin, who, "ship.amin"); need_ptr(S.aov, who, "ship.aov");
  need_ptr(S.segStart, who, "ship.segStart"); need_ptr(S.segCount, who, "ship.segCount");
  if (S.nseg > 0) { need_ptr(S.segCol, who, "ship.segCol"); need_ptr(S.segA, who, "ship.segA"); need_ptr(S.segE, who, "ship.segE"); }
  for (int32_t j = 0; j < S.nseg; ++j) need(S.segCol[j] >= 0 && S.segCol[j] < S.ncol, who + ": segment column out of range");
  for (int32_t n = 0; n < S.nnode; ++n)
    need(S.segStart[n] >= 0 && S.segCount[n] >= 0 && static_cast<int64_t>(S.segStart[n]) + S.segCount[n] <= S.nseg,
         who + ": node " + std::to_string(n) + " segment range exceeds the segment table");
  need(M.dt > 0, who + ": dt must be positive");
  need(M.ngroup >= 0 && M.nconn >= 0 && M.nmonitor >= 0 && M.nmark >= 0, who + ": negative count");
  need_ptr(M.vol0, who, "sim.vol0"); need_ptr(M.level0, who, "sim.level0"); need_ptr(M.cx0, who, "sim.cx0"); need_ptr(M.cy0, who, "sim.cy0"); need_ptr(M.cz0, who, "sim.cz0");
  if (M.ngroup > 0) {
    need_ptr(M.gNodeStart, who, "sim.gNodeStart"); need_ptr(M.gNodeCount, who, "sim.gNodeCount"); need_ptr(M.gNodes, who, "sim.gNodes");
    need_ptr(M.gMode, who, "sim.gMode"); need_ptr(M.gpx, who, "sim.gpx"); need_ptr(M.gpy, who, "sim.gpy"); need_ptr(M.gpz, who, "sim.gpz");
  }
  // every node must be solved by exactly one group, or its level is never set
  std::vector<int> covered(S.nnode, 0);
  for (int32_t g = 0; g < M.ngroup; ++g) {
    need(M.gNodeCount[g] >= 1 && M.gNodeStart[g] >= 0, who + ": empty group");
    for (int32_t k = 0; k < M.gNodeCount[g]; ++k) {
      const int32_t n = M.gNodes[M.gNodeStart[g] + k];
      need(n >= 0 && n < S.nnode, who + ": group member out of range");
      covered[n]++;
    }
  }
  for (int32_t n = 0; n < S.nnode; ++n) need(covered[n] == 1, who + ": node " + std::to_string(n) + " is not in exactly one group");
  if (M.nconn > 0) {
    need_ptr(M.ca, who, "sim.ca"); need_ptr(M.cb, who, "sim.cb"); need_ptr(M.law, who, "sim.law"); need_ptr(M.cx, who, "sim.cx"); need_ptr(M.cy, who, "sim.cy");
    need_ptr(M.cz, who, "sim.cz"); need_ptr(M.area, who, "sim.area"); need_ptr(M.coef, who, "sim.coef"); need_ptr(M.en, who, "sim.en"); need_ptr(M.tOn, who, "sim.tOn");
    need_ptr(M.skipGroup, who, "sim.skipGroup"); need_ptr(M.monitor, who, "sim.monitor");
  }
  for (int32_t c = 0; c < M.nconn; ++c) {
    need(M.ca[c] >= -1 && M.ca[c] < S.nnode && M.cb[c] >= 0 && M.cb[c] < S.nnode, who + ": connection end out of range");
    need(M.skipGroup[c] >= -1 && M.skipGroup[c] < M.ngroup, who + ": connection skipGroup out of range");
    need(M.monitor[c] >= -1 && M.monitor[c] < M.nmonitor, who + ": connection monitor out of range");
  }
  if (M.nmark > 0) { need_ptr(M.mkx, who, "sim.mkx"); need_ptr(M.mky, who, "sim.mky"); need_ptr(M.mkz, who, "sim.mkz"); need_ptr(M.mkWhen, who, "sim.mkWhen"); }
}

// The forward kernel's LDS bytes per simulation: the largest model's layout at the batch's workgroup size.
inline size_t forward_lds_bytes(const ss_batch& b) {
  size_t n = 0;
  for (const auto& m : b.models) n = std::max(n, m->lds_bytes(b.tpb));
  return n;
}

// The largest dynamic LDS allocation a workgroup may request on the current device.
size_t lds_limit();

// Uploads a batch's staged instances, its models (at the batch's workgroup size) and its view for a launch.
void upload(ss_batch* b);

// The per-instance connection products coef * area of the staged tables (the kernel's flow laws take them).
inline std::vector<double> coef_area(const ss_batch* b) {
  std::vector<double> ca(b->hArea.size());
  for (size_t i = 0; i < static_cast<size_t>(b->opt.count); ++i) {
    const Model& m = b->model(static_cast<int32_t>(i));
    const size_t o = static_cast<size_t>(b->connOff[i]);
    for (int32_t c = 0; c < m.NC; ++c) ca[o + c] = m.coef0[static_cast<size_t>(c)] * b->hArea[o + c];
  }
  return ca;
}

}  // namespace ssa
