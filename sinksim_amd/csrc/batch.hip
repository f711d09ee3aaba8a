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
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sinksim_hip.h"
#include "batch_host.hpp"
#include "step_dev.hpp"

namespace ssa {

// ------------------------------------------------------------------------------------------- the kernel

// Writes the complete state record (the trace layout, 10 + 5 * nodes values), one thread; the out-of-line part takes
// the LDS arrays it reads rather than the block's view, which a call would have to copy to scratch.
__device__ __noinline__ void write_record_arrays(SSG double* rec, const SSL double* scal, const SSL double* vol, const SSL double* level,
                                                const SSL double* cx, const SSL double* cy, const SSL double* cz, int nn) {
  for (int k = 0; k < 7; ++k) rec[k] = scal[S_T + k];
  rec[7] = scal[O_INFLOW]; rec[8] = scal[O_VB]; rec[9] = scal[O_TOP];
  SSG double* p = rec + 10;
  for (int n = 0; n < nn; ++n) p[n] = vol[n];
  p += nn;
  for (int n = 0; n < nn; ++n) p[n] = level[n];
  p += nn;
  for (int n = 0; n < nn; ++n) p[n] = cx[n];
  p += nn;
  for (int n = 0; n < nn; ++n) p[n] = cy[n];
  p += nn;
  for (int n = 0; n < nn; ++n) p[n] = cz[n];
}
__device__ __forceinline__ void write_record(SSG double* rec, const Blk& B, int nn) {
  write_record_arrays(rec, B.scal, B.vol, B.level, B.cx, B.cy, B.cz, nn);
}

// The thread's index for one step. In the portable32 kernel it is hidden from the compiler at each step (an empty
// asm), so that the lane conditions the step derives from it are evaluated inside the step, one compare each,
// instead of being hoisted out of the step loop into scalar registers that spill to vector lanes and cost two lane
// reads per use; the FP64 kernel keeps them hoisted (measured faster there).
template <bool Mixed>
__device__ __forceinline__ int step_thread(int tid) {
  if constexpr (Mixed) asm volatile("" : "+v"(tid));
  return tid;
}

// One instance on one workgroup: load it into LDS, step it until it founders, reaches tMax or has taken maxSteps
// steps in this launch, store it back. Traces and curves are recorded as in the reference
// (device_step.cuh:379-416); the readouts of the state at the end of the launch go to final7.
// Inlined into the kernel: as a called function it saved and restored its callee-saved registers in scratch memory
// for every instance.
template <bool Mixed, int TPB>
__device__ __attribute__((always_inline)) void run_instance(PC& P, BC& Bt, int inst, long long maxSteps, double tMax, SSL double* lds, int tid) {
  const int NN = P.S.nn, NC = P.M.nc, evCap = P.M.evCap, nmon = P.M.nmon, nmark = P.M.nmark;
  const size_t iN = static_cast<size_t>(Bt.nodeOff[inst]), iC = static_cast<size_t>(Bt.connOff[inst]);
  const size_t iE = static_cast<size_t>(Bt.evOff[inst]), iMon = static_cast<size_t>(Bt.monOff[inst]), iMark = static_cast<size_t>(Bt.markOff[inst]);
  const Blk B = carve(lds, P);

  // load the instance, and the model's node and group tables
  if (tid < 7) {
    B.scal[S_T + tid] = Bt.st[inst * 7 + tid];
    B.scal[O_INFLOW + tid] = Bt.out[inst * 7 + tid];
  }
  if (tid == 0) {
    *B.nextGroup = 0;
    B.scal[S_EVCOUNT] = Bt.evCount[inst];
    B.scal[S_FOUNDERED] = Bt.foundered[inst];
    B.scal[S_FOUNDERT] = Bt.founderT[inst];
    B.scal[S_SEA] = 0;
    B.scal[S_CALM] = Bt.stopWhenStable ? Bt.calm[inst] : 0.0;
    if (Bt.curveEvery > 0) {
      B.scal[S_NEXTCURVE] = Bt.nextCurve[inst];
      B.scal[S_CURVECOUNT] = Bt.curveCount[inst];
    }
  }
  for (int n = tid; n < NN; n += TPB) {
    B.vol[n] = Bt.vol[iN + n]; B.level[n] = Bt.level[iN + n];
    B.cx[n] = Bt.cx[iN + n]; B.cy[n] = Bt.cy[iN + n]; B.cz[n] = Bt.cz[iN + n]; B.cvol[n] = 0;
    B.area[n] = 0; B.zmin[n] = 0; B.zmax[n] = 0; B.hEff[n] = 0; B.aEff[n] = 0; B.excess[n] = 0; B.deg[n] = 0;
    B.tX[n] = 0; B.tY[n] = 0; B.tF[n] = 0; B.tM[n] = 0; B.tIp[n] = 0; B.tIr[n] = 0;
  }
  for (int c = tid; c < NC; c += TPB) {
    B.cen[c] = Bt.en[iC + c]; B.carea[c] = Bt.area[iC + c]; B.ctOn[c] = Bt.tOn[iC + c];
    B.dV[c] = 0; B.q[c] = 0;
  }
  for (int g = tid; g < P.M.ng; g += TPB) {
    B.merged[g] = 0;
    B.gStart[g] = P.M.gStart[g]; B.gCount[g] = P.M.gCount[g]; B.gMode[g] = P.M.gMode[g];
    B.gpx[g] = P.M.gpx[g]; B.gpy[g] = P.M.gpy[g]; B.gpz[g] = P.M.gpz[g];
  }
  for (int k = tid; k < P.M.nMembers; k += TPB) B.gNodes[k] = P.M.gNodes[k];
  for (int n = tid; n < NN; n += TPB) {
    B.nSegStart[n] = P.S.segStart[n]; B.nSegCount[n] = P.S.segCount[n];
    B.nMu[n] = P.S.mu[n]; B.nVmax[n] = P.S.vmax[n]; B.nAov[n] = P.S.aov[n]; B.nAmin[n] = P.S.amin[n];
  }
  for (int e = tid; e < evCap; e += TPB) { B.evT[e] = Bt.evT[iE + e]; B.evId[e] = Bt.evId[iE + e]; B.evKind[e] = Bt.evKind[iE + e]; }
  for (int m = tid; m < nmon; m += TPB) { B.overT[m] = Bt.overT[iMon + m]; B.over[m] = 0; }
  for (int m = tid; m < nmark; m += TPB) B.markT[m] = Bt.markT[iMark + m];
  __syncthreads();
  if (tid < 64) pose_wave0(P, B, tid, B.scal[S_ZO], B.scal[S_PITCH], B.scal[S_ROLL]);  // the loaded state's pose
  __syncthreads();

  const bool traced = inst < Bt.traced;
  long long done = Bt.steps[inst];
  if (done == 0 && traced) {
    // the initial record: buoyancy at the initial pose, no inflow yet (simulation.cpp:137-145, 204-206); the
    // canonical partials are the first 256 threads'
    if (tid < kThreads) buoyancy_partials<Mixed>(P, B, tid);
    __syncthreads();
    if (tid == 0) {
      double V, bx, by, bz, top;
      finish_buoyancy(B, V, bx, by, bz, top);
      B.scal[O_INFLOW] = 0; B.scal[O_VB] = V; B.scal[O_TOP] = top; B.scal[O_BX] = bx; B.scal[O_BY] = by; B.scal[O_BZ] = bz;
      const size_t rs = static_cast<size_t>(Bt.recSize);
      write_record(Bt.denseRec + static_cast<size_t>(inst) * (Bt.denseSteps + 1) * rs, B, NN);
      write_record(Bt.sampleRec + static_cast<size_t>(inst) * Bt.maxSamples * rs, B, NN);
      Bt.denseCount[inst] = 1; Bt.sampleCount[inst] = 1; Bt.nextSample[inst] = Bt.sampleEvery;
    }
    __syncthreads();
  }

  const bool curves = Bt.curveEvery > 0;
  const bool exact = Bt.exact != 0;
  const long long limit = exact && Bt.stepLimit ? Bt.stepLimit[inst] : maxSteps;
  uint64_t pa = uniform64(reinterpret_cast<uint64_t>(&P));  // the same in every lane, held in scalar registers
  for (long long s = 0; s < limit; ++s) {
    const int t = step_thread<Mixed>(tid);
    PC& Q = model_at(pa);
    const Blk Bq = carve(lds, Q);
    // uniform: LDS values after a barrier
    if (!exact && (Bq.scal[S_FOUNDERED] != 0 || Bq.scal[S_T] >= tMax || Bq.scal[S_CALM] > kCalmStop)) break;
    // a readout-curve record when due (simulation.cpp:175-176). Every wave evaluates the condition on the same
    // LDS values; the barrier inside orders all of those reads (and the record's reads of the volumes) before
    // thread 0 advances the bookkeeping, so no wave can see the advanced values and skip the barrier.
    if (curves && Bq.scal[S_T] >= Bq.scal[S_NEXTCURVE] && Bq.scal[S_CURVECOUNT] < Bt.curveMax) {
      if (t == 0) {
        SSG double* rec = Bt.curveRec + (static_cast<size_t>(inst) * Bt.curveMax + static_cast<size_t>(Bq.scal[S_CURVECOUNT])) * kCurveStride;
        readouts7(Q, Bq.scal, Bq.vol, rec, 1);
      }
      __syncthreads();
      if (t == 0) {
        Bq.scal[S_CURVECOUNT] += 1;
        Bq.scal[S_NEXTCURVE] += Bt.curveEvery;
      }
      // read again only at the next iteration, after the step's barriers
    }
    device_step<Mixed, TPB>(pa, lds, t);
    ++done;
    if (Bt.stopWhenStable) {
      // simulation.cpp:179-182: past 30 minutes, calm accumulates while the inflow, the pitch rate and the heave
      // velocity after the step stay small, and resets otherwise; the loop ends once it exceeds 20 minutes
      if (t == 0 && Bq.scal[S_T] > kCalmStart) {
        const bool still = __builtin_fabs(Bq.scal[O_INFLOW]) < 0.02 && __builtin_fabs(Bq.scal[S_WTH]) < 1e-5 && __builtin_fabs(Bq.scal[S_VZ]) < 1e-3;
        Bq.scal[S_CALM] = still ? Bq.scal[S_CALM] + Q.M.dt : 0.0;
      }
      __syncthreads();
    }
    if (traced) {
      if (t == 0) {
        const size_t rs = static_cast<size_t>(Bt.recSize);
        if (done <= Bt.denseSteps) {
          const int k = Bt.denseCount[inst]++;
          write_record(Bt.denseRec + (static_cast<size_t>(inst) * (Bt.denseSteps + 1) + k) * rs, Bq, NN);
        }
        if (Bq.scal[S_T] >= Bt.nextSample[inst] - 1e-9 && Bt.sampleCount[inst] < Bt.maxSamples) {
          const int k = Bt.sampleCount[inst]++;
          write_record(Bt.sampleRec + (static_cast<size_t>(inst) * Bt.maxSamples + k) * rs, Bq, NN);
          Bt.nextSample[inst] += Bt.sampleEvery;
        }
      }
      __syncthreads();
    }
  }

  // store the instance
  if (tid < 7) {
    Bt.st[inst * 7 + tid] = B.scal[S_T + tid];
    Bt.out[inst * 7 + tid] = B.scal[O_INFLOW + tid];
  }
  if (tid == 0) {
    Bt.evCount[inst] = static_cast<int32_t>(B.scal[S_EVCOUNT]);
    Bt.foundered[inst] = static_cast<uint8_t>(B.scal[S_FOUNDERED]);
    Bt.founderT[inst] = B.scal[S_FOUNDERT];
    Bt.steps[inst] = done;
    if (Bt.stopWhenStable) Bt.calm[inst] = B.scal[S_CALM];
    if (Bt.curveEvery > 0) {
      Bt.nextCurve[inst] = B.scal[S_NEXTCURVE];
      Bt.curveCount[inst] = static_cast<int32_t>(B.scal[S_CURVECOUNT]);
    }
    readouts7(P, B.scal, B.vol, Bt.final7 + static_cast<size_t>(inst) * 7, 0);
  }
  for (int n = tid; n < NN; n += TPB) {
    Bt.vol[iN + n] = B.vol[n]; Bt.level[iN + n] = B.level[n];
    Bt.cx[iN + n] = B.cx[n]; Bt.cy[iN + n] = B.cy[n]; Bt.cz[iN + n] = B.cz[n];
  }
  for (int e = tid; e < evCap; e += TPB) { Bt.evT[iE + e] = B.evT[e]; Bt.evId[iE + e] = B.evId[e]; Bt.evKind[iE + e] = B.evKind[e]; }
  for (int m = tid; m < nmon; m += TPB) Bt.overT[iMon + m] = B.overT[m];
  for (int m = tid; m < nmark; m += TPB) Bt.markT[iMark + m] = B.markT[m];
}

// An epoch's plan (ss_batch_run in the resident form): [0] the count of live instances, [1] the workgroup size that
// runs them in this epoch, then the live instances in work order.
constexpr int kEpochHead = 2;

// The batch kernel. Persistent: each workgroup takes the next work item from the batch's counter, runs that
// instance (of its own model) to the end of the launch, and takes another, until the queue is empty; a workgroup
// whose instance founders or reaches tMax early refills itself with the next one, and every workgroup leaves the
// loop once the counter has passed the item count. Otherwise (when every instance has a resident workgroup of its
// own, so there is nothing to refill) workgroup k runs work item k; that form schedules a little better. With an
// epoch plan (resident form only), workgroup k runs the plan's k-th live instance, and a launch whose workgroup size
// the plan does not name, or a workgroup past the live count, returns at once.
template <bool Mixed, bool Persistent, int TPB>
__global__ __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(4, 8))) void run_batch(const ModelParams* __restrict__ models,
                                                                                                  const DevBatch* __restrict__ batch,
                                                                                                  long long maxSteps, double tMax,
                                                                                                  const int32_t* __restrict__ epochPlan) {
  BC& Bt = *(BC*)batch;
  extern __shared__ double lds_raw[];
  __shared__ int32_t item;
  SSL double* lds = (SSL double*)lds_raw;
  const int tid = threadIdx.x;
  if constexpr (!Persistent) {
    int inst;
    if (epochPlan) {
      const SSG int32_t* plan = (const SSG int32_t*)epochPlan;
      if (__builtin_amdgcn_readfirstlane(plan[1]) != TPB) return;
      if (static_cast<int>(blockIdx.x) >= __builtin_amdgcn_readfirstlane(plan[0])) return;
      inst = __builtin_amdgcn_readfirstlane(plan[kEpochHead + blockIdx.x]);
    } else {
      inst = __builtin_amdgcn_readfirstlane(Bt.order[blockIdx.x]);
    }
    PC& P = *(PC*)((const SSC char*)models + static_cast<size_t>(Bt.instModel[inst]) * sizeof(ModelParams));
    run_instance<Mixed, TPB>(P, Bt, inst, maxSteps, tMax, lds, tid);
    return;
  }
  for (;;) {
    if (tid == 0) item = atomicAdd((int32_t*)Bt.counter, 1);
    __syncthreads();
    const int k = __builtin_amdgcn_readfirstlane(item);  // uniform: the model's address stays in scalar registers
    __syncthreads();  // every wave has read the item before thread 0 takes the next one
    if (k >= Bt.count) break;
    const int inst = __builtin_amdgcn_readfirstlane(Bt.order[k]);
    PC& P = *(PC*)((const SSC char*)models + static_cast<size_t>(Bt.instModel[inst]) * sizeof(ModelParams));
    run_instance<Mixed, TPB>(P, Bt, inst, maxSteps, tMax, lds, tid);
    __syncthreads();  // the instance's LDS is free for the next one
  }
}

// Resets the batch's work counter before a launch (a kernel rather than a memset, so that the pair can be captured
// in a graph and replayed).
__global__ void reset_counter(int32_t* counter) { *counter = 0; }

// The packed double records: record i's five values from its 48-byte slot (6 doubles) to its 40-byte one.
__global__ void pack_records(const double* __restrict__ src, double* __restrict__ dst, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    for (int k = 0; k < 5; ++k) dst[5 * i + k] = src[6 * i + k];
}

// Plans an epoch of ss_batch_run: the instances of the work order that have not ended by the step loop's own test
// (foundered, t at or past tMax, or stopped by the calm rule), in work order, and their count; every instance for
// the first epoch (everyInstance), so that each is loaded and stored at least once, as by a single launch. The epoch
// runs them on lastTpb threads per simulation once all of them fit resident at that size (at most lastFit), else on
// tailTpb once they fit at that size (at most tailFit), else on mainTpb (a size of 0: that form is not used). One workgroup of 1,024 threads compacts the order a tile at a time (a ballot per wave, the waves' counts
// in LDS).
__global__ __launch_bounds__(1024) void plan_epoch(const DevBatch* __restrict__ batch, double tMax, int32_t* __restrict__ epochPlan,
                                                   int32_t mainTpb, int32_t tailTpb, int32_t tailFit, int32_t lastTpb, int32_t lastFit,
                                                   int32_t everyInstance) {
  BC& Bt = *(BC*)batch;
  SSG int32_t* plan = (SSG int32_t*)epochPlan;
  __shared__ int32_t waveLive[16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int32_t base = 0;
  for (int k0 = 0; k0 < Bt.count; k0 += 1024) {
    const int k = k0 + tid;
    bool live = false;
    int32_t inst = 0;
    if (k < Bt.count) {
      inst = Bt.order[k];
      const bool calm = Bt.stopWhenStable && Bt.calm[inst] > kCalmStop;
      live = everyInstance || !(Bt.foundered[inst] != 0 || Bt.st[static_cast<size_t>(inst) * 7] >= tMax || calm);
    }
    const uint64_t m = __ballot(live);
    if (lane == 0) waveLive[wave] = __popcll(m);
    __syncthreads();
    int32_t at = base, total = 0;
    for (int w = 0; w < 16; ++w) {
      if (w < wave) at += waveLive[w];
      total += waveLive[w];
    }
    if (live) plan[kEpochHead + at + __popcll(m & ((1ull << lane) - 1))] = inst;
    base += total;
    __syncthreads();  // every wave has read the counts before the next tile writes them
  }
  if (tid == 0) {
    plan[0] = base;
    plan[1] = lastTpb > 0 && base <= lastFit ? lastTpb : tailTpb > 0 && base <= tailFit ? tailTpb : mainTpb;
  }
}

}  // namespace ssa


namespace {


template <class T>
std::vector<T> repeat(const T* v, size_t n, size_t times) {
  std::vector<T> out;
  out.reserve(n * times);
  for (size_t i = 0; i < times; ++i) out.insert(out.end(), v, v + n);
  return out;
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

}  // extern "C"

namespace {

void check_options(const ss_batch_options* opt, const char* who) {
  const std::string w(who);
  need(opt != nullptr, w + ": null options");
  need(opt->count >= 1, w + ": count must be at least 1");
  need(opt->numerics == SS_NUMERICS_PORTABLE || opt->numerics == SS_NUMERICS_PORTABLE32,
       w + ": numerics must be portable (1) or portable32 (2); the oracle and std numerics are CPU-only");
  need(opt->traced >= 0 && opt->denseSteps >= 0, w + ": negative trace option");
  need(opt->traced == 0 || opt->sampleEvery > 0, w + ": sampleEvery must be positive when tracing");
  need(opt->tMax >= 0 && std::isfinite(opt->tMax), w + ": tMax must be finite and not negative");
  need(opt->curveEvery >= 0, w + ": curveEvery must not be negative");
}

// Uploads one model: geometry records, network tables (model.cpp:26-56 for the incidence), LDS layout.
std::shared_ptr<Model> build_model(const ss_ship& S, const ss_sim& M, size_t ldsLimit) {
  std::shared_ptr<Model> mp(new Model);
  Model& b = *mp;
  b.NN = S.nnode; b.NC = M.nconn; b.NG = M.ngroup; b.nmon = M.nmonitor; b.nmark = M.nmark; b.ncol = S.ncol; b.nseg = S.nseg;
  b.evCap = b.nmon + b.nmark + 1;
  for (int32_t g = 0; g < b.NG; ++g) b.maxGroup = std::max(b.maxGroup, M.gNodeCount[g]);
  size_t members = 0;  // length of the group member list the groups index into
  for (int32_t g = 0; g < b.NG; ++g) members = std::max(members, static_cast<size_t>(M.gNodeStart[g] + M.gNodeCount[g]));
  b.members = static_cast<int32_t>(members);
  b.nsea = static_cast<int32_t>(std::count_if(M.ca, M.ca + M.nconn, [](int32_t a) { return a < 0; }));
  if (b.lds_bytes(kThreads) > ldsLimit)
    fail(SS_RUNTIME, "ss_batch_create: this model needs " + std::to_string(b.lds_bytes(kThreads)) +
                         " bytes of LDS per simulation, the device allows " + std::to_string(ldsLimit));
  b.cost = static_cast<double>(S.ncol) + 2.0 * static_cast<double>(S.nseg) + 50.0 * static_cast<double>(S.nnode);

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

  const size_t ncn = static_cast<size_t>(M.nconn), ng = static_cast<size_t>(M.ngroup);
  b.gStart.upload(M.gNodeStart, ng); b.gCount.upload(M.gNodeCount, ng); b.gNodes.upload(M.gNodes, members); b.gMode.upload(M.gMode, ng);
  b.gpx.upload(M.gpx, ng); b.gpy.upload(M.gpy, ng); b.gpz.upload(M.gpz, ng);
  std::vector<int4> conI(std::max<size_t>(ncn, 1));
  std::vector<double4> conD(std::max<size_t>(ncn, 1));
  int32_t seaSlots = 0;  // a sea connection's a end carries -1 - its place in the sea list (increasing connection order)
  for (size_t c = 0; c < ncn; ++c) {
    conI[c] = make_int4(M.ca[c] >= 0 ? M.ca[c] : -1 - seaSlots++, M.cb[c], M.skipGroup[c], M.law[c]);
    conD[c] = make_double4(M.cx[c], M.cy[c], M.cz[c], M.coef[c]);
  }
  b.conI.upload(conI); b.conD.upload(conD);
  std::vector<int32_t> flowOrder;  // weirs (law 1) first, then orifices, each in increasing connection order
  for (int law = 1; law >= 0; --law)
    for (size_t c = 0; c < ncn; ++c)
      if ((M.law[c] != 0 ? 1 : 0) == law) flowOrder.push_back(static_cast<int32_t>(c));
  b.flowOrder.upload(flowOrder);
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
  d.ng = M.ngroup; d.nMembers = b.members; d.gStart = G(b.gStart.p); d.gCount = G(b.gCount.p); d.gNodes = G(b.gNodes.p); d.gMode = G(b.gMode.p);
  d.gpx = G(b.gpx.p); d.gpy = G(b.gpy.p); d.gpz = G(b.gpz.p); d.gMargin = M.gMargin;
  d.nc = M.nconn; d.conI = (const SSG i4v*)b.conI.p; d.conD = (const SSG d4v*)b.conD.p; d.flowOrder = G(b.flowOrder.p);
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

  b.en0.assign(M.en, M.en + ncn); b.area0.assign(M.area, M.area + ncn); b.tOn0.assign(M.tOn, M.tOn + ncn);
  b.coef0.assign(M.coef, M.coef + ncn);
  b.state0.assign(M.state0, M.state0 + 7);
  b.vol0.assign(M.vol0, M.vol0 + nn); b.level0.assign(M.level0, M.level0 + nn);
  b.cx0.assign(M.cx0, M.cx0 + nn); b.cy0.assign(M.cy0, M.cy0 + nn); b.cz0.assign(M.cz0, M.cz0 + nn);
  return mp;
}

// A batch over uploaded models: per-instance offsets, the work order (largest models first, so that the long
// simulations start early and the small ones fill in at the end), and the host staging of every instance's tables
// and state from its model's defaults.
std::unique_ptr<ss_batch> make_batch(const ss_batch_options& opt, std::vector<std::shared_ptr<const Model>> models, std::vector<int32_t> instModel) {
  std::unique_ptr<ss_batch> B(new ss_batch);
  ss_batch& b = *B;
  b.opt = opt;
  if (b.opt.traced > b.opt.count) b.opt.traced = b.opt.count;
  b.mixed = opt.numerics == SS_NUMERICS_PORTABLE32;
  b.models = std::move(models);
  b.instModel = std::move(instModel);
  const size_t count = static_cast<size_t>(b.opt.count);
  b.nodeOff.resize(count); b.connOff.resize(count); b.evOff.resize(count); b.monOff.resize(count); b.markOff.resize(count);
  for (size_t i = 0; i < count; ++i) {
    const Model& m = b.model(static_cast<int32_t>(i));
    b.nodeOff[i] = static_cast<int32_t>(b.nodeTotal); b.nodeTotal += static_cast<size_t>(m.NN);
    b.connOff[i] = static_cast<int32_t>(b.connTotal); b.connTotal += static_cast<size_t>(m.NC);
    b.evOff[i] = static_cast<int32_t>(b.evTotal); b.evTotal += static_cast<size_t>(m.evCap);
    b.monOff[i] = static_cast<int32_t>(b.monTotal); b.monTotal += static_cast<size_t>(m.nmon);
    b.markOff[i] = static_cast<int32_t>(b.markTotal); b.markTotal += static_cast<size_t>(m.nmark);
    b.hEn.insert(b.hEn.end(), m.en0.begin(), m.en0.end());
    b.hArea.insert(b.hArea.end(), m.area0.begin(), m.area0.end());
    b.hTOn.insert(b.hTOn.end(), m.tOn0.begin(), m.tOn0.end());
    b.hSt.insert(b.hSt.end(), m.state0.begin(), m.state0.end());
    b.hVol.insert(b.hVol.end(), m.vol0.begin(), m.vol0.end());
    b.hLevel.insert(b.hLevel.end(), m.level0.begin(), m.level0.end());
    b.hCx.insert(b.hCx.end(), m.cx0.begin(), m.cx0.end());
    b.hCy.insert(b.hCy.end(), m.cy0.begin(), m.cy0.end());
    b.hCz.insert(b.hCz.end(), m.cz0.begin(), m.cz0.end());
  }
  if (b.nodeTotal > static_cast<size_t>(INT32_MAX) || b.connTotal > static_cast<size_t>(INT32_MAX))
    fail(SS_INVALID_ARGUMENT, "ss_batch_create: the batch's node or connection rows exceed 2^31 entries");
  b.order.resize(count);
  for (size_t i = 0; i < count; ++i) b.order[i] = static_cast<int32_t>(i);
  std::stable_sort(b.order.begin(), b.order.end(), [&](int32_t x, int32_t y) { return b.model(x).cost > b.model(y).cost; });
  b.recSize = 10 + 5 * b.models[0]->NN;
  b.ldsBytes = forward_lds_bytes(b);
  b.maxSamples = b.opt.traced > 0 ? static_cast<int32_t>(std::floor(b.opt.tMax / b.opt.sampleEvery)) + 2 : 0;
  return B;
}

}  // namespace

size_t ssa::lds_limit() {
  hipDeviceProp_t p{};
  hip_check(hipGetDeviceProperties(&p, 0), "hipGetDeviceProperties");
  return p.sharedMemPerBlock;
}

extern "C" {

ss_status ss_batch_create(const ss_ship* ship, const ss_sim* sim, const ss_batch_options* opt, ss_batch** out) {
  return guarded([&] {
    need(out != nullptr, "ss_batch_create: null output handle");
    *out = nullptr;
    need(ship != nullptr, "ss_batch_create: null ship");
    need(sim != nullptr, "ss_batch_create: null simulation");
    check_options(opt, "ss_batch_create");
    validate(*ship, *sim, "ss_batch_create");
    std::string why;
    if (device_count(&why) == 0) fail(SS_RUNTIME, "ss_batch_create: " + why);
    std::vector<std::shared_ptr<const Model>> models;
    models.push_back(build_model(*ship, *sim, lds_limit()));
    *out = make_batch(*opt, std::move(models), std::vector<int32_t>(static_cast<size_t>(opt->count), 0)).release();
  });
}

ss_status ss_batch_create_fleet(const ss_ship* ships, const ss_sim* sims, int32_t nmodels, const int32_t* instModel,
                                const ss_batch_options* opt, ss_batch** out) {
  return guarded([&] {
    need(out != nullptr, "ss_batch_create_fleet: null output handle");
    *out = nullptr;
    need(ships != nullptr && sims != nullptr && instModel != nullptr, "ss_batch_create_fleet: null argument");
    need(nmodels >= 1, "ss_batch_create_fleet: at least one model");
    check_options(opt, "ss_batch_create_fleet");
    need(opt->traced == 0 || nmodels == 1, "ss_batch_create_fleet: traces are recorded for single-model batches only");
    for (int32_t k = 0; k < nmodels; ++k) validate(ships[k], sims[k], "ss_batch_create_fleet");
    std::vector<int32_t> im(instModel, instModel + opt->count);
    for (int32_t m : im) need(m >= 0 && m < nmodels, "ss_batch_create_fleet: instance model out of range");
    std::string why;
    if (device_count(&why) == 0) fail(SS_RUNTIME, "ss_batch_create_fleet: " + why);
    const size_t limit = lds_limit();
    std::vector<std::shared_ptr<const Model>> models;
    for (int32_t k = 0; k < nmodels; ++k) models.push_back(build_model(ships[k], sims[k], limit));
    *out = make_batch(*opt, std::move(models), std::move(im)).release();
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
    const size_t i = static_cast<size_t>(instance), nn = static_cast<size_t>(b->model(instance).NN), o = static_cast<size_t>(b->nodeOff[i]);
    std::copy(state7, state7 + 7, b->hSt.begin() + i * 7);
    std::copy(vol, vol + nn, b->hVol.begin() + o);
    std::copy(level, level + nn, b->hLevel.begin() + o);
    std::copy(cx, cx + nn, b->hCx.begin() + o);
    std::copy(cy, cy + nn, b->hCy.begin() + o);
    std::copy(cz, cz + nn, b->hCz.begin() + o);
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

}  // extern "C"

// Uploads the staged instances (connection tables as the products coef * area that the flow law multiplies first,
// initial states), clears every result array, and uploads the models at the batch's workgroup size and the batch
// view. The products are formed here exactly as the kernel formed them (one rounded multiplication), so the flows
// are unchanged bit for bit.
void Model::pack40() const {
  auto pack = [](const Dev<double>& from, Dev<double>& to) {
    if (to.p || !from.p) return;
    const size_t n = from.n / 6;
    to.alloc(n * 5);
    hipLaunchKernelGGL(pack_records, dim3(static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 4096))), dim3(256), 0, nullptr,
                       from.p, to.p, static_cast<int64_t>(n));
    hip_check(hipGetLastError(), "pack_records launch");
  };
  pack(col, col40);
  pack(seg, seg40);
  hip_check(hipDeviceSynchronize(), "pack_records");
}

void ssa::upload(ss_batch* b) {
  const size_t count = static_cast<size_t>(b->opt.count);
  b->ldsBytes = 0;
  for (const auto& m : b->models) b->ldsBytes = std::max(b->ldsBytes, m->lds_bytes(b->tpb));
  if (b->ldsBytes > lds_limit())
    fail(SS_RUNTIME, "ss_batch_run: " + std::to_string(b->ldsBytes) + " bytes of LDS per simulation at " + std::to_string(b->tpb) +
                         " threads exceed the device's limit");
  b->en.upload(b->hEn); b->area.upload(coef_area(b)); b->tOn.upload(b->hTOn);
  b->st.upload(b->hSt);
  b->out.fill(0.0, count * 7);
  b->vol.upload(b->hVol); b->level.upload(b->hLevel); b->ccx.upload(b->hCx); b->ccy.upload(b->hCy); b->ccz.upload(b->hCz);
  b->evT.fill(0.0, b->evTotal); b->evId.fill(0, b->evTotal); b->evKind.fill(0, b->evTotal); b->evCount.fill(0, count);
  b->overT.fill(-1.0, b->monTotal);
  b->markT.fill(-1.0, b->markTotal);
  b->foundered.fill(0, count); b->founderT.fill(-1.0, count); b->steps.fill(0, count); b->final7.fill(0.0, count * 7);
  b->calm.fill(0.0, b->stopWhenStable ? count : 1);
  const size_t traced = static_cast<size_t>(b->opt.traced), rs = static_cast<size_t>(b->recSize);
  b->denseRec.alloc(traced * static_cast<size_t>(b->opt.denseSteps + 1) * rs);
  b->sampleRec.alloc(traced * static_cast<size_t>(b->maxSamples) * rs);
  b->denseCount.fill(0, traced); b->sampleCount.fill(0, traced); b->nextSample.fill(b->opt.sampleEvery, traced);
  if (b->opt.curveEvery > 0) {
    b->curveMax = static_cast<int32_t>(std::floor(b->opt.tMax / b->opt.curveEvery)) + 2;
    b->curveRec.alloc(count * static_cast<size_t>(b->curveMax) * kCurveStride);
    b->curveCount.fill(0, count);
    b->nextCurve.fill(0.0, count);  // the first record is the initial state, as in the oracle's run
  }
  b->dInstModel.upload(b->instModel); b->dOrder.upload(b->order);
  b->dNodeOff.upload(b->nodeOff); b->dConnOff.upload(b->connOff); b->dEvOff.upload(b->evOff);
  b->dMonOff.upload(b->monOff); b->dMarkOff.upload(b->markOff);
  b->counter.fill(0, 1);
  std::vector<ModelParams> mp;
  for (const auto& m : b->models) {
    if (b->recBytes == 40) m->pack40();
    mp.push_back(m->params(b->tpb, b->recBytes));
  }
  b->dModels.upload(mp);
  const DevBatch bv = b->batchView();
  b->dBatch.upload(&bv, 1);
}

namespace {

using BatchKernel = void (*)(const ModelParams*, const DevBatch*, long long, double, const int32_t*);

BatchKernel batch_kernel(bool mixed, bool persistent, int tpb) {
  if (tpb == 1024) {
    if (persistent) return mixed ? run_batch<true, true, 1024> : run_batch<false, true, 1024>;
    return mixed ? run_batch<true, false, 1024> : run_batch<false, false, 1024>;
  }
  if (tpb == 512) {
    if (persistent) return mixed ? run_batch<true, true, 512> : run_batch<false, true, 512>;
    return mixed ? run_batch<true, false, 512> : run_batch<false, false, 512>;
  }
  if (persistent) return mixed ? run_batch<true, true, kThreads> : run_batch<false, true, kThreads>;
  return mixed ? run_batch<true, false, kThreads> : run_batch<false, false, kThreads>;
}

// The launch shape of a batch after its upload: the kernel (numerics, refilling or not, workgroup size) and the
// grid. Every kernel instantiation that may run gets the batch's dynamic LDS size allowed first.
struct LaunchPlan {
  BatchKernel kernel;
  int grid;
};

LaunchPlan plan_launch(ss_batch* b) {
  const size_t count = static_cast<size_t>(b->opt.count);
  for (int tpb : {kThreads, 512, 1024})
    for (bool pers : {false, true})
      for (bool mixed : {false, true})
        if (b->ldsBytes > 64 * 1024)
          hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(batch_kernel(mixed, pers, tpb)), hipFuncAttributeMaxDynamicSharedMemorySize,
                                        static_cast<int>(b->ldsBytes)),
                    "hipFuncSetAttribute");
  // the persistent grid: as many workgroups as are resident on the device at once (or the set count)
  int perCU = 0, cus = 0;
  hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, reinterpret_cast<const void*>(batch_kernel(b->mixed, true, b->tpb)), b->tpb,
                                                         b->ldsBytes),
            "occupancy");
  hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0), "hipDeviceGetAttribute");
  const int resident = std::max(1, perCU) * std::max(1, cus);
  const int grid = static_cast<int>(std::min<long long>(static_cast<long long>(count), b->blocks > 0 ? b->blocks : resident));
  // one workgroup per instance when all of them are resident at once; the refilling loop otherwise
  b->refill = grid < static_cast<int>(count);
  return LaunchPlan{batch_kernel(b->mixed, b->refill, b->tpb), grid};
}

// The steps of an epoch: the batch's stepsPerLaunch when set, else a 32nd of the most steps any instance can still
// take before tMax (at least 256, at most 4,096).
long long epoch_steps(const ss_batch* b) {
  if (b->opt.stepsPerLaunch > 0) return static_cast<long long>(b->opt.stepsPerLaunch);
  double most = 0;
  for (int32_t i = 0; i < b->opt.count; ++i) {
    const double left = b->opt.tMax - b->hSt[static_cast<size_t>(i) * 7];
    if (left > 0) most = std::max(most, std::ceil(left / b->model(i).net.dt) + 2);
  }
  return std::clamp(static_cast<long long>(std::ceil(most / 32)), 256LL, 4096LL);
}

// The resident form of a run as epochs. The batch kernel runs each instance that plan_epoch has listed for up to
// epoch_steps steps on a workgroup of its own, so that the device's dispatcher spreads the live instances evenly
// over the compute units (one long instance on a compute unit with three others that finished early otherwise runs
// alone there while other compute units hold several). With the automatic launch shape (tailForm) at 256 threads,
// once the live instances fit resident at 512 threads the epoch runs them on the 512-thread form, and once they fit
// one per compute unit on the 1,024-thread form: the step of an instance with a compute unit (or half of one) to
// itself is shorter on more threads, its level solve spread over more waves. The epoch (main form, the two tail forms,
// the next epoch's plan) is one graph, replayed in groups of eight without host waits; the plan after each group counts the instances still
// live, and the host stops once a group has left none (seen one group later, so the device never waits for the
// host). Epochs change the launches, not the results: an instance's state is in global memory between launches, as
// in chunked runs.
void run_epochs(ss_batch* b, const LaunchPlan& lp, hipStream_t stream, hipEvent_t e0, hipEvent_t e1) {
  const int32_t count = b->opt.count;
  const long long steps = epoch_steps(b);
  b->epochPlan.alloc(static_cast<size_t>(kEpochHead + count));
  // the tail forms: 512 threads per simulation once the live instances fit resident at that size, 1,024 once they fit
  // one per compute unit (each form only where every model's layout fits its LDS and a workgroup is resident)
  struct Form {
    int32_t tpb = 0, fit = 0;
    size_t lds = 0;
  } forms[2];
  if (b->tailForm && b->tpb == kThreads) {
    Dev<ModelParams>* params[2] = {&b->dModelsTail, &b->dModelsLast};
    const int32_t sizes[2] = {512, 1024};
    for (int f = 0; f < 2; ++f) {
      size_t lds = 0;
      for (const auto& m : b->models) lds = std::max(lds, m->lds_bytes(sizes[f]));
      if (lds > lds_limit()) break;
      const void* k = reinterpret_cast<const void*>(batch_kernel(b->mixed, false, sizes[f]));
      if (lds > 64 * 1024)
        hip_check(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)), "hipFuncSetAttribute");
      int perCU = 0, cus = 0;
      hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k, sizes[f], lds), "occupancy");
      hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0), "hipDeviceGetAttribute");
      if (perCU <= 0) break;
      forms[f].tpb = sizes[f];
      forms[f].fit = perCU * cus;
      forms[f].lds = lds;
      std::vector<ModelParams> mp;
      for (const auto& m : b->models) mp.push_back(m->params(sizes[f], b->recBytes));
      params[f]->upload(mp);
    }
  }
  const int32_t tailTpb = forms[0].tpb, tailFit = forms[0].fit, lastTpb = forms[1].tpb, lastFit = forms[1].fit;
  const DevBatch* dB = static_cast<const DevBatch*>(b->dBatch.p);
  const double tMax = b->opt.tMax;
  int32_t* plan = b->epochPlan.p;

  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t g[2] = {nullptr, nullptr};
  int32_t* live = nullptr;  // pinned: the live count after each group, two groups in flight
  struct Release {
    hipGraph_t& gr; hipGraphExec_t& x; hipEvent_t* ev; int32_t*& h;
    ~Release() {
      if (x) (void)hipGraphExecDestroy(x);
      if (gr) (void)hipGraphDestroy(gr);
      for (int i = 0; i < 2; ++i)
        if (ev[i]) (void)hipEventDestroy(ev[i]);
      if (h) (void)hipHostFree(h);
    }
  } release{graph, exec, g, live};
  for (auto& e : g) hip_check(hipEventCreate(&e), "hipEventCreate");
  hip_check(hipHostMalloc(reinterpret_cast<void**>(&live), 2 * sizeof(int32_t), hipHostMallocDefault), "hipHostMalloc");
  live[0] = live[1] = -1;

  hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  hipLaunchKernelGGL(lp.kernel, dim3(lp.grid), dim3(b->tpb), b->ldsBytes, stream, static_cast<const ModelParams*>(b->dModels.p), dB, steps,
                     tMax, plan);
  if (tailTpb)
    hipLaunchKernelGGL(batch_kernel(b->mixed, false, tailTpb), dim3(std::min(count, tailFit)), dim3(tailTpb), forms[0].lds, stream,
                       static_cast<const ModelParams*>(b->dModelsTail.p), dB, steps, tMax, plan);
  if (lastTpb)
    hipLaunchKernelGGL(batch_kernel(b->mixed, false, lastTpb), dim3(std::min(count, lastFit)), dim3(lastTpb), forms[1].lds, stream,
                       static_cast<const ModelParams*>(b->dModelsLast.p), dB, steps, tMax, plan);
  hipLaunchKernelGGL(plan_epoch, dim3(1), dim3(1024), 0, stream, dB, tMax, plan, b->tpb, tailTpb, tailFit, lastTpb, lastFit, 0);
  hip_check(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
  hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");

  constexpr int kGroup = 8;
  hip_check(hipEventRecord(e0, stream), "hipEventRecord");
  hipLaunchKernelGGL(plan_epoch, dim3(1), dim3(1024), 0, stream, dB, tMax, plan, b->tpb, tailTpb, tailFit, lastTpb, lastFit, 1);
  hip_check(hipGetLastError(), "plan_epoch launch");
  for (int grp = 0;; ++grp) {
    for (int e = 0; e < kGroup; ++e) hip_check(hipGraphLaunch(exec, stream), "hipGraphLaunch");
    b->launches += kGroup;
    hip_check(hipMemcpyAsync(live + (grp & 1), plan, sizeof(int32_t), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
    hip_check(hipEventRecord(g[grp & 1], stream), "hipEventRecord");
    if (grp > 0) {
      hip_check(hipEventSynchronize(g[(grp - 1) & 1]), "kernel execution");
      if (live[(grp - 1) & 1] == 0) break;
    }
  }
  hip_check(hipEventRecord(e1, stream), "hipEventRecord");
  hip_check(hipEventSynchronize(e1), "kernel execution");
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
  b->kernelMs += ms;
}

// Uploads the staged instances, runs them to their end and reads the results back. The launch shape is the batch's
// (tpb, blocks). With `epochs` and every instance resident at once, as epochs (run_epochs); otherwise one graph
// launch, or one per chunk of stepsPerLaunch steps, at most maxLaunches.
void execute(ss_batch* b, double* seconds, bool epochs, int maxLaunches = INT_MAX) {
  const size_t count = static_cast<size_t>(b->opt.count);
  upload(b);

  const long long maxSteps = b->opt.stepsPerLaunch > 0 ? static_cast<long long>(b->opt.stepsPerLaunch) : LLONG_MAX;
  const LaunchPlan lp = plan_launch(b);

  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  struct Release {
    hipStream_t& s; hipGraph_t& g; hipGraphExec_t& x; hipEvent_t& a; hipEvent_t& c;
    ~Release() {
      if (x) (void)hipGraphExecDestroy(x);
      if (g) (void)hipGraphDestroy(g);
      if (a) (void)hipEventDestroy(a);
      if (c) (void)hipEventDestroy(c);
      if (s) (void)hipStreamDestroy(s);
    }
  } release{stream, graph, exec, e0, e1};
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  b->kernelMs = 0;
  b->launches = 0;
  if (epochs && !b->refill) {
    const auto t0 = std::chrono::steady_clock::now();
    run_epochs(b, lp, stream, e0, e1);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    read_back(b);
    if (seconds) *seconds = secs;
    return;
  }

  // The launch pair (counter reset, batch kernel) is captured once in a graph on a stream of the batch's own and
  // replayed for every chunk: identical launches, one submission each.
  hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, stream, b->counter.p);
  hipLaunchKernelGGL(lp.kernel, dim3(lp.grid), dim3(b->tpb), b->ldsBytes, stream, static_cast<const ModelParams*>(b->dModels.p),
                     static_cast<const DevBatch*>(b->dBatch.p), maxSteps, b->opt.tMax, nullptr);
  hip_check(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
  hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<uint8_t> f(count);
  std::vector<double> st(count * 7), calm(count, 0.0);
  for (;;) {
    hip_check(hipEventRecord(e0, stream), "hipEventRecord");
    hip_check(hipGraphLaunch(exec, stream), "hipGraphLaunch");
    hip_check(hipEventRecord(e1, stream), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "kernel execution");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    b->kernelMs += ms;
    b->launches++;
    if (maxSteps == LLONG_MAX || b->launches >= maxLaunches) break;
    b->foundered.download(f.data(), count);
    b->st.download(st.data(), count * 7);
    if (b->stopWhenStable) b->calm.download(calm.data(), count);
    bool active = false;
    for (size_t i = 0; i < count; ++i)
      if (!f[i] && st[i * 7] < b->opt.tMax && !(calm[i] > kCalmStop)) { active = true; break; }
    if (!active) break;
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  read_back(b);
  if (seconds) *seconds = secs;
}

// One launch in exact-step mode: instance i takes exactly limits[i] steps (or k each when limits is null) from the
// device state (uploaded from the staged instances at the first call), founder and tMax notwithstanding.
void step_exact(ss_batch* b, long long k, const int64_t* limits, double* seconds) {
  const size_t count = static_cast<size_t>(b->opt.count);
  if (!b->stepped) upload(b);
  b->stepped = true;
  long long most = k;
  if (limits) {
    std::vector<long long> l(limits, limits + count);
    most = *std::max_element(l.begin(), l.end());
    b->stepLimit.upload(l);
  } else {
    b->stepLimit.release();
  }
  b->exact = true;
  const DevBatch bv = b->batchView();
  b->exact = false;
  b->dBatch.upload(&bv, 1);
  const LaunchPlan lp = plan_launch(b);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  struct Release {
    hipEvent_t& a; hipEvent_t& c;
    ~Release() {
      if (a) (void)hipEventDestroy(a);
      if (c) (void)hipEventDestroy(c);
    }
  } release{e0, e1};
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  const auto t0 = std::chrono::steady_clock::now();
  hip_check(hipEventRecord(e0, nullptr), "hipEventRecord");
  hipLaunchKernelGGL(reset_counter, dim3(1), dim3(1), 0, nullptr, b->counter.p);
  if (most > 0)
    hipLaunchKernelGGL(lp.kernel, dim3(lp.grid), dim3(b->tpb), b->ldsBytes, nullptr, static_cast<const ModelParams*>(b->dModels.p),
                       static_cast<const DevBatch*>(b->dBatch.p), most, DBL_MAX, nullptr);
  hip_check(hipGetLastError(), "kernel launch");
  hip_check(hipEventRecord(e1, nullptr), "hipEventRecord");
  hip_check(hipEventSynchronize(e1), "kernel execution");
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
  b->kernelMs += ms;
  b->launches++;
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  read_back(b);
  if (seconds) *seconds = secs;
}

}  // namespace

extern "C" {

ss_status ss_batch_set_launch(ss_batch* b, int32_t threadsPerSim, int32_t workgroups) {
  return guarded([&] {
    need(b != nullptr, "ss_batch_set_launch: null batch");
    if (b->ran) fail(SS_LOGIC, "ss_batch_set_launch: set after run");
    need(threadsPerSim == 0 || threadsPerSim == kThreads || threadsPerSim == 512 || threadsPerSim == 1024,
         "ss_batch_set_launch: threadsPerSim must be 0, 256, 512 or 1024");
    need(workgroups >= 0, "ss_batch_set_launch: negative workgroup count");
    b->tpb = threadsPerSim == 0 ? kThreads : threadsPerSim;
    b->tailForm = threadsPerSim == 0;
    b->blocks = workgroups;
    b->ldsBytes = forward_lds_bytes(*b);
  });
}

ss_status ss_batch_set_record_layout(ss_batch* b, int32_t recordBytes) {
  return guarded([&] {
    need(b != nullptr, "ss_batch_set_record_layout: null batch");
    if (b->ran || b->stepped) fail(SS_LOGIC, "ss_batch_set_record_layout: set after run or step");
    need(recordBytes == 48 || recordBytes == 40, "ss_batch_set_record_layout: recordBytes must be 48 or 40");
    b->recBytes = recordBytes;
  });
}

ss_status ss_batch_set_stop_when_stable(ss_batch* b, int32_t on) {
  return guarded([&] {
    need(b != nullptr, "ss_batch_set_stop_when_stable: null batch");
    if (b->ran) fail(SS_LOGIC, "ss_batch_set_stop_when_stable: set after run");
    need(on == 0 || on == 1, "ss_batch_set_stop_when_stable: on must be 0 or 1");
    b->stopWhenStable = on == 1;
  });
}

ss_status ss_batch_autotune(ss_batch* b, int32_t sample, int64_t trialSteps, ss_tune_result* out) {
  return guarded([&] {
    need(b != nullptr, "ss_batch_autotune: null batch");
    if (b->ran) fail(SS_LOGIC, "ss_batch_autotune: tune after run");
    need(sample >= 1 && trialSteps >= 1, "ss_batch_autotune: sample and trialSteps must be positive");
    // the trial batch: the first `sample` instances of the work order (the largest models), with their staged
    // tables and states, over the same uploaded models
    const int32_t n = std::min(sample, b->opt.count);
    std::vector<int32_t> im(static_cast<size_t>(n));
    for (int32_t k = 0; k < n; ++k) im[static_cast<size_t>(k)] = b->instModel[static_cast<size_t>(b->order[static_cast<size_t>(k)])];
    ss_batch_options o = b->opt;
    o.count = n;
    o.traced = 0;
    o.curveEvery = 0;
    o.stepsPerLaunch = trialSteps;
    const int tpbs[3] = {kThreads, 512, 1024};
    double best = -1;
    int32_t bestTpb = kThreads, bestBlocks = 0, bestRec = b->recBytes;
    ss_tune_result res{};
    int cus = 0;
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0), "hipDeviceGetAttribute");
    // one launch of trialSteps steps for every sampled instance at a shape and record layout: steps/s
    auto trial = [&](int tpb, int32_t grid, int32_t rec) {
      std::unique_ptr<ss_batch> t = make_batch(o, b->models, im);
      for (int32_t k = 0; k < n; ++k) {  // copy the staged tables and states of the sampled instances
        const size_t src = static_cast<size_t>(b->order[static_cast<size_t>(k)]), dst = static_cast<size_t>(k);
        const Model& m = b->model(static_cast<int32_t>(src));
        std::copy_n(b->hEn.begin() + b->connOff[src], m.NC, t->hEn.begin() + t->connOff[dst]);
        std::copy_n(b->hArea.begin() + b->connOff[src], m.NC, t->hArea.begin() + t->connOff[dst]);
        std::copy_n(b->hTOn.begin() + b->connOff[src], m.NC, t->hTOn.begin() + t->connOff[dst]);
        std::copy_n(b->hSt.begin() + 7 * src, 7, t->hSt.begin() + 7 * dst);
        for (auto [from, to] : {std::make_pair(&b->hVol, &t->hVol), std::make_pair(&b->hLevel, &t->hLevel), std::make_pair(&b->hCx, &t->hCx),
                                std::make_pair(&b->hCy, &t->hCy), std::make_pair(&b->hCz, &t->hCz)})
          std::copy_n(from->begin() + b->nodeOff[src], m.NN, to->begin() + t->nodeOff[dst]);
      }
      t->tpb = tpb;
      t->blocks = grid;
      t->recBytes = rec;
      double secs = 0;
      execute(t.get(), &secs, false, 1);
      const double rate = t->kernelMs > 0 ? static_cast<double>(t->totalSteps) / (t->kernelMs * 1e-3) : 0;
      if (res.trials < 8) {
        res.trialThreads[res.trials] = tpb;
        res.trialWorkgroups[res.trials] = grid;
        res.trialRate[res.trials] = rate;
        res.trialRecordBytes[res.trials] = rec;
        ++res.trials;
      }
      return rate;
    };
    for (int c = 0; c < 3; ++c) {
      size_t lds = 0;
      for (const auto& m : b->models) lds = std::max(lds, m->lds_bytes(tpbs[c]));
      if (lds > lds_limit()) continue;  // this shape does not fit the largest model
      // the grids: the default (every instance resident, or as many workgroups as are resident at once) and, when the
      // sample exceeds the resident count, three quarters and half of it (whole workgroups per compute unit)
      int perCU = 0;
      hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(batch_kernel(b->mixed, true, tpbs[c])),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)),
                "hipFuncSetAttribute");
      hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, reinterpret_cast<const void*>(batch_kernel(b->mixed, true, tpbs[c])),
                                                             tpbs[c], lds),
                "occupancy");
      std::vector<int32_t> grids{0};
      if (perCU * cus < n)
        for (int num : {3, 2})
          if (const int per = perCU * num / 4; per >= 1 && per < perCU && per * cus != grids.back()) grids.push_back(per * cus);
      for (int32_t grid : grids) {
        const double rate = trial(tpbs[c], grid, b->recBytes);
        res.rate[c] = std::max(res.rate[c], rate);
        if (rate > best) { best = rate; bestTpb = tpbs[c]; bestBlocks = grid; }
      }
    }
    // the other layout of the double records at the fastest shape (the float records have one layout)
    if (!b->mixed) {
      const int32_t other = b->recBytes == 48 ? 40 : 48;
      if (const double rate = trial(bestTpb, bestBlocks, other); rate > best) { best = rate; bestRec = other; }
    }
    b->tpb = bestTpb;
    b->tailForm = true;
    b->blocks = bestBlocks;
    b->recBytes = bestRec;
    b->ldsBytes = forward_lds_bytes(*b);
    res.threadsPerSim = bestTpb;
    res.workgroups = bestBlocks;
    res.recordBytes = bestRec;
    if (out) *out = res;
  });
}

ss_status ss_batch_run(ss_batch* b, double* seconds) {
  return guarded([&] {
    need(b != nullptr, "ss_batch_run: null batch");
    if (b->stepped) fail(SS_LOGIC, "ss_batch_run: run after ss_batch_step");
    if (b->ran) fail(SS_LOGIC, "ss_batch_run: run twice");
    b->ran = true;
    execute(b, seconds, true);
  });
}

ss_status ss_batch_step(ss_batch* b, int64_t steps, const int64_t* perInstance, double* seconds) {
  return guarded([&] {
    need(b != nullptr, "ss_batch_step: null batch");
    if (b->ran && !b->stepped) fail(SS_LOGIC, "ss_batch_step: step after run");
    if (perInstance) {
      for (int32_t i = 0; i < b->opt.count; ++i) need(perInstance[i] >= 0, "ss_batch_step: negative step count");
    } else {
      need(steps >= 0, "ss_batch_step: negative step count");
    }
    b->ran = true;
    step_exact(b, steps, perInstance, seconds);
  });
}

ss_status ss_batch_state_records(const ss_batch* b, double* records) {
  return guarded([&] {
    need(b != nullptr && records != nullptr, "ss_batch_state_records: null argument");
    if (!b->ran) fail(SS_LOGIC, "ss_batch_state_records: records before run");
    const std::vector<double> vol = b->vol.all(), level = b->level.all(), cx = b->ccx.all(), cy = b->ccy.all(), cz = b->ccz.all();
    double* r = records;
    for (int32_t i = 0; i < b->opt.count; ++i) {
      const size_t k = static_cast<size_t>(i), nn = static_cast<size_t>(b->model(i).NN), o = static_cast<size_t>(b->nodeOff[k]);
      for (int j = 0; j < 7; ++j) r[j] = b->rSt[k * 7 + j];
      for (int j = 0; j < 3; ++j) r[7 + j] = b->rOut[k * 7 + j];
      r += 10;
      for (const std::vector<double>* a : {&vol, &level, &cx, &cy, &cz}) {
        std::copy(a->begin() + static_cast<std::ptrdiff_t>(o), a->begin() + static_cast<std::ptrdiff_t>(o + nn), r);
        r += nn;
      }
    }
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
    const size_t i = static_cast<size_t>(instance), nn = static_cast<size_t>(b->model(instance).NN), o = static_cast<size_t>(b->nodeOff[i]);
    if (res) {
      res->foundered = b->rFoundered[i];
      res->founderT = b->rFounderT[i];
      res->tEnd = b->rSt[i * 7];
      res->steps = b->rSteps[i];
      for (int k = 0; k < 7; ++k) { res->state[k] = b->rSt[i * 7 + k]; res->outputs[k] = b->rOut[i * 7 + k]; }
      res->events = b->rEvCount[i];
    }
    if (vol) b->vol.download(vol, nn, o);
    if (level) b->level.download(level, nn, o);
    if (cx) b->ccx.download(cx, nn, o);
    if (cy) b->ccy.download(cy, nn, o);
    if (cz) b->ccz.download(cz, nn, o);
    if (events && evCap > 0) {
      const int32_t n = std::min(b->rEvCount[i], evCap);
      std::vector<double> t(n);
      std::vector<int32_t> id(n);
      std::vector<uint8_t> kind(n);
      const size_t off = static_cast<size_t>(b->evOff[i]);
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
    need(denseCap >= 0 && samplesCap >= 0, "ss_batch_trace: negative capacity");
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
      std::vector<double> padded(static_cast<size_t>(keep) * kCurveStride);
      b->curveRec.download(padded.data(), padded.size(), i * static_cast<size_t>(b->curveMax) * kCurveStride);
      for (int32_t k = 0; k < keep; ++k)
        std::copy_n(padded.begin() + static_cast<std::ptrdiff_t>(k) * kCurveStride, kCurveValues, rec + static_cast<size_t>(k) * kCurveValues);
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
    if (conn < 0 || conn >= b->model(instance).NC)
      fail(SS_OUT_OF_RANGE, std::string(what) + ": connection " + std::to_string(conn) + " out of range");
    f(static_cast<size_t>(b->connOff[static_cast<size_t>(instance)]) + static_cast<size_t>(conn));
  });
}

// ------------------------------------------------------------------------------------------- allocation-free stepping
//
// ss_batch_step_noalloc: ss_batch_step for callers that may not allocate device memory. The caller passes device
// inputs and outputs; everything else the launch needs (the batch and model views the kernel reads, the work counter,
// the per-instance offsets, state, event bookkeeping and connection products) lives in one file-scope device array of
// fixed size, `scratch`. Host code never passes the array's name to the device: its device address is looked up once
// per device (hipGetSymbolAddress, with that device current) and kept per device.

namespace ssa {

constexpr size_t kNoAllocScratchBytes = 16u << 20;  // per device, held until the process ends
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// the staging kernel takes both views by value (kernel arguments are limited to 4 KiB)
static_assert(sizeof(ModelParams) + sizeof(DevBatch) + 6 * sizeof(void*) <= 4096, "views too large for kernel arguments");

// Where each per-call array sits in the scratch array (byte offsets, 256-byte aligned).
struct NoAllocPlan {
  size_t models, batch, counter, instModel, order, nodeOff, connOff, evOff, monOff, markOff, carea, st, out, final7, vol, level, cx, cy, cz,
      evT, evId, evKind, evCount, overT, markT, foundered, founderT, steps, calm, total;
};

inline NoAllocPlan noalloc_plan(size_t count, const Model& m) {
  NoAllocPlan p{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = (off + std::max<size_t>(bytes, 1) + 255) & ~static_cast<size_t>(255);
    return at;
  };
  const size_t NN = static_cast<size_t>(m.NN), NC = static_cast<size_t>(m.NC), ev = static_cast<size_t>(m.evCap);
  p.models = take(sizeof(ModelParams)); p.batch = take(sizeof(DevBatch)); p.counter = take(sizeof(int32_t));
  p.instModel = take(count * 4); p.order = take(count * 4); p.nodeOff = take(count * 4); p.connOff = take(count * 4);
  p.evOff = take(count * 4); p.monOff = take(count * 4); p.markOff = take(count * 4);
  p.carea = take(count * NC * 8);
  p.st = take(count * 7 * 8); p.out = take(count * 7 * 8); p.final7 = take(count * 7 * 8);
  p.vol = take(count * NN * 8); p.level = take(count * NN * 8); p.cx = take(count * NN * 8); p.cy = take(count * NN * 8);
  p.cz = take(count * NN * 8);
  p.evT = take(count * ev * 8); p.evId = take(count * ev * 4); p.evKind = take(count * ev); p.evCount = take(count * 4);
  p.overT = take(count * static_cast<size_t>(m.nmon) * 8); p.markT = take(count * static_cast<size_t>(m.nmark) * 8);
  p.foundered = take(count); p.founderT = take(count * 8); p.steps = take(count * 8); p.calm = take(count * 8);
  p.total = off;
  return p;
}

// One workgroup per instance: the views (workgroup 0), the instance's offsets, its state unpacked from its record
// (trace layout), its event bookkeeping cleared and its connection products coef * area (one rounded multiplication,
// as the host forms them for ss_batch_run and ss_batch_step).
__global__ __launch_bounds__(256) void noalloc_stage(ModelParams mp, DevBatch bv, ModelParams* dstM, DevBatch* dstB, int32_t* counter,
                                                      const double* __restrict__ recIn, const double* __restrict__ area, double* carea) {
  const int i = blockIdx.x, t = threadIdx.x;
  const int NN = mp.S.nn, NC = mp.M.nc, ev = mp.M.evCap, nmon = mp.M.nmon, nmark = mp.M.nmark;
  if (i == 0 && t == 0) {
    *dstM = mp;
    *dstB = bv;
    *counter = 0;
  }
  SSG int32_t* im = const_cast<SSG int32_t*>(bv.instModel);
  SSG int32_t* ord = const_cast<SSG int32_t*>(bv.order);
  if (t == 0) {
    im[i] = 0; ord[i] = i;
    const_cast<SSG int32_t*>(bv.nodeOff)[i] = i * NN; const_cast<SSG int32_t*>(bv.connOff)[i] = i * NC;
    const_cast<SSG int32_t*>(bv.evOff)[i] = i * ev; const_cast<SSG int32_t*>(bv.monOff)[i] = i * nmon;
    const_cast<SSG int32_t*>(bv.markOff)[i] = i * nmark;
    bv.evCount[i] = 0; bv.foundered[i] = 0; bv.founderT[i] = -1.0; bv.steps[i] = 0; bv.calm[i] = 0.0;
  }
  const size_t rs = static_cast<size_t>(10 + 5 * NN);
  const double* r = recIn + static_cast<size_t>(i) * rs;
  if (t < 7) {
    bv.st[i * 7 + t] = r[t];
    bv.out[i * 7 + t] = t < 3 ? r[7 + t] : 0.0;
    bv.final7[i * 7 + t] = 0.0;
  }
  const size_t iN = static_cast<size_t>(i) * NN;
  for (int n = t; n < NN; n += 256) {
    bv.vol[iN + n] = r[10 + n]; bv.level[iN + n] = r[10 + NN + n];
    bv.cx[iN + n] = r[10 + 2 * NN + n]; bv.cy[iN + n] = r[10 + 3 * NN + n]; bv.cz[iN + n] = r[10 + 4 * NN + n];
  }
  const size_t iC = static_cast<size_t>(i) * NC;
  for (int c = t; c < NC; c += 256) carea[iC + c] = mp.M.conD[c].w * area[iC + c];
  for (int e = t; e < ev; e += 256) {
    bv.evT[static_cast<size_t>(i) * ev + e] = 0.0; bv.evId[static_cast<size_t>(i) * ev + e] = 0; bv.evKind[static_cast<size_t>(i) * ev + e] = 0;
  }
  for (int k = t; k < nmon; k += 256) bv.overT[static_cast<size_t>(i) * nmon + k] = -1.0;
  for (int k = t; k < nmark; k += 256) bv.markT[static_cast<size_t>(i) * nmark + k] = -1.0;
}

// One workgroup per instance: its state after the launch, packed into its record (trace layout).
__global__ __launch_bounds__(256) void noalloc_pack(const DevBatch* __restrict__ batch, int32_t nn, double* __restrict__ recOut) {
  const DevBatch& bv = *batch;
  const int i = blockIdx.x, t = threadIdx.x;
  const size_t rs = static_cast<size_t>(10 + 5 * nn);
  double* r = recOut + static_cast<size_t>(i) * rs;
  if (t < 7) r[t] = bv.st[i * 7 + t];
  if (t < 3) r[7 + t] = bv.out[i * 7 + t];
  const size_t iN = static_cast<size_t>(i) * nn;
  for (int n = t; n < nn; n += 256) {
    r[10 + n] = bv.vol[iN + n]; r[10 + nn + n] = bv.level[iN + n];
    r[10 + 2 * nn + n] = bv.cx[iN + n]; r[10 + 3 * nn + n] = bv.cy[iN + n]; r[10 + 4 * nn + n] = bv.cz[iN + n];
  }
}

}  // namespace ssa

namespace {

std::mutex g_scratchLock;
void* g_scratchAt[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up at the first call on that device.
unsigned char* scratch_address() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) fail(SS_RUNTIME, "ss_batch_step_noalloc: device index beyond the library's table");
  std::lock_guard<std::mutex> hold(g_scratchLock);
  if (!g_scratchAt[dev]) {
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratchAt[dev] = p;
  }
  return static_cast<unsigned char*>(g_scratchAt[dev]);
}

}  // namespace

extern "C" {

ss_status ss_batch_step_noalloc(const ss_batch* b, int32_t count, int64_t steps, const double* recIn, const uint8_t* en, const double* area,
                                const double* tOn, double* recOut, void* stream) {
  return guarded([&] {
    need(b != nullptr, "ss_batch_step_noalloc: null batch");
    need(b->models.size() == 1, "ss_batch_step_noalloc: single-model batches only");
    need(count >= 1, "ss_batch_step_noalloc: count must be at least 1");
    need(steps >= 0, "ss_batch_step_noalloc: negative step count");
    need(recIn && en && area && tOn && recOut, "ss_batch_step_noalloc: null array");
    const Model& m = *b->models[0];
    const NoAllocPlan plan = noalloc_plan(static_cast<size_t>(count), m);
    if (plan.total > kNoAllocScratchBytes)
      fail(SS_INVALID_ARGUMENT, "ss_batch_step_noalloc: " + std::to_string(count) + " instances need " + std::to_string(plan.total) +
                                    " bytes of scratch, the maximum is " + std::to_string(kNoAllocScratchBytes));
    const size_t lds = m.lds_bytes(b->tpb);
    if (lds > lds_limit()) fail(SS_RUNTIME, "ss_batch_step_noalloc: the model's LDS exceeds the device's limit at this launch shape");
    unsigned char* base = scratch_address();
    const hipStream_t s = static_cast<hipStream_t>(stream);

    // the views the kernel reads, pointing into the scratch array and at the caller's tables
    ModelParams mp = m.params(b->tpb, 48);  // the padded records, which every model has (the packed ones would be allocated)
    DevBatch bv{};
    bv.count = count;
    bv.instModel = G(reinterpret_cast<const int32_t*>(base + plan.instModel));
    bv.order = G(reinterpret_cast<const int32_t*>(base + plan.order));
    bv.nodeOff = G(reinterpret_cast<const int32_t*>(base + plan.nodeOff));
    bv.connOff = G(reinterpret_cast<const int32_t*>(base + plan.connOff));
    bv.evOff = G(reinterpret_cast<const int32_t*>(base + plan.evOff));
    bv.monOff = G(reinterpret_cast<const int32_t*>(base + plan.monOff));
    bv.markOff = G(reinterpret_cast<const int32_t*>(base + plan.markOff));
    bv.counter = W(reinterpret_cast<int32_t*>(base + plan.counter));
    bv.en = G(en); bv.area = G(reinterpret_cast<const double*>(base + plan.carea)); bv.tOn = G(tOn);
    bv.st = W(reinterpret_cast<double*>(base + plan.st)); bv.out = W(reinterpret_cast<double*>(base + plan.out));
    bv.vol = W(reinterpret_cast<double*>(base + plan.vol)); bv.level = W(reinterpret_cast<double*>(base + plan.level));
    bv.cx = W(reinterpret_cast<double*>(base + plan.cx)); bv.cy = W(reinterpret_cast<double*>(base + plan.cy));
    bv.cz = W(reinterpret_cast<double*>(base + plan.cz));
    bv.evT = W(reinterpret_cast<double*>(base + plan.evT)); bv.evId = W(reinterpret_cast<int32_t*>(base + plan.evId));
    bv.evKind = W(reinterpret_cast<uint8_t*>(base + plan.evKind)); bv.evCount = W(reinterpret_cast<int32_t*>(base + plan.evCount));
    bv.overT = W(reinterpret_cast<double*>(base + plan.overT)); bv.markT = W(reinterpret_cast<double*>(base + plan.markT));
    bv.foundered = W(reinterpret_cast<uint8_t*>(base + plan.foundered)); bv.founderT = W(reinterpret_cast<double*>(base + plan.founderT));
    bv.steps = W(reinterpret_cast<long long*>(base + plan.steps)); bv.final7 = W(reinterpret_cast<double*>(base + plan.final7));
    bv.calm = W(reinterpret_cast<double*>(base + plan.calm));
    bv.exact = 1;
    bv.stepLimit = nullptr;
    ModelParams* dM = reinterpret_cast<ModelParams*>(base + plan.models);
    DevBatch* dB = reinterpret_cast<DevBatch*>(base + plan.batch);

    // the launch shape: one workgroup per instance when all are resident at once, else the refilling form
    if (lds > 64 * 1024)
      for (bool pers : {false, true})
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(batch_kernel(b->mixed, pers, b->tpb)), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(lds)),
                  "hipFuncSetAttribute");
    int perCU = 0, cus = 0;
    hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, reinterpret_cast<const void*>(batch_kernel(b->mixed, true, b->tpb)), b->tpb, lds),
              "occupancy");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0), "hipDeviceGetAttribute");
    const int grid = std::min(count, std::max(1, perCU) * std::max(1, cus));
    const BatchKernel kernel = batch_kernel(b->mixed, grid < count, b->tpb);

    hipLaunchKernelGGL(noalloc_stage, dim3(count), dim3(256), 0, s, mp, bv, dM, dB, reinterpret_cast<int32_t*>(base + plan.counter), recIn, area,
                       reinterpret_cast<double*>(base + plan.carea));
    hip_check(hipGetLastError(), "ss_batch_step_noalloc: staging launch");
    if (steps > 0) {
      hipLaunchKernelGGL(kernel, dim3(grid), dim3(b->tpb), lds, s, static_cast<const ModelParams*>(dM), static_cast<const DevBatch*>(dB),
                         static_cast<long long>(steps), DBL_MAX, nullptr);
      hip_check(hipGetLastError(), "ss_batch_step_noalloc: batch launch");
    }
    hipLaunchKernelGGL(noalloc_pack, dim3(count), dim3(256), 0, s, static_cast<const DevBatch*>(dB), m.NN, recOut);
    hip_check(hipGetLastError(), "ss_batch_step_noalloc: packing launch");
  });
}

}  // extern "C"
