/*
 * sinksim_hip.h — the C ABI of the MI355X batch flooding engine (libsinksim_hip.so).
 *
 * One opaque batch handle owns every device buffer of a batch of independent simulations of one compiled ship
 * and scenario. Plain pointers and sizes cross this boundary, never C++ or torch types. Every call returns an
 * ss_status; on failure ss_last_error() holds the message (thread-local, valid until the next call on the
 * same thread). The status classes are the exception types of the reference's C++ API so that the C++
 * wrapper (sinksim_amd/csrc/host/batch.hpp) can rethrow the same type:
 *   SS_INVALID_ARGUMENT  std::invalid_argument   SS_RUNTIME  std::runtime_error
 *   SS_LOGIC             std::logic_error        SS_OUT_OF_RANGE std::out_of_range
 *
 * Each entry point names the reference interface it replaces (paths relative to the reference checkout).
 * The device computes the reference's "portable" (double) and "portable32" (mixed) numerics bit for bit.
 */
#ifndef SINKSIM_HIP_H
#define SINKSIM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ss_status {
  SS_OK = 0,
  SS_INVALID_ARGUMENT = 1,
  SS_RUNTIME = 2,
  SS_LOGIC = 3,
  SS_OUT_OF_RANGE = 4
} ss_status;

/* The reference's numerics, Numerics::{Oracle, Portable, PortableMixed, Std} (engine/include/sinksim/simulation.hpp:26).
 * A batch on the device computes portable and portable32 (BatchOptions::mixedPrecision false / true,
 * engine/include/sinksim/cuda/batch.hpp:29); the host CPU engine (ss_cpu_*) computes all four. */
enum { SS_NUMERICS_ORACLE = 0, SS_NUMERICS_PORTABLE = 1, SS_NUMERICS_PORTABLE32 = 2, SS_NUMERICS_STD = 3 };

/* A compiled ship (docs/FORMATS.md "sinksim.compiled-ship"; engine/include/sinksim/model.hpp:15-35). */
typedef struct ss_ship {
  int32_t ncol;
  const double *x, *y, *dxdy, *zlo, *zhi; /* [ncol] */
  int32_t nseg;
  const int32_t* segCol;                  /* [nseg] */
  const double *segA, *segE;              /* [nseg] */
  int32_t nnode;
  const double *mu, *vmax, *amin, *aov;   /* [nnode] */
  const int32_t *segStart, *segCount;     /* [nnode] */
} ss_ship;

/* A compiled simulation (docs/FORMATS.md "sinksim.compiled-sim"; engine/include/sinksim/model.hpp:49-96). */
typedef struct ss_sim {
  int32_t nnode;     /* must equal the ship's node count (engine/cuda/batch.cu:183) */
  double rho, g, dt;
  int32_t ngroup;
  const int32_t *gNodeStart, *gNodeCount, *gNodes; /* [ngroup], [ngroup], [sum of counts] */
  const uint8_t* gMode;
  const double *gpx, *gpy, *gpz;
  double gMargin;
  int32_t nconn;
  const int32_t *ca, *cb;                          /* a = -1 is the sea */
  const uint8_t* law;                              /* 0 orifice, 1 weir */
  const double *cx, *cy, *cz, *area, *coef;
  const uint8_t* kind;
  const int32_t* idx;
  const uint8_t* en;
  const double* tOn;
  const int32_t *skipGroup, *monitor;
  int32_t nmonitor;
  double monitorThreshold;
  int32_t nmark;
  const double *mkx, *mky, *mkz;
  const uint8_t* mkWhen;
  double founder[4]; /* hullTopBelow, pitchAbsAbove, buoyancyBelow, minT */
  double body[16];   /* ms KG g0x g0y g0z XO YO mh0 Ith0 Iph0 cDz cDth cDph qz qth qph */
  double readout[2]; /* xFP xAP */
  double state0[7];  /* t zO pitch roll vz wth wph */
  const double *vol0, *level0, *cx0, *cy0, *cz0; /* [nnode] */
} ss_sim;

/* BatchOptions (engine/include/sinksim/cuda/batch.hpp:27-37), with the numerics named explicitly. */
typedef struct ss_batch_options {
  int32_t count;          /* simulations in the batch, >= 1 */
  int32_t numerics;       /* SS_NUMERICS_PORTABLE or SS_NUMERICS_PORTABLE32 */
  int32_t traced;         /* leading instances that record traces */
  int32_t denseSteps;     /* dense records at the start of each trace */
  double sampleEvery;     /* s between sampled records */
  double curveEvery;      /* s between readout-curve records of every instance; 0 = none */
  double tMax;            /* s; every instance stops here if it has not foundered */
  int64_t stepsPerLaunch; /* steps per kernel launch (per epoch of a resident run); <= 0: the library's choice */
} ss_batch_options;

/* DeviceInfo (engine/include/sinksim/cuda/batch.hpp:19-25) */
typedef struct ss_device_properties {
  char name[256];
  char arch[64];
  int32_t computeUnits;
  int32_t major, minor;
  size_t globalMemory, sharedPerBlock;
  int32_t runtimeVersion, driverVersion;
} ss_device_properties;

/* InstanceResult (engine/include/sinksim/cuda/batch.hpp:45-54) without the per-node arrays, which
 * ss_batch_result fills separately. */
typedef struct ss_instance_result {
  int32_t foundered;
  double founderT;
  double tEnd;
  int64_t steps;
  double state[7];   /* t zO pitch roll vz wth wph */
  double outputs[7]; /* inflow Vb hullTop mw bx by bz */
  int32_t events;    /* number of events recorded */
} ss_instance_result;

/* Event (engine/include/sinksim/kernel/types.hpp:224-228): kind 0 overflow (id = monitor), 1 mark (id = mark),
 * 2 founder (id = -1). */
typedef struct ss_event {
  double t;
  int32_t id;
  int32_t kind;
} ss_event;

typedef struct ss_batch ss_batch;

/* The message of the last failed call on this thread. */
const char* ss_last_error(void);

/* cuda::available (engine/cuda/batch.cu:71-83): SS_OK when a HIP device is present; otherwise `why` receives
 * the reason (truncated to whyCap bytes). */
ss_status ss_available(char* why, size_t whyCap);

/* cuda::device_info (engine/cuda/batch.cu:85-95), for device 0. */
ss_status ss_device_info(ss_device_properties* out);

/* Batch::Batch (engine/cuda/batch.cu:170-209): uploads the ship and the scenario; every instance starts from
 * the compiled initial state and connection table. SS_INVALID_ARGUMENT on a null ship, count < 1, an unknown
 * numerics value or inconsistent arrays; SS_RUNTIME when no device is present or the model does not fit the
 * kernel's LDS budget. */
ss_status ss_batch_create(const ss_ship* ship, const ss_sim* sim, const ss_batch_options* opt, ss_batch** out);

/* No counterpart in the reference (whose batch runs one hull): a batch of opt->count instances over nmodels
 * models (ships[k], sims[k]); instance i runs model instModel[i] from that model's initial state and connection
 * table. One persistent launch hands the instances out (largest models first) and refills each workgroup as its
 * simulation founders or reaches tMax. Every per-instance entry point below works on a fleet as on a batch;
 * ss_batch_set_connections takes each instance's row of its own model's connection count, concatenated in
 * instance order. Traces need a single model (SS_INVALID_ARGUMENT otherwise). */
ss_status ss_batch_create_fleet(const ss_ship* ships, const ss_sim* sims, int32_t nmodels, const int32_t* instModel,
                                const ss_batch_options* opt, ss_batch** out);

/* Batch::~Batch: frees every device buffer. Null is ignored. */
void ss_batch_destroy(ss_batch* b);

/* Batch::restore (engine/cuda/batch.cu:217-227): state[7] = t zO pitch roll vz wth wph, per-node arrays of
 * nnode values. SS_LOGIC after run, SS_OUT_OF_RANGE on a bad instance. */
ss_status ss_batch_restore(ss_batch* b, int32_t instance, const double* state7, const double* vol, const double* level,
                           const double* cx, const double* cy, const double* cz);

/* Batch::set_connection_enabled / set_connection_area / set_connection_activation (engine/cuda/batch.cu:229-243).
 * SS_LOGIC after run, SS_OUT_OF_RANGE on a bad instance or connection. */
ss_status ss_batch_set_connection_enabled(ss_batch* b, int32_t instance, int32_t conn, int32_t enabled);
ss_status ss_batch_set_connection_area(ss_batch* b, int32_t instance, int32_t conn, double area);
ss_status ss_batch_set_connection_activation(ss_batch* b, int32_t instance, int32_t conn, double tOn);

/* The three setters above for every instance at once: en, area, tOn are [count * nconn] tables (instance-major).
 * Any pointer may be null to keep that field. No reference counterpart: it is the bulk form a driver that
 * applies a whole sweep uses instead of count * nconn single calls. */
ss_status ss_batch_set_connections(ss_batch* b, const uint8_t* en, const double* area, const double* tOn);

/* No counterpart in the reference: the launch shape. threadsPerSim is the workgroup size of one simulation: 256,
 * 512 or 1,024 (beyond 256 the other waves start the level solve while the first 256 threads compute the buoyancy
 * partials), or 0, the default: 256, then 512 and 1,024 for the last epochs of a run once its live instances fit
 * resident at those sizes (see ss_batch_run).
 * workgroups is the number of workgroups of the launch (0, the default: as many as are resident on the device at
 * once, from the kernel's occupancy at this batch's LDS size); fewer than the instance count makes the kernel refill
 * them. The shape changes the time, never the results. SS_LOGIC after run. */
ss_status ss_batch_set_launch(ss_batch* b, int32_t threadsPerSim, int32_t workgroups);

/* No counterpart in the reference: the device layout of the hull's double geometry records (portable numerics).
 * recordBytes 48 (the default) pads each record's five doubles to 48 bytes, 16-byte aligned; 40 packs them, 8-byte
 * aligned. The padded records are read faster while a hull stays in the caches (one hull shared by the batch), the
 * packed ones carry a sixth fewer bytes where hulls stream from memory (fleets of large hulls); ss_batch_autotune
 * measures both. The layout changes the time, never the results. ss_batch_step_noalloc always reads the padded
 * records. SS_INVALID_ARGUMENT for another size; SS_LOGIC after run or step. */
ss_status ss_batch_set_record_layout(ss_batch* b, int32_t recordBytes);

/* The oracle's stopWhenStable rule on the device (RunOptions::stopWhenStable, engine/include/sinksim/simulation.hpp:34;
 * engine/src/simulation.cpp:179-182): past 30 minutes an instance accumulates calm time (dt per step) while the
 * step's sea inflow stays under 0.02, the pitch rate under 1e-5 and the heave velocity under 1e-3 in magnitude,
 * resets it otherwise, and stops once it exceeds 20 minutes. The reference's GPU batch refuses such sweeps
 * (apps/sinksim_cuda_run.cpp:28; its sweep.hpp:41 says CPU only); here every instance applies it when on is 1.
 * SS_LOGIC after run. */
ss_status ss_batch_set_stop_when_stable(ss_batch* b, int32_t on);

/* No counterpart in the reference: what ss_batch_autotune measured and chose. threadsPerSim and workgroups: the shape
 * set on the batch (workgroups 0: one per instance, or as many as are resident at once, the default). rate[k]: the
 * best simulation steps per second measured at 256 (k = 0), 512 (k = 1) and 1,024 (k = 2) threads per simulation (0
 * if that shape does not fit). trial*: every shape measured, in order (at most 8). */
typedef struct ss_tune_result {
  int32_t threadsPerSim, workgroups;
  double rate[3];
  int32_t trials;
  int32_t trialThreads[8], trialWorkgroups[8];
  double trialRate[8];
  int32_t recordBytes;          /* the double geometry records' layout chosen (48 or 40; see ss_batch_set_record_layout) */
  int32_t trialRecordBytes[8];  /* each trial's record layout */
} ss_tune_result;

/* No counterpart in the reference: the autotuner. Runs the first `sample` instances of the work order (the largest
 * models) for trialSteps steps at each launch shape on a trial copy (this batch's staged data is not touched), and
 * sets the fastest shape on the batch. The shapes: 256, 512 and 1,024 threads per simulation, each with the default
 * grid and, where the sample holds more instances than are resident at once (the refilling kernel) and a compute unit
 * holds more than one workgroup, with three quarters and half of the resident workgroups; then, for a portable (FP64)
 * batch, the other layout of the double geometry records at the fastest shape, kept when faster. SS_LOGIC after run. */
ss_status ss_batch_autotune(ss_batch* b, int32_t sample, int64_t trialSteps, ss_tune_result* out);

/* Batch::run (engine/cuda/batch.cu:245-308): runs every instance to its founder or tMax; *seconds receives the
 * wall time of the device work (uploads excluded, as in the reference). When every instance has a resident workgroup
 * of its own, the run is a sequence of epochs (stepsPerLaunch steps each, or by default a 32nd of the longest
 * remaining run, 256 to 4,096 steps): before each epoch the device lists the instances that have not ended, so that
 * they spread evenly over the compute units as others end, and, with the automatic launch shape, runs them at 512
 * threads per simulation once they all fit resident at that size and at 1,024 once they fit one per compute unit.
 * Otherwise one refilling launch (or one per
 * stepsPerLaunch steps). Results are the same bit for bit whatever the launches. SS_LOGIC when called twice. */
ss_status ss_batch_run(ss_batch* b, double* seconds);

/* Layer 1 of the boundary: kernel::step (engine/include/sinksim/kernel/step.hpp:41-54) over the batch, as the
 * reference's Simulation::step (engine/src/simulation.cpp:147-154) calls it after a restore in the acceptance procedure
 * (apps/sinksim_check.cpp:110-126). Every instance takes exactly `steps` steps (perInstance[i] for instance i when
 * perInstance is not null) from its staged or restored state, whether or not it founders or passes tMax, as repeated
 * calls of step() do. The first call uploads the staged instances; later calls continue from the device state. Results,
 * traces and state records read as after ss_batch_run; restore, the setters and ss_batch_run are refused afterwards
 * (SS_LOGIC), and ss_batch_step is refused after ss_batch_run. *seconds (may be null) receives the wall time of the
 * launch. */
ss_status ss_batch_step(ss_batch* b, int64_t steps, const int64_t* perInstance, double* seconds);

/* Every instance's state after ss_batch_run or ss_batch_step as a record of the trace layout (10 + 5 * nnode values:
 * t zO pitch roll vz wth wph inflow Vb hullTop, then vol level cx cy cz per node; the columns of io::Trace,
 * engine/include/sinksim/io/trace.hpp:15-20), instance after instance (a fleet's records have their own model's node
 * count). No reference counterpart: one download in place of a Batch::result call per instance. */
ss_status ss_batch_state_records(const ss_batch* b, double* records);

/* No counterpart in the reference: ss_batch_step for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). It steps `count` instances of the batch's model (a
 * single-model batch; its numerics and launch shape) exactly `steps` steps each, from the device records recIn
 * [count * (10 + 5 * nnode)] (trace layout, as ss_batch_state_records returns them) with the device connection tables
 * en [count * nconn] (0/1), area [count * nconn] and tOn [count * nconn], and writes the states after the steps to
 * recOut (device, same layout; may equal recIn). The batch's own instances and staged tables are not used or changed.
 * Everything else the launch needs lives in one fixed device array of 16 MiB per device (about 5.5 KB per titanic64
 * instance): a count that needs more is SS_INVALID_ARGUMENT. No device or host memory is allocated on success. The
 * work is queued on `stream` (a hipStream_t; null for the default stream) and not waited for: recOut is ready when
 * the stream's work is. Every call on a device shares that device's array, so calls that overlap on one device race
 * (the caller serialises them, e.g. on one stream). */
ss_status ss_batch_step_noalloc(const ss_batch* b, int32_t count, int64_t steps, const double* recIn, const uint8_t* en, const double* area,
                                const double* tOn, double* recOut, void* stream);

/* Batch::total_steps / shared_memory_bytes (engine/include/sinksim/cuda/batch.hpp:73-74). The LDS size is the forward
 * kernel's per simulation at the batch's launch shape, known from creation on. */
ss_status ss_batch_total_steps(const ss_batch* b, int64_t* steps);
ss_status ss_batch_shared_memory_bytes(const ss_batch* b, size_t* bytes);

/* Batch::result (engine/cuda/batch.cu:310-331). Any of the per-node pointers (nnode values each) and `events`
 * (evCap entries) may be null. SS_LOGIC before run, SS_OUT_OF_RANGE on a bad instance. */
ss_status ss_batch_result(const ss_batch* b, int32_t instance, ss_instance_result* res, double* vol, double* level,
                          double* cx, double* cy, double* cz, ss_event* events, int32_t evCap);

/* Every instance's summary at once ([count] arrays; any may be null): the download the reference repeats per
 * instance (engine/cuda/batch.cu:310-331), done once. */
ss_status ss_batch_results(const ss_batch* b, int64_t* steps, int32_t* foundered, double* founderT, double* tEnd);

/* Batch::trace (engine/cuda/batch.cu:333-367): the records of a traced instance, each 10 + 5 * nnode values
 * (t zO pitch roll vz wth wph inflow Vb hullTop, then vol level cx cy cz per node). Call with null buffers to
 * learn the counts. SS_OUT_OF_RANGE when the instance was not traced. */
ss_status ss_batch_trace(const ss_batch* b, int32_t instance, double* dense, int32_t denseCap, int32_t* denseCount,
                         double* samples, int32_t samplesCap, int32_t* samplesCount);

/* Batch::curve (engine/cuda/batch.cu:384-400): the readout curve (7 values per record: t trim list water
 * inflow draftF draftA), closed by the final record. Call with a null buffer to learn the count. SS_LOGIC when
 * curves were not recorded. */
ss_status ss_batch_curve(const ss_batch* b, int32_t instance, double* rec, int32_t cap, int32_t* count);

/* Batch::final_readouts (engine/cuda/batch.cu:369-382): the readouts at the end of the run (7 values as in a
 * curve record), computed on the device with the portable numerics. */
ss_status ss_batch_final_readouts(const ss_batch* b, int32_t instance, double* rec7);

/* Device time of run(): the summed HIP-event duration of its kernel launches (ms) and the number of launches (of
 * epochs, for a run in epochs; each epoch's graph holds the batch kernel at both workgroup sizes and the next plan).
 * No reference counterpart: the measurement hook of the benchmark. */
ss_status ss_batch_kernel_time(const ss_batch* b, double* ms, int32_t* launches);


/* ---------------------------------------------------------------------------------------------------------------
 * The gradient (no counterpart in the reference; the reference calibrates by batched search over sweeps,
 * tools/py/calibrate_gpu.py:95-124, which this replaces with gradients).
 *
 * Parameters theta are positive scale factors: connection c of instance i takes area * theta[i][pa[c]] and discharge
 * coefficient coef * theta[i][pc[c]] (-1: no factor); pa and pc are laid out like the connection tables.
 *
 * The loss of an instance is a weighted sum of squared errors against observations (extending the reference's
 * calibration objective, tools/py/calibrate_gpu.py:35-64, to terms with derivatives):
 *   readout terms j: the trim (rdKind 0: pitch * 180 / pi) or list (rdKind 1: roll * 180 / pi) after step rdStep[j];
 *                    rdAfter[j] when the run ended before that step (the reference's values after foundering);
 *   event terms e:   the time at which monitor evIdx (evKind 0) first exceeds the threshold, mark evIdx (evKind 1)
 *                    first goes under (or above), or the ship founders (evKind 2), interpolated linearly between the
 *                    step before and the step at which it fires from the event's signed signal s (monitor: flow sum
 *                    - threshold; mark: -world z, or +world z for an 'above' mark; founder: the largest margin of the
 *                    founder rule): t* = t_k - dt * s_k / (s_k - s_{k-1}). An event that never fires scores the founder
 *                    time if the ship foundered, else evMissing[e]. */
typedef struct ss_loss {
  int32_t nRead;
  const int64_t* rdStep;
  const int32_t* rdKind;
  const double* rdW;
  const double* rdAfter;
  int32_t nEvent;
  const int32_t* evKind;
  const int32_t* evIdx;
  const double* evW;
  const double* evMissing;
} ss_loss;

/* checkpointEvery: K, the steps between stored states (the reverse sweep recomputes each segment of K steps from its
 * checkpoint; K = 1 stores every state and recomputes nothing; 0 chooses 1 when every instance's states fit in a
 * third of the free device memory, else 256); relaxed: 0 for the derivative of the discrete program, 1 for relaxed level derivatives (each level's
 * tangent taken at the exact root of its volume, as if the level solve had converged); tMax (0: the batch's) and
 * maxSteps (<= 0: none) end the run. */
typedef struct ss_gradient_options {
  int32_t checkpointEvery;
  int32_t relaxed;
  double tMax;
  int64_t maxSteps;
} ss_gradient_options;

/* Runs every instance from its staged initial state and returns its loss [count], the loss's gradient with respect to
 * theta [count * nparams] (null: forward only, no reverse sweep), the loss's signals [count * (nRead + nEvent)]: readout
 * values then event times (may be null), the steps run [count], the recomputed steps whose values differed from the
 * forward's [count] (always 0 for a correct kernel; may be null) and the kernel time (ms, may be null). The forward
 * runs in the batch's numerics: a portable batch is FP64 throughout; a portable32 batch runs the forward and its
 * recomputations with the portable32 step (the batch kernel's trajectory bit for bit) and the reverse in FP64, with
 * the local Jacobians of the FP64 step at the portable32 states; its mismatch count checks each recomputed segment's
 * last state against the forward's. The launch takes 512 threads per simulation where the forward layout and the whole tape
 * fit in LDS, else 256, else the same two with the tape's Jacobian rows and per-connection partials in global memory
 * (ss_batch_gradient_launch reports which). The batch may be run any number of times this way; its own run() is
 * unaffected. */
ss_status ss_batch_gradient(ss_batch* b, int32_t nparams, const int32_t* pa, const int32_t* pc, const double* theta,
                            const ss_loss* loss, const double* observations, const ss_gradient_options* opt, double* lossOut,
                            double* grad, double* signals, int64_t* steps, int32_t* mismatches, double* kernelMs);

/* One step of instance 0 in reverse, the test hook of the adjoint: from a state record (the trace layout), the
 * adjoint of the state after the step (same layout; entries 8 and 9 seed the step's buoyancy volume and hull top,
 * entries 0 and 7 are ignored) and seeds on the step's monitor flow
 * sums [nmonitor] and marks' world z [nmark] (either may be null), the adjoint of the state before it (lambdaOut,
 * trace layout), the loss's derivative with respect to log theta (gradLog [nparams]) and the recomputed state after
 * the step (recordOut, trace layout; inflow and hull top 0). FP64 (portable) batches only. */
ss_status ss_batch_adjoint_step(ss_batch* b, int32_t nparams, const int32_t* pa, const int32_t* pc, const double* theta, int32_t relaxed,
                                const double* record, const double* lambda, const double* monitorSeed, const double* markSeed,
                                double* lambdaOut, double* gradLog, double* recordOut);

/* The launch shape of the last ss_batch_gradient or ss_batch_adjoint_step call: threads per simulation, 1 when the
 * tape's Jacobian rows and per-connection partials were in global memory (else 0), LDS bytes per simulation. SS_LOGIC
 * before any such call. No reference counterpart: a measurement and test hook. */
ss_status ss_batch_gradient_launch(const ss_batch* b, int32_t* threads, int32_t* globalTape, int64_t* ldsBytes);

/* ---------------------------------------------------------------------------------------------------------------
 * The host CPU engine: the reference's four numerics on the CPU (the oracle's serial order with V8-semantics pow, the
 * GPU's canonical tree order in double and in single precision, and the platform math), for the CPU tools and for
 * callers without a device. No device is needed or touched. */

/* RunOptions (engine/include/sinksim/simulation.hpp:31-35) with the numerics and the thread count of sweep::run_cpu
 * (engine/include/sinksim/sweep.hpp:95-96). */
typedef struct ss_cpu_options {
  int32_t numerics;       /* SS_NUMERICS_ORACLE, _PORTABLE, _PORTABLE32 or _STD */
  int32_t threads;        /* host threads; <= 0: as many as the hardware runs at once */
  double tMax;            /* s: each instance runs until it founders or reaches tMax */
  double curveEvery;      /* s between readout-curve records (RunOptions::every) */
  int32_t stopWhenStable; /* 1: stop an instance after 20 min of calm past 30 min (engine/src/simulation.cpp:179-182) */
} ss_cpu_options;

/* sweep::run_cpu (engine/src/sweep.cpp:121-161): count instances of one ship and scenario, each running
 * Simulation::run (engine/src/simulation.cpp:167-192) from the compiled initial state with its own connection tables
 * en, area, tOn ([count * nconn] each, instance-major; null: the scenario's for every instance), on opt->threads
 * host threads. Per instance: results[i] (may be null), up to evCap events at events + i * evCap (may be null), up to
 * curveCap curve records of 7 values (t trim list water inflow draftF draftA, closed by the final record, as the
 * reference's run history) at curves + i * curveCap * 7 with their full count in curveCount[i], and the final
 * readouts final7 + i * 7 (each may be null). *seconds receives the wall time of the runs (no file IO). */
ss_status ss_cpu_sweep(const ss_ship* ship, const ss_sim* sim, const ss_cpu_options* opt, int32_t count, const uint8_t* en,
                       const double* area, const double* tOn, ss_instance_result* results, ss_event* events, int32_t evCap,
                       double* curves, int32_t curveCap, int32_t* curveCount, double* final7, double* seconds);

/* Simulation::record_trace (engine/src/simulation.cpp:194-221) on the host: one instance (connection tables of nconn
 * values, null: the scenario's) from the compiled initial state to its founder or tMax; the initial record, a dense
 * record after each of the first denseSteps steps and a sampled record whenever t reaches the next multiple of
 * sampleEvery, each 10 + 5 * nnode values in the trace layout (at most denseCap / samplesCap kept; the counts are the
 * full ones), the result and up to evCap events. */
ss_status ss_cpu_trace(const ss_ship* ship, const ss_sim* sim, int32_t numerics, double tMax, int32_t denseSteps, double sampleEvery,
                       const uint8_t* en, const double* area, const double* tOn, double* dense, int32_t denseCap, int32_t* denseCount,
                       double* samples, int32_t samplesCap, int32_t* samplesCount, ss_instance_result* result, ss_event* events,
                       int32_t evCap);

#ifdef __cplusplus
}
#endif

#endif /* SINKSIM_HIP_H */
