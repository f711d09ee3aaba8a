// sinksim_hip_run: the drop-in for the reference's GPU runner (apps/sinksim_cuda_run.cpp): a batch of simulations on
// the GPU with its throughput, optionally the trace of the first instance, or a sweep file run to a results file.
//   sinksim_hip_run <ship.json> <sim.json> [--count N] [--precision mixed|double | --numerics portable32|portable]
//                   [--tmax S] [--trace out.json] [--dense N] [--sample-every S] [--steps-per-launch K] [--perturb]
//   sinksim_hip_run <ship.json> <sim.json> --sweep sweep.json --results out.json [--precision mixed|double]
// The options, printed lines and files are the reference's. --perturb makes the batch a sweep of the damage size: the
// breaches of instance i (i >= 1) open to (1 + 0.02 i / count) times their compiled area. stepsPerLaunch defaults to 0
// (one launch runs every instance to its end); sweeps that set stopWhenStable run with the rule on the device.
#include <algorithm>
#include <cstdio>
#include <exception>
#include <string>
#include <vector>

#include "../host/batch.hpp"
#include "../host/runfiles.hpp"
#include "options.hpp"
#include "runs.hpp"

using namespace sinksim;

namespace {

const char kUsage[] =
    "usage: sinksim_hip_run <ship.json> <sim.json> [--count N] [--precision mixed|double] [--tmax S] [--trace out.json] [--dense N]\n"
    "                       [--sample-every S] [--steps-per-launch K] [--perturb]\n"
    "       sinksim_hip_run <ship.json> <sim.json> --sweep sweep.json --results out.json [--precision mixed|double]\n";

struct Settings {
  bool mixed = true;
  int count = 1, dense = 400;
  double tMax = 5 * 3600, sampleEvery = 60;
  std::int64_t stepsPerLaunch = 0;
};

int sweep_mode(const tools::CommandLine& cl, std::shared_ptr<const CompiledShip> ship, const CompiledSim& sim, const Settings& set,
               const hip::DeviceInfo& dev) {
  const files::SweepSpec sw = files::read_sweep(cl.text("sweep", ""), sim);
  if (sw.instances.empty()) throw std::runtime_error("the sweep has no instances");
  const char* numerics = set.mixed ? "portable32" : "portable";
  std::printf("sweep of %zu instances on %s / %s, numerics %s, tMax %s, curve every %g s\n", sw.instances.size(), ship->id.c_str(), sim.id.c_str(),
              numerics, files::minutes(sw.tMax).c_str(), sw.every);
  tools::SweepRun run = tools::gpu_sweep(ship, sim, sw, set.mixed, set.stepsPerLaunch);
  std::printf("%lld steps in %.2f s (%.0f steps/s aggregate)\n", static_cast<long long>(run.steps), run.seconds,
              static_cast<double>(run.steps) / run.seconds);
  files::ResultsFile& r = run.results;
  r.ship = ship->id;
  r.sim = sim.id;
  r.numerics = numerics;
  r.engine = "sinksim_amd HIP engine on " + dev.name;
  r.tMax = sw.tMax;
  r.every = sw.every;
  std::printf("%d of %zu instances foundered\n", tools::foundered_count(r), r.instances.size());
  tools::print_instances(r);
  if (cl.has("results")) {
    files::write_results(r, cl.text("results", ""));
    std::printf("wrote %s\n", cl.text("results", "").c_str());
  }
  return 0;
}

int batch_mode(const tools::CommandLine& cl, std::shared_ptr<const CompiledShip> ship, const CompiledSim& sim, const Settings& set,
               const hip::DeviceInfo& dev) {
  hip::BatchOptions opt;
  opt.count = set.count;
  opt.mixedPrecision = set.mixed;
  opt.tMax = set.tMax;
  opt.tracedInstances = cl.has("trace") ? 1 : 0;
  opt.denseSteps = set.dense;
  opt.sampleEvery = set.sampleEvery;
  opt.stepsPerLaunch = set.stepsPerLaunch;
  std::printf("ship %s: %d columns, %d segments, %d nodes; sim %s: %d groups, %d connections; batch of %d\n", ship->id.c_str(), ship->columns(),
              ship->segments(), ship->nodes(), sim.id.c_str(), sim.groups(), sim.connections(), opt.count);
  hip::Batch batch(ship, sim, opt);
  std::printf("numerics %s; shared memory per block %zu bytes, %lld steps per launch\n",
              set.mixed ? "portable32 (mixed precision)" : "portable (double precision)", batch.shared_memory_bytes(),
              static_cast<long long>(opt.stepsPerLaunch));
  if (cl.has("perturb")) {
    const Index breach = sim.kindIndex("breach");
    const size_t nc = static_cast<size_t>(sim.connections());
    std::vector<Real> area;
    area.reserve(static_cast<size_t>(opt.count) * nc);
    for (Index i = 0; i < opt.count; ++i) {
      const Real factor = 1 + 0.02 * static_cast<Real>(i) / static_cast<Real>(opt.count);
      for (size_t c = 0; c < nc; ++c)
        area.push_back(i > 0 && breach >= 0 && sim.cKind[c] == static_cast<Byte>(breach) ? sim.cArea[c] * factor : sim.cArea[c]);
    }
    batch.set_connections(nullptr, area.data(), nullptr);
  }
  const double seconds = batch.run();
  const std::int64_t steps = batch.total_steps();
  const hip::InstanceResult first = batch.result(0);
  const double rate = static_cast<double>(steps) / seconds;
  std::printf("%lld steps over %d instances in %.3f s: %.0f steps/s aggregate, %.1f simulated hours per wall second, %.0f full sinkings per minute at this length\n",
              static_cast<long long>(steps), opt.count, seconds, rate, static_cast<double>(steps) * sim.dt / 3600.0 / seconds,
              rate / (first.steps > 0 ? static_cast<double>(first.steps) : 1.0) * 60.0);
  int foundered = 0;
  Real earliest = 0, latest = 0;
  for (Index i = 0; i < opt.count; ++i) {
    const hip::InstanceResult r = i == 0 ? first : batch.result(i);
    if (!r.foundered) continue;
    earliest = foundered ? std::min(earliest, r.founderT) : r.founderT;
    latest = foundered ? std::max(latest, r.founderT) : r.founderT;
    ++foundered;
  }
  if (foundered)
    std::printf("%d of %d instances foundered, between %s and %s\n", foundered, opt.count, files::minutes(earliest).c_str(), files::minutes(latest).c_str());
  else
    std::printf("no instance foundered within %s\n", files::minutes(opt.tMax).c_str());
  for (const hip::Event& e : first.events) std::printf("  %9.2f min  %s\n", e.t / 60, batch.event_label(e).c_str());
  if (cl.has("trace")) {
    const std::string path = cl.text("trace", "hip.trace.json");
    const hip::TraceRecords rec = batch.trace(0);
    files::TraceFile tr;
    tr.ship = ship->id;
    tr.sim = sim.id;
    tr.title = sim.title;
    tr.producer = std::string("sinksim_amd HIP engine, numerics ") + (set.mixed ? "portable32" : "portable");
    tr.engine = dev.name + " (" + dev.arch + ")";
    tr.dt = sim.dt;
    tr.nodes = ship->nodes();
    tr.denseSteps = set.dense;
    tr.sampleEvery = set.sampleEvery;
    tr.steps = first.steps;
    tr.tEnd = first.tEnd;
    tr.foundered = first.foundered;
    tr.founderT = first.founderT;
    for (const hip::Event& e : first.events) tr.events.push_back({e.t, batch.event_id(e), batch.event_label(e)});
    tr.dense = rec.dense;
    tr.samples = rec.samples;
    files::write_trace(tr, path);
    std::printf("wrote %s with %d dense and %d sampled records\n", path.c_str(), rec.denseCount, rec.sampleCount);
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  tools::CommandLine cl(argc, argv, {"perturb"});
  Settings set;
  set.count = static_cast<int>(cl.number("count", 1));
  set.tMax = cl.number("tmax", 5 * 3600);
  set.dense = static_cast<int>(cl.number("dense", 400));
  set.sampleEvery = cl.number("sample-every", 60);
  set.stepsPerLaunch = static_cast<std::int64_t>(cl.number("steps-per-launch", 0));
  if (!cl.error().empty()) {
    std::fprintf(stderr, "%s\n", cl.error().c_str());
    return 2;
  }
  if (cl.positional().size() < 2) {
    std::fputs(kUsage, stderr);
    return 2;
  }
  if (!tools::device_numerics(cl, "sinksim_hip_run", set.mixed)) return 2;
  try {
    std::string why;
    if (!hip::available(&why)) {
      std::fprintf(stderr, "sinksim_hip_run: no usable HIP device (%s)\n", why.c_str());
      return 1;
    }
    const hip::DeviceInfo dev = hip::device_info();
    const std::shared_ptr<const CompiledShip> ship = io::load_ship(cl.positional()[0]);
    const CompiledSim sim = io::load_sim(cl.positional()[1]);
    io::check_pairing(*ship, sim);
    std::printf("device %s (%s, %d compute units, %.1f GiB); runtime %d, driver %d\n", dev.name.c_str(), dev.arch.c_str(), dev.computeUnits,
                static_cast<double>(dev.globalMemory) / 1073741824.0, dev.runtimeVersion, dev.driverVersion);
    return cl.has("sweep") ? sweep_mode(cl, ship, sim, set, dev) : batch_mode(cl, ship, sim, set, dev);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "sinksim_hip_run: %s\n", e.what());
    return 1;
  }
}
