// sinksim_sweep: the drop-in for the reference's CPU sweep runner (apps/sinksim_sweep.cpp): run a sweep file's
// instances and write their results, or compare two results files exactly.
//   sinksim_sweep <ship.json> <sim.json> <sweep.json> [--numerics oracle|portable|portable32|std] [--threads N]
//                 [--results out.json] [--device cpu|gpu] [--steps-per-launch K]
//   sinksim_sweep --compare a.results.json b.results.json
// As in the reference the instances run on the host CPU (the numerics default to oracle, --threads 0 takes every
// hardware thread). --device gpu runs the portable and portable32 numerics on the GPU batch instead, with the same
// results bit for bit. Exit codes: 0 done or identical, 1 error or different, 2 usage.
#include <cstdio>
#include <exception>
#include <string>

#include "../host/batch.hpp"
#include "../host/runfiles.hpp"
#include "options.hpp"
#include "runs.hpp"

using namespace sinksim;

namespace {

const char kUsage[] =
    "usage: sinksim_sweep <ship.json> <sim.json> <sweep.json> [--numerics oracle|portable|portable32|std] [--threads N] [--results out.json]\n"
    "                     [--device cpu|gpu]\n"
    "       sinksim_sweep --compare a.results.json b.results.json\n";

int compare_files(const std::string& a, const std::string& b) {
  const files::ResultsFile x = files::read_results(a), y = files::read_results(b);
  const files::Difference d = files::compare(x, y);
  std::printf("compared %zu values over %zu instances: %zu differ", d.compared, x.instances.size(), d.differing);
  if (d.differing) std::printf("; largest |diff| %.3e; first: %s", d.largest, d.first.c_str());
  std::printf("\nRESULT: %s\n", d.differing ? "DIFFERENT" : "IDENTICAL");
  return d.differing ? 1 : 0;
}

}  // namespace

int main(int argc, char** argv) {
  tools::CommandLine cl(argc, argv, {});
  const std::vector<std::string>& pos = cl.positional();
  const bool comparing = cl.has("compare");
  int numerics = SS_NUMERICS_ORACLE;
  const std::string device = cl.text("device", "cpu");
  const double threadsArg = cl.number("threads", 0), chunk = cl.number("steps-per-launch", 0);
  if (!cl.error().empty()) {
    std::fprintf(stderr, "%s\n", cl.error().c_str());
    return 2;
  }
  if (comparing ? pos.empty() : pos.size() < 3) {
    std::fputs(comparing ? "usage: sinksim_sweep --compare a.results.json b.results.json\n" : kUsage, stderr);
    return 2;
  }
  if (!comparing) {
    const std::string name = cl.text("numerics", "oracle");
    if (!tools::numerics_code(name, numerics)) {
      std::fprintf(stderr, "unknown --numerics %s (use oracle, portable, portable32 or std)\n", name.c_str());
      return 2;
    }
    if (device != "cpu" && device != "gpu") {
      std::fprintf(stderr, "unknown --device %s (use cpu or gpu)\n", device.c_str());
      return 2;
    }
    if (device == "gpu" && numerics != SS_NUMERICS_PORTABLE && numerics != SS_NUMERICS_PORTABLE32) {
      std::fprintf(stderr, "--device gpu computes the portable and portable32 numerics only (oracle and std run on the CPU)\n");
      return 2;
    }
  }
  try {
    if (comparing) return compare_files(cl.text("compare", ""), pos[0]);
    const std::shared_ptr<const CompiledShip> ship = io::load_ship(pos[0]);
    const CompiledSim sim = io::load_sim(pos[1]);
    io::check_pairing(*ship, sim);
    const files::SweepSpec sw = files::read_sweep(pos[2], sim);
    const int threads = static_cast<int>(threadsArg);
    std::printf("sweep of %zu instances on %s / %s, numerics %s, tMax %s, curve every %g s, %d threads\n", sw.instances.size(), ship->id.c_str(),
                sim.id.c_str(), tools::numerics_name(numerics), files::minutes(sw.tMax).c_str(), sw.every, threads);
    tools::SweepRun run;
    if (device == "gpu") {
      std::string why;
      if (!hip::available(&why)) {
        std::fprintf(stderr, "sinksim_sweep: no usable HIP device (%s)\n", why.c_str());
        return 1;
      }
      run = tools::gpu_sweep(ship, sim, sw, numerics == SS_NUMERICS_PORTABLE32, static_cast<std::int64_t>(chunk));
      run.results.engine = "sinksim_amd HIP engine on " + hip::device_info().name;
    } else {
      run = tools::cpu_sweep(*ship, sim, sw, numerics, threads);
      run.results.engine = "sinksim_amd CPU engine";
    }
    files::ResultsFile& r = run.results;
    r.ship = ship->id;
    r.sim = sim.id;
    r.numerics = tools::numerics_name(numerics);
    r.tMax = sw.tMax;
    r.every = sw.every;
    std::printf("%lld steps in %.2f s (%.0f steps/s aggregate); %d of %zu instances foundered\n", static_cast<long long>(run.steps), run.seconds,
                run.seconds > 0 ? static_cast<double>(run.steps) / run.seconds : 0.0, tools::foundered_count(r), r.instances.size());
    tools::print_instances(r);
    if (cl.has("results")) {
      files::write_results(r, cl.text("results", ""));
      std::printf("wrote %s\n", cl.text("results", "").c_str());
    }
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "sinksim_sweep: %s\n", e.what());
    return 1;
  }
}
