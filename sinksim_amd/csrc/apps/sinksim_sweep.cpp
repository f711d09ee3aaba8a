// sinksim_sweep: run a sweep file and write the results, or compare two results files exactly. The drop-in for the
// reference's sinksim_sweep (apps/sinksim_sweep.cpp:16-65), with the same options and output lines; the instances run
// on the GPU in the portable or portable32 numerics (the oracle and std numerics are CPU-only and refused).
//   sinksim_sweep <ship.json> <sim.json> <sweep.json> [--numerics portable|portable32] [--threads N] [--results out.json]
//   sinksim_sweep --compare a.results.json b.results.json
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <exception>
#include <string>

#include "../host/batch.hpp"
#include "../host/sweep.hpp"
#include "cli.hpp"

using namespace sinksim;

int main(int argc, char** argv) {
  const cli::Args args = cli::Args::parse(argc, argv, {});
  try {
    if (args.has("compare")) {
      if (args.positional.size() < 1) {
        std::fprintf(stderr, "usage: sinksim_sweep --compare a.results.json b.results.json\n");
        return 2;
      }
      const sweep::Results a = sweep::load_results(args.get("compare", ""));
      const sweep::Results b = sweep::load_results(args.positional[0]);
      const sweep::Diff d = sweep::compare(a, b);
      std::printf("compared %zu values over %zu instances: %zu differ", d.compared, a.instances.size(), d.differing);
      if (d.differing) std::printf("; largest |diff| %.3e; first: %s", d.maxAbs, d.first.c_str());
      std::printf("\nRESULT: %s\n", d.differing == 0 ? "IDENTICAL" : "DIFFERENT");
      return d.differing == 0 ? 0 : 1;
    }
    if (args.positional.size() < 3) {
      std::fprintf(stderr,
                   "usage: sinksim_sweep <ship.json> <sim.json> <sweep.json> [--numerics portable|portable32] [--threads N] [--results out.json]\n"
                   "       sinksim_sweep --compare a.results.json b.results.json\n");
      return 2;
    }
    bool mixed = true;
    if (!args.has("numerics")) {
      std::fprintf(stderr, "sinksim_sweep: give --numerics portable or portable32 (the default oracle numerics run on the CPU only)\n");
      return 2;
    }
    if (!cli::gpu_numerics(args, "sinksim_sweep", "oracle", mixed)) return 2;
    std::string why;
    if (!hip::available(&why)) {
      std::fprintf(stderr, "sinksim_sweep: no usable HIP device (%s)\n", why.c_str());
      return 1;
    }
    const auto ship = io::load_ship(args.positional[0]);
    const CompiledSim sim = io::load_sim(args.positional[1]);
    io::check_pairing(*ship, sim);
    const sweep::Sweep sw = sweep::load_sweep(args.positional[2], sim);
    // stopWhenStable (the oracle's calm rule) runs on the device
    if (sw.instances.empty()) throw std::runtime_error("the sweep has no instances");
    std::printf("sweep of %zu instances on %s / %s, numerics %s, tMax %s, curve every %g s, on the GPU\n", sw.instances.size(), ship->id.c_str(),
                sim.id.c_str(), mixed ? "portable32" : "portable", cli::minutes(sw.tMax).c_str(), sw.every);
    const auto t0 = std::chrono::steady_clock::now();
    hip::BatchOptions opt;
    opt.count = static_cast<Index>(sw.instances.size());
    opt.mixedPrecision = mixed;
    opt.tracedInstances = 0;
    opt.tMax = sw.tMax;
    opt.curveEvery = sw.every;
    hip::Batch batch(ship, sim, opt);
    batch.set_stop_when_stable(sw.stopWhenStable);
    for (Index i = 0; i < opt.count; ++i)
      sweep::apply(sw.instances[static_cast<size_t>(i)], sim, [&](Index c, bool en) { batch.set_connection_enabled(i, c, en); },
                   [&](Index c, Real a) { batch.set_connection_area(i, c, a); }, [&](Index c, Real t) { batch.set_connection_activation(i, c, t); });
    batch.run();
    const hip::DeviceInfo info = hip::device_info();
    sweep::Results r;
    r.ship = ship->id; r.sim = sim.id; r.numerics = mixed ? "portable32" : "portable"; r.engine = "sinksim HIP engine on " + info.name;
    r.tMax = sw.tMax; r.every = sw.every;
    std::int64_t steps = 0;
    int foundered = 0;
    for (Index i = 0; i < opt.count; ++i) {
      const hip::InstanceResult ir = batch.result(i);
      sweep::InstanceResult res;
      res.id = sw.instances[static_cast<size_t>(i)].id;
      res.foundered = ir.foundered; res.founderT = ir.founderT; res.tEnd = ir.tEnd; res.steps = ir.steps;
      const hip::CurveSample f = batch.final_readouts(i);
      res.final.t = f.t; res.final.trimDeg = f.trim; res.final.listDeg = f.list; res.final.waterT = f.water; res.final.inflowTpm = f.inflow;
      res.final.draftF = f.draftF; res.final.draftA = f.draftA; res.final.waterM3 = f.water / sim.rho * 1000;
      for (const hip::Event& e : ir.events) res.events.push_back(sweep::EventRecord{e.t, batch.event_id(e), batch.event_label(e)});
      for (const hip::CurveSample& c : batch.curve(i)) res.curve.push_back(sweep::CurveRecord{c.t, c.trim, c.list, c.water, c.inflow, c.draftF, c.draftA});
      steps += res.steps;
      if (res.foundered) ++foundered;
      r.instances.push_back(std::move(res));
    }
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("%lld steps in %.2f s (%.0f steps/s aggregate); %d of %zu instances foundered\n", static_cast<long long>(steps), s, steps / s, foundered,
                r.instances.size());
    for (size_t i = 0; i < std::min<size_t>(r.instances.size(), 12); ++i) {
      const sweep::InstanceResult& in = r.instances[i];
      std::printf("  %-16s %s  trim %.2f  list %.2f  water %.0f t\n", in.id.c_str(), in.foundered ? ("founders at " + cli::minutes(in.founderT)).c_str() : "afloat",
                  in.final.trimDeg, in.final.listDeg, in.final.waterT);
    }
    if (r.instances.size() > 12) std::printf("  ... %zu more\n", r.instances.size() - 12);
    if (args.has("results")) {
      sweep::save_results(r, args.get("results", ""));
      std::printf("wrote %s\n", args.get("results", "").c_str());
    }
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "sinksim_sweep: %s\n", e.what());
    return 1;
  }
}
