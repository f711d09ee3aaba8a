// Sweeps run by the tools, on the host CPU engine or on the GPU batch, gathered into a results file; and the lines the
// tools print about them.
#pragma once

#include <algorithm>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/sinksim_hip.h"
#include "../host/batch.hpp"
#include "../host/runfiles.hpp"

namespace sinksim::tools {

struct SweepRun {
  files::ResultsFile results;
  double seconds = 0;        // wall time of the runs, no file IO
  std::int64_t steps = 0;    // summed over the instances
};

inline std::vector<files::Outcome::Event> named_events(const CompiledSim& sim, const ss_event* ev, int n) {
  std::vector<files::Outcome::Event> out;
  for (int k = 0; k < n; ++k) out.push_back({ev[k].t, event_id(sim, ev[k].kind, ev[k].id), event_label(sim, ev[k].kind, ev[k].id)});
  return out;
}

// Every instance on the host CPU engine in any of the four numerics, `threads` at a time (0: all the hardware's).
inline SweepRun cpu_sweep(const CompiledShip& ship, const CompiledSim& sim, const files::SweepSpec& sw, int numerics, int threads) {
  const files::Tables t = files::connection_tables(sw, sim);
  const int32_t count = static_cast<int32_t>(sw.instances.size());
  const int32_t evCap = sim.monitors() + sim.marks() + 1;
  const int32_t curveCap = static_cast<int32_t>(sw.tMax / sw.every) + 3;
  std::vector<ss_instance_result> res(static_cast<size_t>(count));
  std::vector<ss_event> ev(static_cast<size_t>(count) * static_cast<size_t>(evCap));
  std::vector<double> curves(static_cast<size_t>(count) * static_cast<size_t>(curveCap) * 7), final7(static_cast<size_t>(count) * 7);
  std::vector<int32_t> ncurve(static_cast<size_t>(count));
  const ss_ship s = ship.view();
  const ss_sim m = sim.view();
  const ss_cpu_options o{numerics, threads, sw.tMax, sw.every, sw.stopWhenStable ? 1 : 0};
  SweepRun run;
  hip::check(ss_cpu_sweep(&s, &m, &o, count, t.en.data(), t.area.data(), t.tOn.data(), res.data(), ev.data(), evCap, curves.data(), curveCap,
                          ncurve.data(), final7.data(), &run.seconds));
  for (int32_t i = 0; i < count; ++i) {
    const size_t k = static_cast<size_t>(i);
    files::Outcome o2;
    o2.id = sw.instances[k].id;
    o2.foundered = res[k].foundered != 0;
    o2.founderT = res[k].founderT;
    o2.tEnd = res[k].tEnd;
    o2.steps = res[k].steps;
    std::copy(final7.begin() + static_cast<std::ptrdiff_t>(7 * k), final7.begin() + static_cast<std::ptrdiff_t>(7 * k + 7), o2.final.begin());
    o2.events = named_events(sim, ev.data() + k * static_cast<size_t>(evCap), std::min(res[k].events, evCap));
    const int32_t n = std::min(ncurve[k], curveCap);
    for (int32_t r = 0; r < n; ++r) {
      const double* p = curves.data() + (k * static_cast<size_t>(curveCap) + static_cast<size_t>(r)) * 7;
      o2.curve.push_back({p[0], p[1], p[2], p[3], p[4], p[5], p[6]});
    }
    run.steps += res[k].steps;
    run.results.instances.push_back(std::move(o2));
  }
  return run;
}

// Every instance on the GPU batch (portable or portable32), the stopWhenStable rule on the device when the sweep asks.
inline SweepRun gpu_sweep(std::shared_ptr<const CompiledShip> ship, const CompiledSim& sim, const files::SweepSpec& sw, bool mixed,
                          std::int64_t stepsPerLaunch) {
  hip::BatchOptions opt;
  opt.count = static_cast<Index>(sw.instances.size());
  opt.mixedPrecision = mixed;
  opt.tracedInstances = 0;
  opt.tMax = sw.tMax;
  opt.curveEvery = sw.every;
  opt.stepsPerLaunch = stepsPerLaunch;
  hip::Batch batch(std::move(ship), sim, opt);
  const files::Tables t = files::connection_tables(sw, sim);
  batch.set_connections(t.en.data(), t.area.data(), t.tOn.data());
  if (sw.stopWhenStable) batch.set_stop_when_stable(true);
  SweepRun run;
  run.seconds = batch.run();
  run.steps = batch.total_steps();
  for (Index i = 0; i < opt.count; ++i) {
    const hip::InstanceResult r = batch.result(i);
    files::Outcome o;
    o.id = sw.instances[static_cast<size_t>(i)].id;
    o.foundered = r.foundered;
    o.founderT = r.founderT;
    o.tEnd = r.tEnd;
    o.steps = r.steps;
    const hip::CurveSample f = batch.final_readouts(i);
    o.final = {f.t, f.trim, f.list, f.water, f.inflow, f.draftF, f.draftA};
    for (const hip::Event& e : r.events) o.events.push_back({e.t, batch.event_id(e), batch.event_label(e)});
    for (const hip::CurveSample& c : batch.curve(i)) o.curve.push_back({c.t, c.trim, c.list, c.water, c.inflow, c.draftF, c.draftA});
    run.results.instances.push_back(std::move(o));
  }
  return run;
}

// The lines both tools print after a sweep: the foundered count is the caller's; then up to twelve instances.
inline void print_instances(const files::ResultsFile& r) {
  const size_t shown = std::min<size_t>(r.instances.size(), 12);
  for (size_t i = 0; i < shown; ++i) {
    const files::Outcome& o = r.instances[i];
    const std::string fate = o.foundered ? "founders at " + files::minutes(o.founderT) : "afloat";
    std::printf("  %-16s %s  trim %.2f  list %.2f  water %.0f t\n", o.id.c_str(), fate.c_str(), o.final[1], o.final[2], o.final[3]);
  }
  if (r.instances.size() > shown) std::printf("  ... %zu more\n", r.instances.size() - shown);
}

inline int foundered_count(const files::ResultsFile& r) {
  return static_cast<int>(std::count_if(r.instances.begin(), r.instances.end(), [](const files::Outcome& o) { return o.foundered; }));
}

}  // namespace sinksim::tools
