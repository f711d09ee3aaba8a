// The run files of the engine's tools (formats in the reference's docs/FORMATS.md, version 1):
//   sinksim.sweep          instances of one scenario, each a list of kind scales and connection overrides
//   sinksim.batch-results  per instance: the end of its run, its events and its readout curve
//   sinksim.trace          the complete state at every dense and sampled record of one run
// A sweep is turned into whole connection tables (every instance's en, area and tOn rows), the form the batch and the
// CPU engine take; results are compared value for value, as the reference's exact comparison does.
#pragma once

#include <array>
#include <cstdint>
#include <optional>
#include <string>
#include <utility>
#include <vector>

#include "model.hpp"

namespace sinksim::files {

// One override: the connections a selector picks (one connection by index, or every connection of a kind, or of a
// kind and idx) get the fields that are set.
struct Override {
  Index conn = -1;
  Byte kind = 255;
  Index idx = -1;
  std::optional<bool> enabled;
  std::optional<Real> tOn, area;
};

struct Variant {
  std::string id;
  std::vector<std::pair<Byte, Real>> scales;  // kind -> factor on the compiled area
  std::vector<Override> overrides;            // in file order, after the scales
};

struct SweepSpec {
  std::string ship, sim;
  Real tMax = 5 * 3600, every = 30;
  bool stopWhenStable = false;
  std::vector<Variant> instances;
};

SweepSpec read_sweep(const std::string& path, const CompiledSim& sim);

// Every instance's connection table, instance after instance ([count * nconn] each): the scenario's table, each kind
// scale multiplying the compiled area of that kind, then the overrides in file order (a later one wins); the order of
// the reference's sweep::apply (engine/include/sinksim/sweep.hpp:49-65), which bit-exact sweeps depend on.
struct Tables {
  std::vector<Byte> en;
  std::vector<Real> area, tOn;
};
Tables connection_tables(const SweepSpec& sw, const CompiledSim& sim);

// readouts of a curve record: t trim list water(t) inflow(t/min) draftF draftA
using Curve = std::array<Real, 7>;

struct Outcome {
  std::string id;
  bool foundered = false;
  Real founderT = -1, tEnd = 0;
  std::int64_t steps = 0;
  Curve final{};
  struct Event {
    Real t;
    std::string id, label;
  };
  std::vector<Event> events;
  std::vector<Curve> curve;
};

struct ResultsFile {
  std::string ship, sim, numerics, engine;
  Real tMax = 0, every = 0;
  std::vector<Outcome> instances;
};

void write_results(const ResultsFile& r, const std::string& path);
ResultsFile read_results(const std::string& path);

// Exact comparison: per instance the founder flag and time, the end time, the step count, the final trim, list, water
// and drafts, the events (count, times, ids) and the curve (length and every value); NaN equals NaN.
struct Difference {
  size_t compared = 0, differing = 0;
  double largest = 0;  // largest |a - b| over the differing values
  std::string first;   // where the first difference is, with both values
};
Difference compare(const ResultsFile& a, const ResultsFile& b);

// A trace: records in the engine's layout (t zO pitch roll vz wth wph inflow Vb hullTop, then vol level cx cy cz per
// node), dense then sampled.
struct TraceFile {
  std::string role = "trace", ship, sim, title, producer, engine;
  Real dt = 0, sampleEvery = 0, tEnd = 0, founderT = -1;
  Index nodes = 0;
  int denseSteps = 0;
  std::int64_t steps = 0;
  bool foundered = false;
  std::vector<Outcome::Event> events;
  std::vector<Real> dense, samples;  // row-major records of 10 + 5 * nodes values
};
void write_trace(const TraceFile& tr, const std::string& path);

// "%.2f min", the tools' way of printing a time
std::string minutes(double seconds);

}  // namespace sinksim::files
