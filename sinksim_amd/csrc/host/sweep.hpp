// Sweep and results files on the host (C++), and trace files: reference engine/include/sinksim/sweep.hpp:17-108,
// engine/src/sweep.cpp:140-391, engine/src/trace.cpp:189-225 (formats in docs/FORMATS.md).
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "model.hpp"

namespace sinksim::sweep {

struct ConnectionOverride {
  Index conn = -1;
  Byte kind = 255;
  Index idx = -1;
  bool hasEnabled = false, enabled = false;
  bool hasTOn = false;
  Real tOn = 0;
  bool hasArea = false;
  Real area = 0;
};

struct Instance {
  std::string id;
  std::vector<std::pair<Byte, Real>> kindScale;
  std::vector<ConnectionOverride> overrides;
};

struct Sweep {
  std::string ship, sim;
  Real tMax = 5 * 3600;
  Real every = 30;
  bool stopWhenStable = false;
  std::vector<Instance> instances;
};

Sweep load_sweep(const std::string& path, const CompiledSim& sim);
void save_sweep(const Sweep& sw, const std::string& path, const CompiledSim& sim);

// sweep.hpp:49-65: kind scales multiply the compiled area, then the overrides in order, through three setters.
template <class SetEnabled, class SetArea, class SetTOn>
void apply(const Instance& in, const CompiledSim& base, SetEnabled setEnabled, SetArea setArea, SetTOn setTOn) {
  const Index nc = base.connections();
  for (const auto& [kind, factor] : in.kindScale)
    for (Index c = 0; c < nc; ++c)
      if (base.cKind[static_cast<size_t>(c)] == kind) setArea(c, base.cArea[static_cast<size_t>(c)] * factor);
  for (const ConnectionOverride& o : in.overrides) {
    for (Index c = 0; c < nc; ++c) {
      const auto i = static_cast<size_t>(c);
      const bool match = o.conn >= 0 ? c == o.conn : (base.cKind[i] == o.kind && (o.idx < 0 || base.cIdx[i] == o.idx));
      if (!match) continue;
      if (o.hasEnabled) setEnabled(c, o.enabled);
      if (o.hasArea) setArea(c, o.area);
      if (o.hasTOn) setTOn(c, o.tOn);
    }
  }
}

struct Readouts {
  Real t = 0, trimDeg = 0, listDeg = 0, draftF = 0, draftA = 0, waterM3 = 0, waterT = 0, inflowTpm = 0;
};

struct CurveRecord {
  Real t, trim, list, water, inflow, dF, dA;
};

struct EventRecord {
  Real t;
  std::string id, label;
};

struct InstanceResult {
  std::string id;
  bool foundered = false;
  Real founderT = -1;
  Real tEnd = 0;
  std::int64_t steps = 0;
  Readouts final{};
  std::vector<EventRecord> events;
  std::vector<CurveRecord> curve;
};

struct Results {
  std::string ship, sim, numerics, engine;
  Real tMax = 0, every = 0;
  std::vector<InstanceResult> instances;
};

void save_results(const Results& r, const std::string& path);
Results load_results(const std::string& path);

struct Diff {
  size_t differing = 0;
  size_t compared = 0;
  double maxAbs = 0;
  std::string first;
};
Diff compare(const Results& a, const Results& b);

}  // namespace sinksim::sweep

namespace sinksim::io {

// A trace in the golden layout; records are 10 + 5 * nodes values each (t zO pitch roll vz wth wph inflow Vb hullTop,
// then vol level cx cy cz per node).
struct TraceEvent {
  Real t;
  std::string id, label;
};

struct Trace {
  std::string role = "trace", ship, sim, title, producer, engine;
  Real dt = 0;
  Index nodes = 0;
  int denseSteps = 0;
  Real sampleEvery = 0;
  std::int64_t steps = 0;
  Real tEnd = 0;
  bool foundered = false;
  Real founderT = -1;
  std::vector<TraceEvent> events;
  std::vector<Real> dense, samples;  // row-major records
};

void save_trace(const Trace& tr, const std::string& path);

}  // namespace sinksim::io
