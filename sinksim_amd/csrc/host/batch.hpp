// sinksim::hip::Batch — the reference's sinksim::cuda::Batch surface (engine/include/sinksim/cuda/batch.hpp:14-87)
// over the C ABI of include/sinksim_hip.h, with the same exception types: std::invalid_argument, std::runtime_error,
// std::logic_error and std::out_of_range. Header-only; link libsinksim_hip.so.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/sinksim_hip.h"
#include "model.hpp"

namespace sinksim::hip {

inline void check(ss_status s) {
  if (s == SS_OK) return;
  const std::string msg = ss_last_error();
  switch (s) {
    case SS_INVALID_ARGUMENT: throw std::invalid_argument(msg);
    case SS_LOGIC: throw std::logic_error(msg);
    case SS_OUT_OF_RANGE: throw std::out_of_range(msg);
    default: throw std::runtime_error(msg);
  }
}

// Is a HIP device present and usable? On failure `why` receives the reason.
inline bool available(std::string* why = nullptr) {
  char buf[512] = {0};
  const bool ok = ss_available(buf, sizeof buf) == SS_OK;
  if (!ok && why) *why = buf;
  return ok;
}

struct DeviceInfo {
  std::string name, arch;
  int computeUnits = 0, major = 0, minor = 0;
  size_t globalMemory = 0, sharedPerBlock = 0;
  int runtimeVersion = 0, driverVersion = 0;
};

inline DeviceInfo device_info() {
  ss_device_properties p{};
  check(ss_device_info(&p));
  return DeviceInfo{p.name, p.arch, p.computeUnits, p.major, p.minor, p.globalMemory, p.sharedPerBlock, p.runtimeVersion, p.driverVersion};
}

struct BatchOptions {
  Index count = 1;
  bool mixedPrecision = true;          // portable32; false computes the portable (double) numerics
  int tracedInstances = 1;
  int denseSteps = 400;
  Real sampleEvery = 60;
  Real curveEvery = 0;
  Real tMax = 5 * 3600;
  std::int64_t stepsPerLaunch = 0;     // <= 0: one launch runs every instance to its end
};

struct CurveSample {
  Real t, trim, list, water, inflow, draftF, draftA;
};

struct Event {
  Real t;
  Index id;
  int kind;  // 0 overflow, 1 mark, 2 founder
};

struct InstanceResult {
  bool foundered = false;
  Real founderT = -1;
  Real tEnd = 0;
  std::int64_t steps = 0;
  Real t = 0, zO = 0, pitch = 0, roll = 0, vz = 0, wth = 0, wph = 0;
  std::vector<Real> vol, level, cx, cy, cz;
  Real outputs[7] = {};  // inflow Vb hullTop mw bx by bz
  std::vector<Event> events;
};

// Per-trace records in the engine's record layout (t zO pitch roll vz wth wph inflow Vb hullTop, vol level cx cy cz).
struct TraceRecords {
  int nodes = 0;
  std::vector<Real> dense, samples;
  int denseCount = 0, sampleCount = 0;
};

class Batch {
 public:
  Batch(std::shared_ptr<const CompiledShip> ship, const CompiledSim& sim, const BatchOptions& opt)
      : ship_(std::move(ship)), sim_(sim), opt_(opt) {
    if (!ship_) throw std::invalid_argument("hip::Batch: null ship");
    const ss_ship s = ship_->view();
    const ss_sim m = sim_.view();
    ss_batch_options o{};
    o.count = opt.count;
    o.numerics = opt.mixedPrecision ? SS_NUMERICS_PORTABLE32 : SS_NUMERICS_PORTABLE;
    o.traced = opt.tracedInstances;
    o.denseSteps = opt.denseSteps;
    o.sampleEvery = opt.sampleEvery;
    o.curveEvery = opt.curveEvery;
    o.tMax = opt.tMax;
    o.stepsPerLaunch = opt.stepsPerLaunch;
    ss_batch* h = nullptr;
    check(ss_batch_create(&s, &m, &o, &h));
    h_.reset(h);
    if (opt_.tracedInstances > opt_.count) opt_.tracedInstances = opt_.count;
  }

  void restore(Index i, const Real* state7, const Real* vol, const Real* level, const Real* cx, const Real* cy, const Real* cz) {
    check(ss_batch_restore(h_.get(), i, state7, vol, level, cx, cy, cz));
  }
  void set_connection_enabled(Index i, Index c, bool en) { check(ss_batch_set_connection_enabled(h_.get(), i, c, en ? 1 : 0)); }
  void set_connection_area(Index i, Index c, Real area) { check(ss_batch_set_connection_area(h_.get(), i, c, area)); }
  void set_connection_activation(Index i, Index c, Real tOn) { check(ss_batch_set_connection_activation(h_.get(), i, c, tOn)); }
  void set_connections(const Byte* en, const Real* area, const Real* tOn) { check(ss_batch_set_connections(h_.get(), en, area, tOn)); }
  // the oracle's stopWhenStable rule on the device (no counterpart in the reference's GPU batch)
  void set_stop_when_stable(bool on) { check(ss_batch_set_stop_when_stable(h_.get(), on ? 1 : 0)); }

  double run() {
    double s = 0;
    check(ss_batch_run(h_.get(), &s));
    return s;
  }

  const BatchOptions& options() const { return opt_; }
  std::int64_t total_steps() const {
    int64_t v = 0;
    check(ss_batch_total_steps(h_.get(), &v));
    return v;
  }
  size_t shared_memory_bytes() const {
    size_t v = 0;
    check(ss_batch_shared_memory_bytes(h_.get(), &v));
    return v;
  }

  InstanceResult result(Index i) const {
    InstanceResult r;
    const size_t nn = static_cast<size_t>(ship_->nodes());
    r.vol.resize(nn); r.level.resize(nn); r.cx.resize(nn); r.cy.resize(nn); r.cz.resize(nn);
    const Index cap = sim_.monitors() + sim_.marks() + 1;
    std::vector<ss_event> ev(static_cast<size_t>(cap));
    ss_instance_result x{};
    check(ss_batch_result(h_.get(), i, &x, r.vol.data(), r.level.data(), r.cx.data(), r.cy.data(), r.cz.data(), ev.data(), cap));
    r.foundered = x.foundered != 0; r.founderT = x.founderT; r.tEnd = x.tEnd; r.steps = x.steps;
    r.t = x.state[0]; r.zO = x.state[1]; r.pitch = x.state[2]; r.roll = x.state[3]; r.vz = x.state[4]; r.wth = x.state[5]; r.wph = x.state[6];
    for (int k = 0; k < 7; ++k) r.outputs[k] = x.outputs[k];
    for (Index k = 0; k < std::min(x.events, cap); ++k) r.events.push_back(Event{ev[static_cast<size_t>(k)].t, ev[static_cast<size_t>(k)].id, ev[static_cast<size_t>(k)].kind});
    return r;
  }

  TraceRecords trace(Index i) const {
    TraceRecords t;
    t.nodes = ship_->nodes();
    int32_t nd = 0, ns = 0;
    check(ss_batch_trace(h_.get(), i, nullptr, 0, &nd, nullptr, 0, &ns));
    const size_t rs = static_cast<size_t>(10 + 5 * t.nodes);
    t.dense.resize(static_cast<size_t>(nd) * rs);
    t.samples.resize(static_cast<size_t>(ns) * rs);
    check(ss_batch_trace(h_.get(), i, t.dense.data(), nd, &nd, t.samples.data(), ns, &ns));
    t.denseCount = nd;
    t.sampleCount = ns;
    return t;
  }

  std::vector<CurveSample> curve(Index i) const {
    int32_t n = 0;
    check(ss_batch_curve(h_.get(), i, nullptr, 0, &n));
    std::vector<Real> buf(static_cast<size_t>(n) * 7);
    check(ss_batch_curve(h_.get(), i, buf.data(), n, &n));
    std::vector<CurveSample> out;
    for (int32_t k = 0; k < n; ++k) {
      const Real* p = buf.data() + static_cast<size_t>(k) * 7;
      out.push_back(CurveSample{p[0], p[1], p[2], p[3], p[4], p[5], p[6]});
    }
    return out;
  }

  CurveSample final_readouts(Index i) const {
    Real r[7];
    check(ss_batch_final_readouts(h_.get(), i, r));
    return CurveSample{r[0], r[1], r[2], r[3], r[4], r[5], r[6]};
  }

  std::string event_id(const Event& e) const { return sinksim::event_id(sim_, e.kind, e.id); }
  std::string event_label(const Event& e) const { return sinksim::event_label(sim_, e.kind, e.id); }

 private:
  struct Free {
    void operator()(ss_batch* b) const { ss_batch_destroy(b); }
  };
  std::shared_ptr<const CompiledShip> ship_;
  CompiledSim sim_;
  BatchOptions opt_;
  std::unique_ptr<ss_batch, Free> h_;
};

}  // namespace sinksim::hip
