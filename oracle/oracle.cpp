// ORACLE — test infrastructure only (tests/, smoke() and the CPU-baseline leg of bench.py load it as the
// checker; the product never links it).
//
// The host side of the CPU restatement: a simulation holder with the reference's Simulation semantics
// (engine/src/simulation.cpp:34-221: reset, restore, step, run with the 30 s history, record_trace) and the
// threaded sweep runner of engine/src/sweep.cpp:121-161, exported as a C API for ctypes; and the driver of the
// dual-number build of the portable step (dual_step.hpp): a run with its loss and one step with its tangent.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

#include "dual_step.hpp"
#include "oracle_step.hpp"

using namespace orc;

extern "C" {

// Compiled ship arrays (docs/FORMATS.md "sinksim.compiled-ship").
struct orc_ship {
  int32_t ncol;
  const double *x, *y, *dxdy, *zlo, *zhi;
  int32_t nseg;
  const int32_t* segCol;
  const double *segA, *segE;
  int32_t nn;
  const double *mu, *vmax, *amin, *aov;
  const int32_t *segStart, *segCount;
};

// Compiled simulation arrays (docs/FORMATS.md "sinksim.compiled-sim").
struct orc_sim {
  double rho, g, dt;
  int32_t ng;
  const int32_t *gStart, *gCount, *gNodes;
  const uint8_t* gMode;
  const double *gpx, *gpy, *gpz;
  double gMargin;
  int32_t nc;
  const int32_t *ca, *cb;
  const uint8_t* law;
  const double *cx, *cy, *cz, *area, *coef;
  const uint8_t* en;
  const double* tOn;
  const int32_t *skipGroup, *monitor;
  int32_t nmon;
  double monThreshold;
  int32_t nmark;
  const double *mkx, *mky, *mkz;
  const uint8_t* mkWhen;
  double founder[4];  // hullTopBelow, pitchAbsAbove, buoyancyBelow, minT
  double body[16];    // ms KG g0x g0y g0z XO YO mh0 Ith0 Iph0 cDz cDth cDph qz qth qph
  double readout[2];  // xFP xAP
  double state0[7];   // t zO pitch roll vz wth wph
  const double *vol0, *level0, *cx0, *cy0, *cz0;
};

}  // extern "C"

namespace {

enum Numerics { kOracle = 0, kPortable = 1, kPortable32 = 2, kStd = 3 };

class Sim {
 public:
  Sim(const orc_ship* s, const orc_sim* m, int numerics) : numerics_(numerics) {
    ship_.ncol = s->ncol; ship_.x = s->x; ship_.y = s->y; ship_.dxdy = s->dxdy; ship_.zlo = s->zlo; ship_.zhi = s->zhi;
    ship_.nseg = s->nseg; ship_.segCol = s->segCol; ship_.segA = s->segA; ship_.segE = s->segE;
    ship_.nn = s->nn; ship_.mu = s->mu; ship_.vmax = s->vmax; ship_.amin = s->amin; ship_.aov = s->aov;
    ship_.segStart = s->segStart; ship_.segCount = s->segCount;
    sim_ = *m;
    const Index NN = s->nn, NC = m->nc;
    en_.assign(m->en, m->en + NC);
    area_.assign(m->area, m->area + NC);
    tOn_.assign(m->tOn, m->tOn + NC);
    build_incidence();
    vol_.assign(NN, 0); level_.assign(NN, 0);
    sArea_.assign(NN, 0); cx_.assign(NN, 0); cy_.assign(NN, 0); cz_.assign(NN, 0); zmin_.assign(NN, 0); zmax_.assign(NN, 0);
    hEff_.assign(NN, 0); aEff_.assign(NN, 0); acc_.assign(NN, 0); excess_.assign(NN, 0); deg_.assign(NN, 0);
    merged_.assign(m->ng, 0);
    Index maxGroup = 1;
    for (Index g = 0; g < m->ng; ++g) maxGroup = std::max(maxGroup, m->gCount[g]);
    passes_.assign(maxGroup, Pass{});
    over_.assign(std::max<Index>(m->nmon, 1), 0);
    q_.assign(NC, 0); dV_.assign(NC, 0);
    evItems_.assign(m->nmon + m->nmark + 1, Event{});
    overT_.assign(std::max<Index>(m->nmon, 1), -1); markT_.assign(std::max<Index>(m->nmark, 1), -1);
    bind();
    reset_to_initial();
  }

  void set_conn(Index c, int en, double area, double tOn) { en_[c] = en ? 1 : 0; area_[c] = area; tOn_[c] = tOn; }

  void reset_to_initial() {  // simulation.cpp:83-90
    restore(sim_.state0, sim_.vol0, sim_.level0, sim_.cx0, sim_.cy0, sim_.cz0);
    std::fill(q_.begin(), q_.end(), 0.0);
    std::fill(dV_.begin(), dV_.end(), 0.0);
    std::fill(merged_.begin(), merged_.end(), 0);
    std::fill(sArea_.begin(), sArea_.end(), 0.0);
    evaluate_hydro();
  }

  void restore(const double* s7, const double* vol, const double* level, const double* cx, const double* cy, const double* cz) {
    st_.t = s7[0]; st_.zO = s7[1]; st_.pitch = s7[2]; st_.roll = s7[3]; st_.vz = s7[4]; st_.wth = s7[5]; st_.wph = s7[6];
    for (Index n = 0; n < ship_.nn; ++n) { vol_[n] = vol[n]; level_[n] = level[n]; cx_[n] = cx[n]; cy_[n] = cy[n]; cz_[n] = cz[n]; }
    ev_.count = 0; ev_.foundered = 0; ev_.founderT = -1;
    std::fill(overT_.begin(), overT_.end(), -1.0);
    std::fill(markT_.begin(), markT_.end(), -1.0);
  }

  void evaluate_hydro() {  // simulation.cpp:137-145
    const Frame F = frame_now();
    Buoy hb;
    switch (numerics_) {
      case kPortable: hb = buoyancy<Order::Tree>(ship_, F); break;
      case kPortable32: hb = buoyancy<Order::TreeMixed>(ship_, F); break;
      default: hb = buoyancy<Order::Serial>(ship_, F); break;
    }
    out_.inflow = 0;
    out_.Vb = hb.V; out_.hullTop = hb.top; out_.bx = hb.x; out_.by = hb.y; out_.bz = hb.z;
    Real mw = 0;
    for (Index n = 0; n < ship_.nn; ++n) mw += net_.rho * vol_[n];
    out_.mw = mw;
  }

  void step() {  // simulation.cpp:147-155
    switch (numerics_) {
      case kOracle: orc::step<MathOracle, Order::Serial>(ship_, net_, st_, sc_, out_, ev_); break;
      case kPortable: orc::step<MathPortable, Order::Tree>(ship_, net_, st_, sc_, out_, ev_); break;
      case kPortable32: orc::step<MathPortable, Order::TreeMixed>(ship_, net_, st_, sc_, out_, ev_); break;
      default: orc::step<MathStd, Order::Serial>(ship_, net_, st_, sc_, out_, ev_); break;
    }
  }

  Readouts readouts() const {
    switch (numerics_) {
      case kOracle: return orc::readouts<MathOracle>(ship_, net_, sim_.readout[0], sim_.readout[1], st_, out_);
      case kStd: return orc::readouts<MathStd>(ship_, net_, sim_.readout[0], sim_.readout[1], st_, out_);
      default: return orc::readouts<MathPortable>(ship_, net_, sim_.readout[0], sim_.readout[1], st_, out_);
    }
  }

  // simulation.cpp:167-192: a history record whenever t >= next, then a step; a closing record at the end
  long long run(double tMax, double every, int stopWhenStable, std::vector<double>& curve) {
    long long steps = 0;
    Real next = 0, calm = 0;
    auto record = [&]() {
      const Readouts r = readouts();
      const double rec[7] = {r.t, r.trimDeg, r.listDeg, r.waterT, r.inflowTpm, r.draftF, r.draftA};
      curve.insert(curve.end(), rec, rec + 7);
    };
    while (st_.t < tMax && !ev_.foundered) {
      if (st_.t >= next) { record(); next += every; }
      step();
      steps++;
      if (stopWhenStable && st_.t > 1800) {
        calm = (std::fabs(out_.inflow) < 0.02 && std::fabs(st_.wth) < 1e-5 && std::fabs(st_.vz) < 1e-3) ? calm + sim_.dt : 0;
        if (calm > 1200) break;
      }
    }
    record();
    return steps;
  }

  // one record of the complete state, the trace layout (trace.cpp:11-21)
  void write_record(double* rec) const {
    const Index NN = ship_.nn;
    rec[0] = st_.t; rec[1] = st_.zO; rec[2] = st_.pitch; rec[3] = st_.roll; rec[4] = st_.vz; rec[5] = st_.wth; rec[6] = st_.wph;
    rec[7] = out_.inflow; rec[8] = out_.Vb; rec[9] = out_.hullTop;
    for (Index n = 0; n < NN; ++n) {
      rec[10 + n] = vol_[n]; rec[10 + NN + n] = level_[n]; rec[10 + 2 * NN + n] = cx_[n]; rec[10 + 3 * NN + n] = cy_[n]; rec[10 + 4 * NN + n] = cz_[n];
    }
  }

  // the world z of mark i at the current state's pose, and monitor m's flow sum of the last step
  Real mark_z(Index i) const { return world_z(frame_now(), sim_.mkx[i], sim_.mky[i], sim_.mkz[i]); }
  Real monitor_sum(Index m) const { return over_[m]; }

  Index nodes() const { return ship_.nn; }
  const State& state() const { return st_; }
  const Outputs& outputs() const { return out_; }
  const Events& events() const { return ev_; }
  const Scratch& scratch() const { return sc_; }
  const std::vector<Real>& flows() const { return q_; }
  const Ship& ship_view() const { return ship_; }
  const Net& net_view() const { return net_; }
  Index max_group() const { return static_cast<Index>(passes_.size()); }

 private:
  int numerics_;
  Ship ship_{};
  orc_sim sim_{};
  Net net_{};
  std::vector<Byte> en_;
  std::vector<Real> area_, tOn_;
  std::vector<Index> incStart_, incCount_, incConn_, sea_;
  std::vector<Byte> incSide_;
  std::vector<Real> vol_, level_, sArea_, cx_, cy_, cz_, zmin_, zmax_, hEff_, aEff_, acc_, excess_, over_, q_, dV_;
  std::vector<Index> deg_;
  std::vector<Byte> merged_;
  std::vector<Pass> passes_;
  std::vector<Event> evItems_;
  std::vector<Real> overT_, markT_;
  State st_{};
  Scratch sc_{};
  Outputs out_{};
  Events ev_{};

  // model.cpp:26-56: node -> connection CSR in increasing connection order, sea connections in order
  void build_incidence() {
    const Index NN = ship_.nn, NC = sim_.nc;
    incCount_.assign(NN, 0);
    for (Index c = 0; c < NC; ++c) {
      if (sim_.ca[c] >= 0) incCount_[sim_.ca[c]]++;
      incCount_[sim_.cb[c]]++;
    }
    incStart_.assign(NN, 0);
    Index total = 0;
    for (Index n = 0; n < NN; ++n) { incStart_[n] = total; total += incCount_[n]; }
    incConn_.assign(total, -1);
    incSide_.assign(total, 0);
    std::vector<Index> fill(NN, 0);
    for (Index c = 0; c < NC; ++c) {
      if (sim_.ca[c] >= 0) {
        const Index n = sim_.ca[c];
        const Index k = incStart_[n] + fill[n]++;
        incConn_[k] = c; incSide_[k] = 0;
      } else {
        sea_.push_back(c);
      }
      const Index n = sim_.cb[c];
      const Index k = incStart_[n] + fill[n]++;
      incConn_[k] = c; incSide_[k] = 1;
    }
  }

  void bind() {
    const orc_sim& m = sim_;
    net_.rho = m.rho; net_.g = m.g; net_.sq2g = std::sqrt(2 * m.g); net_.dt = m.dt;
    net_.ng = m.ng; net_.gStart = m.gStart; net_.gCount = m.gCount; net_.gNodes = m.gNodes; net_.gMode = m.gMode;
    net_.gpx = m.gpx; net_.gpy = m.gpy; net_.gpz = m.gpz; net_.gMargin = m.gMargin;
    net_.nc = m.nc; net_.ca = m.ca; net_.cb = m.cb; net_.law = m.law; net_.cx = m.cx; net_.cy = m.cy; net_.cz = m.cz;
    net_.area = area_.data(); net_.coef = m.coef; net_.en = en_.data(); net_.tOn = tOn_.data();
    net_.skipGroup = m.skipGroup; net_.monitor = m.monitor; net_.q = q_.data();
    net_.incStart = incStart_.data(); net_.incCount = incCount_.data(); net_.incConn = incConn_.data(); net_.incSide = incSide_.data();
    net_.seaCount = static_cast<Index>(sea_.size()); net_.sea = sea_.data();
    net_.nmon = m.nmon; net_.monThreshold = m.monThreshold;
    net_.nmark = m.nmark; net_.mkx = m.mkx; net_.mky = m.mky; net_.mkz = m.mkz; net_.mkWhen = m.mkWhen;
    net_.hullTopBelow = m.founder[0]; net_.pitchAbsAbove = m.founder[1]; net_.buoyancyBelow = m.founder[2]; net_.minT = m.founder[3];
    const double* b = m.body;
    net_.body = Body{b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], b[8], b[9], b[10], b[11], b[12], b[13], b[14], b[15]};
    st_.vol = vol_.data(); st_.level = level_.data();
    sc_ = Scratch{sArea_.data(), cx_.data(), cy_.data(), cz_.data(), zmin_.data(), zmax_.data(), hEff_.data(), aEff_.data(),
                  acc_.data(), excess_.data(), deg_.data(), merged_.data(), passes_.data(), over_.data(), dV_.data()};
    ev_.items = evItems_.data(); ev_.capacity = static_cast<Index>(evItems_.size());
    ev_.overT = overT_.data(); ev_.markT = markT_.data();
  }

  Frame frame_now() const {
    switch (numerics_) {
      case kStd: return pose<MathStd>(net_.body, st_.zO, st_.pitch, st_.roll);
      default: return pose<MathPortable>(net_.body, st_.zO, st_.pitch, st_.roll);
    }
  }
};

}  // namespace

extern "C" {

// --------------------------------------------------------------------------- simulation handle

void* orc_new(const orc_ship* s, const orc_sim* m, int numerics) { return new Sim(s, m, numerics); }
void orc_free(void* h) { delete static_cast<Sim*>(h); }
void orc_set_conn(void* h, int32_t c, int en, double area, double tOn) { static_cast<Sim*>(h)->set_conn(c, en, area, tOn); }
void orc_reset(void* h) { static_cast<Sim*>(h)->reset_to_initial(); }
void orc_restore(void* h, const double* s7, const double* vol, const double* level, const double* cx, const double* cy, const double* cz) {
  static_cast<Sim*>(h)->restore(s7, vol, level, cx, cy, cz);
}
void orc_step(void* h) { static_cast<Sim*>(h)->step(); }
void orc_steps(void* h, int64_t n) {
  Sim* S = static_cast<Sim*>(h);
  for (int64_t i = 0; i < n; ++i) S->step();
}
void orc_evaluate_hydro(void* h) { static_cast<Sim*>(h)->evaluate_hydro(); }

// the complete record (trace layout, 10 + 5 * nodes values) and the seven outputs
void orc_record(void* h, double* rec) { static_cast<Sim*>(h)->write_record(rec); }
void orc_outputs(void* h, double* o7) {
  const Outputs& o = static_cast<Sim*>(h)->outputs();
  o7[0] = o.inflow; o7[1] = o.Vb; o7[2] = o.hullTop; o7[3] = o.mw; o7[4] = o.bx; o7[5] = o.by; o7[6] = o.bz;
}
void orc_flows(void* h, double* q) {
  const std::vector<Real>& f = static_cast<Sim*>(h)->flows();
  std::copy(f.begin(), f.end(), q);
}
void orc_scratch(void* h, double* area, double* zmin, double* zmax, uint8_t* merged, int32_t ng) {
  Sim* S = static_cast<Sim*>(h);
  const Scratch& sc = S->scratch();
  for (Index n = 0; n < S->nodes(); ++n) { area[n] = sc.area[n]; zmin[n] = sc.zmin[n]; zmax[n] = sc.zmax[n]; }
  for (Index g = 0; g < ng; ++g) merged[g] = sc.merged[g];
}
int32_t orc_events(void* h, double* t, int32_t* id, uint8_t* kind, int32_t cap, int32_t* foundered, double* founderT) {
  const Events& e = static_cast<Sim*>(h)->events();
  const int32_t n = std::min(e.count, cap);
  for (int32_t k = 0; k < n; ++k) { t[k] = e.items[k].t; id[k] = e.items[k].id; kind[k] = e.items[k].kind; }
  *foundered = e.foundered;
  *founderT = e.founderT;
  return e.count;
}
void orc_readouts(void* h, double* r8) {
  const Readouts r = static_cast<Sim*>(h)->readouts();
  r8[0] = r.t; r8[1] = r.trimDeg; r8[2] = r.listDeg; r8[3] = r.draftF; r8[4] = r.draftA; r8[5] = r.waterM3; r8[6] = r.waterT; r8[7] = r.inflowTpm;
}

// simulation.cpp:194-221: the initial record, every step up to `dense`, then every `sampleEvery` seconds.
// Returns the step count; denseN / sampleN receive the record counts (capped by the capacities).
int64_t orc_record_trace(void* h, double tMax, int32_t dense, double sampleEvery, double* denseRec, int32_t denseCap,
                         double* sampleRec, int32_t sampleCap, int32_t* denseN, int32_t* sampleN) {
  Sim* S = static_cast<Sim*>(h);
  const Index rec = 10 + 5 * S->nodes();
  S->evaluate_hydro();
  int32_t nd = 0, ns = 0;
  if (nd < denseCap) S->write_record(denseRec + static_cast<size_t>(nd++) * rec);
  if (ns < sampleCap) S->write_record(sampleRec + static_cast<size_t>(ns++) * rec);
  Real next = sampleEvery;
  int64_t steps = 0;
  while (!S->events().foundered && S->state().t < tMax) {
    S->step();
    ++steps;
    if (steps <= dense && nd < denseCap) S->write_record(denseRec + static_cast<size_t>(nd++) * rec);
    if (S->state().t >= next - 1e-9) {
      if (ns < sampleCap) S->write_record(sampleRec + static_cast<size_t>(ns++) * rec);
      next += sampleEvery;
    }
  }
  *denseN = nd;
  *sampleN = ns;
  return steps;
}

// --------------------------------------------------------------------------- threaded sweep (sweep.cpp:231-271)

// Per instance i: connection fields en/area/tOn at [i * nc]; results: steps, foundered, founderT, tEnd,
// curve records (7 values, at most curveCap) and events (at most evCap).
void orc_sweep(const orc_ship* s, const orc_sim* m, int numerics, int32_t count, const uint8_t* en, const double* area,
               const double* tOn, double tMax, double every, int stopWhenStable, int threads, int64_t* steps,
               uint8_t* foundered, double* founderT, double* tEnd, double* curve, int32_t curveCap, int32_t* curveN,
               double* evT, int32_t* evId, uint8_t* evKind, int32_t evCap, int32_t* evN) {
  std::atomic<int32_t> next{0};
  const int32_t nc = m->nc;
  auto worker = [&]() {
    std::vector<double> cv;
    for (;;) {
      const int32_t i = next.fetch_add(1);
      if (i >= count) return;
      Sim S(s, m, numerics);
      for (int32_t c = 0; c < nc; ++c) {
        const size_t k = static_cast<size_t>(i) * nc + c;
        S.set_conn(c, en[k], area[k], tOn[k]);
      }
      cv.clear();
      steps[i] = S.run(tMax, every, stopWhenStable, cv);
      const Events& e = S.events();
      foundered[i] = e.foundered;
      founderT[i] = e.founderT;
      tEnd[i] = S.state().t;
      const int32_t nrec = static_cast<int32_t>(cv.size() / 7);
      curveN[i] = nrec;
      const int32_t keep = std::min(nrec, curveCap);
      std::copy(cv.begin(), cv.begin() + static_cast<std::ptrdiff_t>(keep) * 7, curve + static_cast<size_t>(i) * curveCap * 7);
      evN[i] = e.count;
      for (int32_t k = 0; k < std::min(e.count, evCap); ++k) {
        const size_t o = static_cast<size_t>(i) * evCap + k;
        evT[o] = e.items[k].t; evId[o] = e.items[k].id; evKind[o] = e.items[k].kind;
      }
    }
  };
  if (threads <= 0) threads = static_cast<int>(std::thread::hardware_concurrency());
  if (threads <= 0) threads = 1;
  const int n = std::min<int>(threads, count);
  if (n <= 1) {  // on the calling thread (its thread-local operation counter then sees the runs)
    worker();
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < n; ++t) pool.emplace_back(worker);
  for (std::thread& t : pool) t.join();
}

// --------------------------------------------------------------------------- kernel-level probes

// order: 0 serial, 1 tree, 2 tree mixed; math: 0 portable, 1 oracle, 2 std
void orc_pose(const double* body16, double zO, double pitch, double roll, int math, double* f13) {
  const double* b = body16;
  const Body B{b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], b[8], b[9], b[10], b[11], b[12], b[13], b[14], b[15]};
  Frame F = math == 2 ? pose<MathStd>(B, zO, pitch, roll) : pose<MathPortable>(B, zO, pitch, roll);
  const double v[13] = {F.R00, F.R01, F.R02, F.R10, F.R11, F.R12, F.R20, F.R21, F.R22, F.tx, F.ty, F.tz, F.cth};
  std::memcpy(f13, v, sizeof v);
}

static Ship ship_of(const orc_ship* s) {
  Ship S{};
  S.ncol = s->ncol; S.x = s->x; S.y = s->y; S.dxdy = s->dxdy; S.zlo = s->zlo; S.zhi = s->zhi;
  S.nseg = s->nseg; S.segCol = s->segCol; S.segA = s->segA; S.segE = s->segE;
  S.nn = s->nn; S.mu = s->mu; S.vmax = s->vmax; S.amin = s->amin; S.aov = s->aov; S.segStart = s->segStart; S.segCount = s->segCount;
  return S;
}

static Frame frame_of(const double* f13) {
  Frame F;
  F.R00 = f13[0]; F.R01 = f13[1]; F.R02 = f13[2]; F.R10 = f13[3]; F.R11 = f13[4]; F.R12 = f13[5];
  F.R20 = f13[6]; F.R21 = f13[7]; F.R22 = f13[8]; F.tx = f13[9]; F.ty = f13[10]; F.tz = f13[11]; F.cth = f13[12];
  return F;
}

void orc_buoyancy(const orc_ship* s, const double* f13, int order, double* b5) {
  const Ship S = ship_of(s);
  const Frame F = frame_of(f13);
  const Buoy b = order == 0 ? buoyancy<Order::Serial>(S, F) : order == 1 ? buoyancy<Order::Tree>(S, F) : buoyancy<Order::TreeMixed>(S, F);
  b5[0] = b.V; b5[1] = b.x; b5[2] = b.y; b5[3] = b.z; b5[4] = b.top;
}

void orc_node_pass(const orc_ship* s, const double* f13, int order, int32_t n, double h, double* p7) {
  const Ship S = ship_of(s);
  const Frame F = frame_of(f13);
  const Pass p = order == 0 ? node_pass<Order::Serial>(S, F, n, h) : order == 1 ? node_pass<Order::Tree>(S, F, n, h) : node_pass<Order::TreeMixed>(S, F, n, h);
  p7[0] = p.vol; p7[1] = p.dv; p7[2] = p.mx; p7[3] = p.my; p7[4] = p.mz; p7[5] = p.zmin; p7[6] = p.zmax;
}

double orc_pow_v8(double x, double y) { return MathOracle::pow(x, y); }

// Operation counting on the calling thread (counts: cwet cdry spart sfull sdry passes steps active pows wetNodes).
void orc_count_begin(void) {
  g_ops = OpCount{};
  g_counting = true;
}
void orc_count_end(int64_t* out10) {
  g_counting = false;
  const long long v[10] = {g_ops.cwet, g_ops.cdry, g_ops.spart, g_ops.sfull, g_ops.sdry, g_ops.passes, g_ops.steps, g_ops.active, g_ops.pows,
                           g_ops.wetNodes};
  for (int k = 0; k < 10; ++k) out10[k] = v[k];
}

}  // extern "C"

// --------------------------------------------------------------------------- the dual-number build's driver

extern "C" {

// Scale-factor parameters: connection c's area takes theta[pa[c]], its coefficient theta[pc[c]] (-1: none).
struct orc_params {
  int32_t np;
  const int32_t *pa, *pc;
  int32_t relaxed;  // 1: relaxed level derivatives (dual_step.hpp relax_level)
};

// The loss of a run (no counterpart in the reference; it extends the reference's calibration objective,
// tools/py/calibrate_gpu.py:35-64, to terms with derivatives): weighted squared errors against observations of
//   readout terms j: the trim (kind 0: pitch * 180 / pi) or list (kind 1: roll * 180 / pi) after step rdStep[j]
//                    (readouts.hpp:26-27); a step after the end of the run scores rdAfter[j] (no derivative), as
//                    the reference objective scores trim and list after foundering;
//   event terms e:   the time at which monitor idx (kind 0) first exceeds the threshold, mark idx (kind 1) first
//                    goes under (or above), or the ship founders (kind 2), from the event's signed signal s_k at
//                    step k: monitor flow sum - threshold; -world z of an 'under' mark, +world z of an 'above' one;
//                    for foundering the largest margin of the founder rule (events.hpp:32-37): hullTopBelow - hull
//                    top, |pitch| - pitchAbsAbove, and after minT buoyancyBelow - V (a tie keeps the earlier). With
//                    k the first step at which s_k > 0, t* = t_k - dt * s_k / (s_k - s_{k-1}), the interpolated
//                    crossing (t_1 if k = 1). An event that never fires scores the founder time if the ship
//                    foundered and evMissing[e] if it stayed afloat, as the reference objective does (no derivative).
struct orc_loss {
  int32_t nRead;
  const int64_t* rdStep;
  const int32_t* rdKind;
  const double *rdObs, *rdW, *rdAfter;
  int32_t nEv;
  const int32_t *evKind, *evIdx;
  const double *evObs, *evW, *evMissing;
};

}  // extern "C"

namespace {

template <class T>
void lift(const double* v, int nd, const double* dv, size_t stride, T* out, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    out[i] = T(v[i]);
    if constexpr (!std::is_same_v<T, double>)
      for (int k = 0; k < nd; ++k) out[i].d[k] = dv ? dv[k * stride + i] : 0.0;
  }
}

template <class T>
void drop(const T& x, int nd, double* v, double* dv, size_t stride, size_t i) {
  v[i] = dual::val(x);
  if constexpr (!std::is_same_v<T, double>)
    if (dv)
      for (int k = 0; k < nd; ++k) dv[k * stride + i] = x.d[k];
}

// the state as a trace record (trace.cpp:11-21 layout; inflow is not carried and stays 0)
template <class T>
void state_record(const dual::Sim<T>& d, Index NN, int nd, double* rec, double* drec) {
  const size_t R = 10 + 5 * static_cast<size_t>(NN);
  if (drec) std::fill(drec, drec + nd * R, 0.0);
  rec[0] = d.t;
  drop(d.zO, nd, rec, drec, R, 1); drop(d.pitch, nd, rec, drec, R, 2); drop(d.roll, nd, rec, drec, R, 3);
  drop(d.vz, nd, rec, drec, R, 4); drop(d.wth, nd, rec, drec, R, 5); drop(d.wph, nd, rec, drec, R, 6);
  rec[7] = 0;
  drop(d.Vb, nd, rec, drec, R, 8); drop(d.hullTop, nd, rec, drec, R, 9);
  for (Index n = 0; n < NN; ++n) {
    drop(d.vol[n], nd, rec, drec, R, 10 + n); drop(d.level[n], nd, rec, drec, R, 10 + NN + n);
    drop(d.cx[n], nd, rec, drec, R, 10 + 2 * NN + n); drop(d.cy[n], nd, rec, drec, R, 10 + 3 * NN + n);
    drop(d.cz[n], nd, rec, drec, R, 10 + 4 * NN + n);
  }
}

template <class T>
void restore_record(dual::Sim<T>& d, Index NN, int nd, const double* rec, const double* drec) {
  const size_t R = 10 + 5 * static_cast<size_t>(NN);
  std::vector<T> s6(6), v(5 * NN);
  const int at6[6] = {1, 2, 3, 4, 5, 6};
  for (int k = 0; k < 6; ++k) {
    s6[k] = T(rec[at6[k]]);
    if constexpr (!std::is_same_v<T, double>)
      for (int j = 0; j < nd; ++j) s6[k].d[j] = drec ? drec[j * R + at6[k]] : 0.0;
  }
  lift(rec + 10, nd, drec ? drec + 10 : nullptr, R, v.data(), 5 * static_cast<size_t>(NN));
  d.restore(rec[0], s6.data(), v.data(), v.data() + NN, v.data() + 2 * NN, v.data() + 3 * NN, v.data() + 4 * NN);
}

// The signed signal of loss event (kind, idx) in the last step (see orc_loss).
template <class T>
T event_signal(const dual::Sim<T>& d, const Net& M, int32_t kind, int32_t idx) {
  if (kind == 0) return d.over_[idx] - M.monThreshold;
  if (kind == 1) return M.mkWhen[idx] == 0 ? -d.markZ_[idx] : d.markZ_[idx];
  T s = M.hullTopBelow - d.hullTop;
  const T sp = dual::dfabs(d.pitch) - M.pitchAbsAbove;
  if (sp > s) s = sp;
  if (d.t > M.minT) {
    const T sb = M.buoyancyBelow - d.Vb;
    if (sb > s) s = sb;
  }
  return s;
}

template <int N>
int64_t dual_run(const orc_ship* s, const orc_sim* m, const uint8_t* en, const double* area, const double* tOn,
                 const orc_params* prm, const double* theta, int32_t nd, const double* seed, double tMax, int64_t maxSteps,
                 const orc_loss* L, double* lossOut, double* sigOut, double* recOut) {
  using T = dual::D<N>;
  Sim o(s, m, kPortable);
  for (int32_t c = 0; c < m->nc; ++c) o.set_conn(c, en[c], area[c], tOn[c]);
  const Ship& S = o.ship_view();
  const Net& M = o.net_view();
  const Index NN = S.nn;
  const Index gmax = o.max_group();
  dual::Sim<T> d(S, M, dual::Params{prm->np, prm->pa, prm->pc, prm->relaxed}, m->coef, &gmax);
  std::vector<T> th(prm->np);
  lift(theta, nd, seed, static_cast<size_t>(prm->np), th.data(), static_cast<size_t>(prm->np));
  d.set_params(th.data());
  std::vector<T> s6(6), v(5 * NN);
  for (int k = 0; k < 6; ++k) s6[k] = T(m->state0[1 + k]);
  lift(m->vol0, 0, nullptr, 0, v.data(), NN);
  lift(m->level0, 0, nullptr, 0, v.data() + NN, NN);
  lift(m->cx0, 0, nullptr, 0, v.data() + 2 * NN, NN);
  lift(m->cy0, 0, nullptr, 0, v.data() + 3 * NN, NN);
  lift(m->cz0, 0, nullptr, 0, v.data() + 4 * NN, NN);
  d.restore(m->state0[0], s6.data(), v.data(), v.data() + NN, v.data() + 2 * NN, v.data() + 3 * NN, v.data() + 4 * NN);

  const int32_t nT = L ? L->nRead : 0, nE = L ? L->nEv : 0;
  std::vector<T> rd(nT), tstar(nE), prev(nE);
  std::vector<char> rdSet(nT, 0), fired(nE, 0);
  int64_t steps = 0;
  while (d.t < tMax && !d.foundered && steps < maxSteps) {
    d.step();
    ++steps;
    for (int32_t j = 0; j < nT; ++j)
      if (L->rdStep[j] == steps) { rd[j] = (L->rdKind[j] == 0 ? d.pitch : d.roll) * 180.0 / kPi; rdSet[j] = 1; }
    for (int32_t e = 0; e < nE; ++e) {
      const T sg = event_signal(d, M, L->evKind[e], L->evIdx[e]);
      if (!fired[e] && sg > 0) {
        fired[e] = 1;
        tstar[e] = steps == 1 ? T(d.t) : d.t - M.dt * sg / (sg - prev[e]);
      }
      prev[e] = sg;
    }
  }
  T loss = T(0);
  for (int32_t j = 0; j < nT; ++j) {
    if (!rdSet[j]) rd[j] = T(L->rdAfter[j]);
    const T r = rd[j] - L->rdObs[j];
    loss += L->rdW[j] * (r * r);
  }
  for (int32_t e = 0; e < nE; ++e) {
    if (!fired[e]) tstar[e] = T(d.foundered ? d.founderT : L->evMissing[e]);
    const T r = tstar[e] - L->evObs[e];
    loss += L->evW[e] * (r * r);
  }
  if (lossOut) drop(loss, nd, lossOut, lossOut + 1, 1, 0);
  if (sigOut) {
    const size_t K = static_cast<size_t>(nT + nE);
    for (int32_t j = 0; j < nT; ++j) drop(rd[j], nd, sigOut, sigOut + K, K, j);
    for (int32_t e = 0; e < nE; ++e) drop(tstar[e], nd, sigOut, sigOut + K, K, nT + e);
  }
  if (recOut) {
    const size_t R = 10 + 5 * static_cast<size_t>(NN);
    state_record(d, NN, nd, recOut, recOut + R);
  }
  return steps;
}

}  // namespace

extern "C" {

// A run of one instance from the scenario's initial state (to tMax, founder or maxSteps) in the dual build, with
// theta and nd <= 24 tangent directions of theta (seed[k * np + i]: direction k, parameter i). Outputs (each
// optional): lossOut[1 + nd] the loss and its directional derivatives; sigOut[(1 + nd) * (nRead + nEv)] the
// readout values and event times, then their derivatives direction by direction; recOut[(1 + nd) * (10 + 5 NN)] the
// final state as a trace record, then its derivatives. Returns the step count, or -1 for nd > 24.
int64_t orc_dual_run(const orc_ship* s, const orc_sim* m, const uint8_t* en, const double* area, const double* tOn,
                     const orc_params* prm, const double* theta, int32_t nd, const double* seed, double tMax, int64_t maxSteps,
                     const orc_loss* L, double* lossOut, double* sigOut, double* recOut) {
  if (nd <= 1) return dual_run<1>(s, m, en, area, tOn, prm, theta, nd, seed, tMax, maxSteps, L, lossOut, sigOut, recOut);
  if (nd <= 12) return dual_run<12>(s, m, en, area, tOn, prm, theta, nd, seed, tMax, maxSteps, L, lossOut, sigOut, recOut);
  if (nd <= 24) return dual_run<24>(s, m, en, area, tOn, prm, theta, nd, seed, tMax, maxSteps, L, lossOut, sigOut, recOut);
  return -1;
}

// The loss of one run (orc_loss) in plain arithmetic and any numerics: the sweep's step with per-connection
// discharge coefficients coef[nc] (the scenario's, scaled) and areas area[nc] (scaled), from the scenario's initial
// state to tMax, founder or maxSteps; the same terms, signals and order as the dual build's run. Writes the loss and
// the signals (readout values, then event times); returns the step count.
int64_t orc_loss_run(const orc_ship* s, const orc_sim* m, int numerics, const uint8_t* en, const double* area, const double* tOn,
                     const double* coef, double tMax, int64_t maxSteps, const orc_loss* L, double* lossOut, double* sigOut) {
  orc_sim mc = *m;
  mc.coef = coef;
  Sim o(s, &mc, numerics);
  for (int32_t c = 0; c < m->nc; ++c) o.set_conn(c, en[c], area[c], tOn[c]);
  const int32_t nT = L ? L->nRead : 0, nE = L ? L->nEv : 0;
  std::vector<double> rd(nT, 0.0), tstar(nE, 0.0), prev(nE, 0.0), zpre(nE, 0.0);
  std::vector<char> rdSet(nT, 0), fired(nE, 0);
  const double dt = m->dt;
  int64_t steps = 0;
  while (o.state().t < tMax && !o.events().foundered && steps < maxSteps) {
    for (int32_t e = 0; e < nE; ++e)
      if (L->evKind[e] == 1) zpre[e] = o.mark_z(L->evIdx[e]);
    o.step();
    ++steps;
    const State& st = o.state();
    for (int32_t j = 0; j < nT; ++j)
      if (L->rdStep[j] == steps) { rd[j] = (L->rdKind[j] == 0 ? st.pitch : st.roll) * 180.0 / kPi; rdSet[j] = 1; }
    for (int32_t e = 0; e < nE; ++e) {
      double sg;
      const int32_t kind = L->evKind[e], idx = L->evIdx[e];
      if (kind == 0) sg = o.monitor_sum(idx) - m->monThreshold;
      else if (kind == 1) sg = m->mkWhen[idx] == 0 ? -zpre[e] : zpre[e];
      else {
        sg = m->founder[0] - o.outputs().hullTop;
        const double sp = std::fabs(st.pitch) - m->founder[1];
        if (sp > sg) sg = sp;
        if (st.t > m->founder[3]) {
          const double sb = m->founder[2] - o.outputs().Vb;
          if (sb > sg) sg = sb;
        }
      }
      if (!fired[e] && sg > 0) {
        fired[e] = 1;
        tstar[e] = steps == 1 ? st.t : st.t - dt * sg / (sg - prev[e]);
      }
      prev[e] = sg;
    }
  }
  const bool foundered = o.events().foundered != 0;
  double loss = 0;
  for (int32_t j = 0; j < nT; ++j) {
    if (!rdSet[j]) rd[j] = L->rdAfter[j];
    const double r = rd[j] - L->rdObs[j];
    loss += L->rdW[j] * (r * r);
  }
  for (int32_t e = 0; e < nE; ++e) {
    if (!fired[e]) tstar[e] = foundered ? o.events().founderT : L->evMissing[e];
    const double r = tstar[e] - L->evObs[e];
    loss += L->evW[e] * (r * r);
  }
  if (lossOut) *lossOut = loss;
  if (sigOut) {
    for (int32_t j = 0; j < nT; ++j) sigOut[j] = rd[j];
    for (int32_t e = 0; e < nE; ++e) sigOut[nT + e] = tstar[e];
  }
  return steps;
}

// One step of the dual build from a state record and its tangent along one direction (drec, same layout;
// entries 0 and 7-9 are ignored), with theta and its tangent dtheta. Writes the state after the step and its
// tangent, and the step's signals: sig[nmon + nmark] = monitor flow sums then mark world z, dsig their tangents.
void orc_dual_step(const orc_ship* s, const orc_sim* m, const uint8_t* en, const double* area, const double* tOn,
                   const orc_params* prm, const double* theta, const double* dtheta, const double* rec, const double* drec,
                   double* recOut, double* drecOut, double* sig, double* dsig) {
  using T = dual::D<1>;
  Sim o(s, m, kPortable);
  for (int32_t c = 0; c < m->nc; ++c) o.set_conn(c, en[c], area[c], tOn[c]);
  const Ship& S = o.ship_view();
  const Net& M = o.net_view();
  const Index NN = S.nn, gmax = o.max_group();
  dual::Sim<T> d(S, M, dual::Params{prm->np, prm->pa, prm->pc, prm->relaxed}, m->coef, &gmax);
  std::vector<T> th(prm->np);
  lift(theta, 1, dtheta, static_cast<size_t>(prm->np), th.data(), static_cast<size_t>(prm->np));
  d.set_params(th.data());
  restore_record(d, NN, 1, rec, drec);
  d.step();
  state_record(d, NN, 1, recOut, drecOut);
  for (Index k = 0; k < M.nmon; ++k) drop(d.over_[k], 1, sig, dsig, 0, k);
  for (Index k = 0; k < M.nmark; ++k) drop(d.markZ_[k], 1, sig, dsig, 0, M.nmon + k);
}

}  // extern "C"
