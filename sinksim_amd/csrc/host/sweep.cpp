#include "sweep.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "json.hpp"

namespace sinksim::sweep {

using json::Value;

namespace {

Byte kind_byte(const CompiledSim& sim, const std::string& name, const std::string& where) {
  const Index k = sim.kindIndex(name);
  if (k < 0) throw std::runtime_error(where + ": unknown connection kind \"" + name + "\"");
  return static_cast<Byte>(k);
}

std::string kind_name(const CompiledSim& sim, Byte kind) {
  const auto i = static_cast<size_t>(kind);
  return i < sim.kindNames.size() ? sim.kindNames[i] : std::to_string(kind);
}

}  // namespace

Sweep load_sweep(const std::string& path, const CompiledSim& sim) {
  const Value doc = json::read_file(path);
  if (json::string_or(doc, "format", "") != "sinksim.sweep") throw std::runtime_error(path + ": format is not sinksim.sweep");
  if (json::integer(doc, "formatVersion", path) != 1) throw std::runtime_error(path + ": unsupported sweep formatVersion");
  Sweep sw;
  sw.ship = json::string_or(doc, "ship", "");
  sw.sim = json::string_or(doc, "sim", "");
  sw.tMax = json::number_or(doc, "tMax", 5 * 3600);
  sw.every = json::number_or(doc, "every", 30);
  const Value* sws = doc.find("stopWhenStable");
  sw.stopWhenStable = sws && sws->is_bool() && sws->as_bool();
  const Value& list = json::member(doc, "instances", path);
  if (!list.is_array()) throw std::runtime_error(path + ": instances must be an array");
  size_t k = 0;
  for (const Value& j : list.items()) {
    Instance in;
    in.id = json::string_or(j, "id", "instance" + std::to_string(k));
    const std::string where = path + " instance " + in.id;
    if (const Value* s = j.find("scale")) {
      if (!s->is_object()) throw std::runtime_error(where + ": scale must be an object of kind -> factor");
      for (const auto& m : s->members()) {
        if (!m.second.is_number()) throw std::runtime_error(where + ": scale factor for " + m.first + " is not a number");
        in.kindScale.emplace_back(kind_byte(sim, m.first, where), m.second.as_double());
      }
    }
    if (const Value* l2 = j.find("connections")) {
      if (!l2->is_array()) throw std::runtime_error(where + ": connections must be an array");
      for (const Value& o : l2->items()) {
        ConnectionOverride ov;
        if (o.contains("conn")) ov.conn = static_cast<Index>(json::integer(o, "conn", where));
        else {
          ov.kind = kind_byte(sim, json::string(o, "kind", where), where);
          if (o.contains("idx")) ov.idx = static_cast<Index>(json::integer(o, "idx", where));
        }
        if (o.contains("enabled")) { ov.hasEnabled = true; ov.enabled = json::boolean(o, "enabled", where); }
        if (o.contains("tOn")) { ov.hasTOn = true; ov.tOn = json::number(o, "tOn", where); }
        if (o.contains("area")) { ov.hasArea = true; ov.area = json::number(o, "area", where); }
        in.overrides.push_back(ov);
      }
    }
    sw.instances.push_back(std::move(in));
    ++k;
  }
  return sw;
}

void save_sweep(const Sweep& sw, const std::string& path, const CompiledSim& sim) {
  Value doc = Value::object();
  doc["format"] = "sinksim.sweep";
  doc["formatVersion"] = 1;
  doc["ship"] = sw.ship;
  doc["sim"] = sw.sim;
  doc["tMax"] = sw.tMax;
  doc["every"] = sw.every;
  doc["stopWhenStable"] = sw.stopWhenStable;
  Value list = Value::array();
  for (const Instance& in : sw.instances) {
    Value j = Value::object();
    j["id"] = in.id;
    if (!in.kindScale.empty()) {
      Value s = Value::object();
      for (const auto& [kind, factor] : in.kindScale) s[kind_name(sim, kind)] = factor;
      j["scale"] = s;
    }
    if (!in.overrides.empty()) {
      Value l2 = Value::array();
      for (const ConnectionOverride& o : in.overrides) {
        Value x = Value::object();
        if (o.conn >= 0) x["conn"] = o.conn;
        else { x["kind"] = kind_name(sim, o.kind); if (o.idx >= 0) x["idx"] = o.idx; }
        if (o.hasEnabled) x["enabled"] = o.enabled;
        if (o.hasTOn) x["tOn"] = o.tOn;
        if (o.hasArea) x["area"] = o.area;
        l2.push_back(x);
      }
      j["connections"] = l2;
    }
    list.push_back(j);
  }
  doc["instances"] = list;
  json::write_file(path, doc, 2);
}

void save_results(const Results& r, const std::string& path) {
  Value doc = Value::object();
  doc["format"] = "sinksim.batch-results";
  doc["formatVersion"] = 1;
  doc["ship"] = r.ship;
  doc["sim"] = r.sim;
  doc["numerics"] = r.numerics;
  doc["engine"] = r.engine;
  doc["tMax"] = r.tMax;
  doc["every"] = r.every;
  Value list = Value::array();
  for (const InstanceResult& in : r.instances) {
    Value j = Value::object();
    j["id"] = in.id;
    j["foundered"] = in.foundered;
    j["founderT"] = in.founderT;
    j["tEnd"] = in.tEnd;
    j["steps"] = static_cast<long long>(in.steps);
    Value f = Value::object();
    f["trimDeg"] = in.final.trimDeg; f["listDeg"] = in.final.listDeg; f["waterT"] = in.final.waterT;
    f["draftF"] = in.final.draftF; f["draftA"] = in.final.draftA; f["inflowTpm"] = in.final.inflowTpm;
    j["final"] = f;
    Value ev = Value::array();
    for (const EventRecord& e : in.events) {
      Value x = Value::object();
      x["t"] = e.t; x["id"] = e.id; x["label"] = e.label;
      ev.push_back(x);
    }
    j["events"] = ev;
    std::vector<Real> t, trim, lst, water, inflow, dF, dA;
    for (const CurveRecord& k : in.curve) {
      t.push_back(k.t); trim.push_back(k.trim); lst.push_back(k.list); water.push_back(k.water); inflow.push_back(k.inflow);
      dF.push_back(k.dF); dA.push_back(k.dA);
    }
    Value c = Value::object();
    c["t"] = Value(t); c["trim"] = Value(trim); c["list"] = Value(lst); c["water"] = Value(water); c["inflow"] = Value(inflow);
    c["draftF"] = Value(dF); c["draftA"] = Value(dA);
    j["curve"] = c;
    list.push_back(j);
  }
  doc["instances"] = list;
  json::write_file(path, doc, 2);
}

Results load_results(const std::string& path) {
  const Value doc = json::read_file(path);
  if (json::string_or(doc, "format", "") != "sinksim.batch-results") throw std::runtime_error(path + ": format is not sinksim.batch-results");
  Results r;
  r.ship = json::string_or(doc, "ship", "");
  r.sim = json::string_or(doc, "sim", "");
  r.numerics = json::string_or(doc, "numerics", "");
  r.engine = json::string_or(doc, "engine", "");
  r.tMax = json::number_or(doc, "tMax", 0);
  r.every = json::number_or(doc, "every", 0);
  for (const Value& j : json::member(doc, "instances", path).items()) {
    InstanceResult in;
    in.id = json::string_or(j, "id", "");
    in.foundered = json::boolean(j, "foundered", path);
    in.founderT = json::number(j, "founderT", path);
    in.tEnd = json::number(j, "tEnd", path);
    in.steps = json::integer(j, "steps", path);
    const Value& f = json::member(j, "final", path);
    in.final.trimDeg = json::number(f, "trimDeg", path); in.final.listDeg = json::number(f, "listDeg", path);
    in.final.waterT = json::number(f, "waterT", path); in.final.draftF = json::number(f, "draftF", path);
    in.final.draftA = json::number(f, "draftA", path); in.final.inflowTpm = json::number(f, "inflowTpm", path);
    in.final.t = in.tEnd;
    for (const Value& e : json::member(j, "events", path).items())
      in.events.push_back(EventRecord{json::number(e, "t", path), json::string(e, "id", path), json::string_or(e, "label", "")});
    const Value& c = json::member(j, "curve", path);
    auto col = [&](const char* k) {
      std::vector<Real> v;
      for (const Value& x : json::member(c, k, path).items()) v.push_back(x.as_double());
      return v;
    };
    const std::vector<Real> t = col("t"), trim = col("trim"), lst = col("list"), water = col("water"), inflow = col("inflow"), dF = col("draftF"),
                            dA = col("draftA");
    for (size_t i = 0; i < t.size(); ++i) in.curve.push_back(CurveRecord{t[i], trim[i], lst[i], water[i], inflow[i], dF[i], dA[i]});
    r.instances.push_back(std::move(in));
  }
  return r;
}

Diff compare(const Results& a, const Results& b) {
  Diff d;
  auto note = [&](const std::string& where, double x, double y) {
    ++d.compared;
    if (x != y && !(std::isnan(x) && std::isnan(y))) {
      ++d.differing;
      const double ad = std::fabs(x - y);
      if (ad > d.maxAbs) d.maxAbs = ad;
      if (d.first.empty()) d.first = where + ": " + json::format_number(x) + " vs " + json::format_number(y);
    }
  };
  if (a.instances.size() != b.instances.size()) {
    d.differing++;
    d.first = "instance count " + std::to_string(a.instances.size()) + " vs " + std::to_string(b.instances.size());
    return d;
  }
  for (size_t i = 0; i < a.instances.size(); ++i) {
    const InstanceResult& x = a.instances[i];
    const InstanceResult& y = b.instances[i];
    const std::string w = "instance " + x.id;
    note(w + " foundered", x.foundered, y.foundered);
    note(w + " founderT", x.founderT, y.founderT);
    note(w + " tEnd", x.tEnd, y.tEnd);
    note(w + " steps", static_cast<double>(x.steps), static_cast<double>(y.steps));
    note(w + " trim", x.final.trimDeg, y.final.trimDeg);
    note(w + " list", x.final.listDeg, y.final.listDeg);
    note(w + " water", x.final.waterT, y.final.waterT);
    note(w + " draftF", x.final.draftF, y.final.draftF);
    note(w + " draftA", x.final.draftA, y.final.draftA);
    note(w + " events", static_cast<double>(x.events.size()), static_cast<double>(y.events.size()));
    for (size_t k = 0; k < std::min(x.events.size(), y.events.size()); ++k) {
      note(w + " event " + x.events[k].id + " t", x.events[k].t, y.events[k].t);
      if (x.events[k].id != y.events[k].id) {
        ++d.differing;
        if (d.first.empty()) d.first = w + " event id " + x.events[k].id + " vs " + y.events[k].id;
      }
    }
    note(w + " curve length", static_cast<double>(x.curve.size()), static_cast<double>(y.curve.size()));
    for (size_t k = 0; k < std::min(x.curve.size(), y.curve.size()); ++k) {
      const std::string c = w + " curve[" + std::to_string(k) + "]";
      note(c + " t", x.curve[k].t, y.curve[k].t);
      note(c + " trim", x.curve[k].trim, y.curve[k].trim);
      note(c + " list", x.curve[k].list, y.curve[k].list);
      note(c + " water", x.curve[k].water, y.curve[k].water);
      note(c + " inflow", x.curve[k].inflow, y.curve[k].inflow);
      note(c + " draftF", x.curve[k].dF, y.curve[k].dF);
      note(c + " draftA", x.curve[k].dA, y.curve[k].dA);
    }
  }
  return d;
}

}  // namespace sinksim::sweep

namespace sinksim::io {

namespace {

constexpr const char* kFields[] = {"t", "zO", "pitch", "roll", "vz", "wth", "wph", "inflow", "Vb", "hullTop", "vol", "level", "cx", "cy", "cz"};

// column f of row-major records: scalars 0..9 one value per record, per-node fields nodes values per record
std::vector<Real> column(const std::vector<Real>& rec, int nrec, Index nodes, int f) {
  const size_t rs = static_cast<size_t>(10 + 5 * nodes);
  std::vector<Real> out;
  for (int i = 0; i < nrec; ++i) {
    const Real* r = rec.data() + static_cast<size_t>(i) * rs;
    if (f < 10) out.push_back(r[f]);
    else out.insert(out.end(), r + 10 + static_cast<size_t>(f - 10) * nodes, r + 10 + static_cast<size_t>(f - 9) * nodes);
  }
  return out;
}

}  // namespace

void save_trace(const Trace& tr, const std::string& path) {
  using json::Value;
  Value doc = Value::object();
  doc["format"] = "sinksim.trace";
  doc["formatVersion"] = 1;
  doc["role"] = tr.role;
  doc["ship"] = tr.ship;
  doc["sim"] = tr.sim;
  Value meta = Value::object();
  meta["title"] = tr.title;
  meta["producer"] = tr.producer;
  meta["engine"] = tr.engine;
  doc["meta"] = meta;
  doc["dt"] = tr.dt;
  doc["nodes"] = tr.nodes;
  doc["denseSteps"] = tr.denseSteps;
  doc["sampleEvery"] = tr.sampleEvery;
  doc["steps"] = static_cast<long long>(tr.steps);
  doc["tEnd"] = tr.tEnd;
  doc["foundered"] = tr.foundered;
  doc["founderT"] = tr.founderT;
  Value ev = Value::array();
  for (const TraceEvent& e : tr.events) {
    Value x = Value::object();
    x["t"] = e.t;
    x["id"] = e.id;
    x["label"] = e.label;
    ev.push_back(x);
  }
  doc["events"] = ev;
  Value hash = Value::object();
  Fnv1a64 chain;
  const size_t rs = static_cast<size_t>(10 + 5 * tr.nodes);
  const std::pair<const char*, const std::vector<Real>*> parts[2] = {{"dense", &tr.dense}, {"samples", &tr.samples}};
  for (const auto& [name, recs] : parts) {
    const int n = static_cast<int>(recs->size() / rs);
    Value v = Value::object();
    v["n"] = n;
    for (int f = 0; f < 15; ++f) {
      const std::vector<Real> col = column(*recs, n, tr.nodes, f);
      v[kFields[f]] = Value(col);
      Fnv1a64 one;
      one.update_f64(col.data(), col.size());
      chain.update_f64(col.data(), col.size());
      hash[std::string(name) + "." + kFields[f]] = one.hex();
    }
    doc[name] = v;
  }
  hash["all"] = chain.hex();
  doc["hash"] = hash;
  json::write_file(path, doc, 2);
}

}  // namespace sinksim::io
