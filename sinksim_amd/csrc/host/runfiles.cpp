#include "runfiles.hpp"

#include <cmath>
#include <cstdio>
#include <functional>
#include <stdexcept>

#include "json.hpp"

namespace sinksim::files {

using json::Value;

namespace {

[[noreturn]] void refuse(const std::string& where, const std::string& what) { throw std::runtime_error(where + ": " + what); }

const Value& array_at(const Value& obj, const char* key, const std::string& where) {
  const Value& v = json::member(obj, key, where);
  if (!v.is_array()) refuse(where, std::string(key) + " must be an array");
  return v;
}

Byte kind_named(const CompiledSim& sim, const std::string& name, const std::string& where) {
  const Index k = sim.kindIndex(name);
  if (k < 0) refuse(where, "unknown connection kind \"" + name + "\"");
  return static_cast<Byte>(k);
}

// the fields of a curve record, as the results file names its columns
const char* const kCurveKeys[7] = {"t", "trim", "list", "water", "inflow", "draftF", "draftA"};

}  // namespace

std::string minutes(double seconds) {
  char buf[48];
  std::snprintf(buf, sizeof buf, "%.2f min", seconds / 60.0);
  return buf;
}

// ------------------------------------------------------------------------------------------- sweeps

SweepSpec read_sweep(const std::string& path, const CompiledSim& sim) {
  const Value doc = json::read_file(path);
  if (json::string_or(doc, "format", "") != "sinksim.sweep") refuse(path, "format is not sinksim.sweep");
  if (json::integer(doc, "formatVersion", path) != 1) refuse(path, "unsupported sweep formatVersion");
  SweepSpec sw;
  sw.ship = json::string_or(doc, "ship", "");
  sw.sim = json::string_or(doc, "sim", "");
  sw.tMax = json::number_or(doc, "tMax", 5 * 3600);
  sw.every = json::number_or(doc, "every", 30);
  if (const Value* s = doc.find("stopWhenStable")) sw.stopWhenStable = s->is_bool() && s->as_bool();
  const Value& list = array_at(doc, "instances", path);
  for (size_t k = 0; k < list.size(); ++k) {
    const Value& j = list[k];
    Variant v;
    v.id = json::string_or(j, "id", "instance" + std::to_string(k));
    const std::string where = path + " instance " + v.id;
    if (const Value* s = j.find("scale")) {
      if (!s->is_object()) refuse(where, "scale must be an object of kind -> factor");
      for (const auto& [name, factor] : s->members()) {
        if (!factor.is_number()) refuse(where, "scale factor for " + name + " is not a number");
        v.scales.emplace_back(kind_named(sim, name, where), factor.as_double());
      }
    }
    if (const Value* c = j.find("connections")) {
      if (!c->is_array()) refuse(where, "connections must be an array");
      for (const Value& o : c->items()) {
        Override ov;
        if (o.contains("conn")) {
          ov.conn = static_cast<Index>(json::integer(o, "conn", where));
        } else {
          ov.kind = kind_named(sim, json::string(o, "kind", where), where);
          if (o.contains("idx")) ov.idx = static_cast<Index>(json::integer(o, "idx", where));
        }
        if (o.contains("enabled")) ov.enabled = json::boolean(o, "enabled", where);
        if (o.contains("tOn")) ov.tOn = json::number(o, "tOn", where);
        if (o.contains("area")) ov.area = json::number(o, "area", where);
        v.overrides.push_back(ov);
      }
    }
    sw.instances.push_back(std::move(v));
  }
  return sw;
}

Tables connection_tables(const SweepSpec& sw, const CompiledSim& sim) {
  const size_t nc = static_cast<size_t>(sim.connections()), count = sw.instances.size();
  Tables t;
  t.en.reserve(count * nc); t.area.reserve(count * nc); t.tOn.reserve(count * nc);
  for (const Variant& v : sw.instances) {
    const size_t row = t.en.size();
    t.en.insert(t.en.end(), sim.cEn.begin(), sim.cEn.end());
    t.area.insert(t.area.end(), sim.cArea.begin(), sim.cArea.end());
    t.tOn.insert(t.tOn.end(), sim.cTOn.begin(), sim.cTOn.end());
    for (const auto& [kind, factor] : v.scales)
      for (size_t c = 0; c < nc; ++c)
        if (sim.cKind[c] == kind) t.area[row + c] = sim.cArea[c] * factor;
    for (const Override& o : v.overrides)
      for (size_t c = 0; c < nc; ++c) {
        const bool picked = o.conn >= 0 ? static_cast<Index>(c) == o.conn : sim.cKind[c] == o.kind && (o.idx < 0 || sim.cIdx[c] == o.idx);
        if (!picked) continue;
        if (o.enabled) t.en[row + c] = *o.enabled ? 1 : 0;
        if (o.area) t.area[row + c] = *o.area;
        if (o.tOn) t.tOn[row + c] = *o.tOn;
      }
  }
  return t;
}

// ------------------------------------------------------------------------------------------- results

void write_results(const ResultsFile& r, const std::string& path) {
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
  for (const Outcome& o : r.instances) {
    Value j = Value::object();
    j["id"] = o.id;
    j["foundered"] = o.foundered;
    j["founderT"] = o.founderT;
    j["tEnd"] = o.tEnd;
    j["steps"] = static_cast<long long>(o.steps);
    Value f = Value::object();
    f["trimDeg"] = o.final[1]; f["listDeg"] = o.final[2]; f["waterT"] = o.final[3];
    f["draftF"] = o.final[5]; f["draftA"] = o.final[6]; f["inflowTpm"] = o.final[4];
    j["final"] = f;
    Value ev = Value::array();
    for (const Outcome::Event& e : o.events) {
      Value x = Value::object();
      x["t"] = e.t; x["id"] = e.id; x["label"] = e.label;
      ev.push_back(x);
    }
    j["events"] = ev;
    Value c = Value::object();
    for (size_t k = 0; k < 7; ++k) {
      std::vector<double> col(o.curve.size());
      for (size_t i = 0; i < o.curve.size(); ++i) col[i] = o.curve[i][k];
      c[kCurveKeys[k]] = Value(col);
    }
    j["curve"] = c;
    list.push_back(j);
  }
  doc["instances"] = list;
  json::write_file(path, doc, 2);
}

ResultsFile read_results(const std::string& path) {
  const Value doc = json::read_file(path);
  if (json::string_or(doc, "format", "") != "sinksim.batch-results") refuse(path, "format is not sinksim.batch-results");
  ResultsFile r;
  r.ship = json::string_or(doc, "ship", "");
  r.sim = json::string_or(doc, "sim", "");
  r.numerics = json::string_or(doc, "numerics", "");
  r.engine = json::string_or(doc, "engine", "");
  r.tMax = json::number_or(doc, "tMax", 0);
  r.every = json::number_or(doc, "every", 0);
  for (const Value& j : array_at(doc, "instances", path).items()) {
    Outcome o;
    o.id = json::string_or(j, "id", "");
    o.foundered = json::boolean(j, "foundered", path);
    o.founderT = json::number(j, "founderT", path);
    o.tEnd = json::number(j, "tEnd", path);
    o.steps = json::integer(j, "steps", path);
    const Value& f = json::member(j, "final", path);
    o.final = Curve{o.tEnd, json::number(f, "trimDeg", path), json::number(f, "listDeg", path), json::number(f, "waterT", path),
                    json::number(f, "inflowTpm", path), json::number(f, "draftF", path), json::number(f, "draftA", path)};
    for (const Value& e : array_at(j, "events", path).items())
      o.events.push_back(Outcome::Event{json::number(e, "t", path), json::string(e, "id", path), json::string_or(e, "label", "")});
    const Value& c = json::member(j, "curve", path);
    std::array<const Value*, 7> col{};
    for (size_t k = 0; k < 7; ++k) col[k] = &array_at(c, kCurveKeys[k], path);
    for (size_t k = 1; k < 7; ++k)
      if (col[k]->size() != col[0]->size()) refuse(path, std::string("curve.") + kCurveKeys[k] + " differs in length from curve.t");
    o.curve.resize(col[0]->size());
    for (size_t i = 0; i < o.curve.size(); ++i)
      for (size_t k = 0; k < 7; ++k) {
        const Value& x = (*col[k])[i];
        if (!x.is_number()) refuse(path, std::string("curve.") + kCurveKeys[k] + " contains a non-number");
        o.curve[i][k] = x.as_double();
      }
    r.instances.push_back(std::move(o));
  }
  return r;
}

Difference compare(const ResultsFile& a, const ResultsFile& b) {
  Difference d;
  const auto check = [&](const std::function<std::string()>& where, double x, double y) {
    ++d.compared;
    if (x == y || (std::isnan(x) && std::isnan(y))) return;
    ++d.differing;
    d.largest = std::fmax(d.largest, std::fabs(x - y));
    if (d.first.empty()) d.first = where() + ": " + json::format_number(x) + " vs " + json::format_number(y);
  };
  if (a.instances.size() != b.instances.size()) {
    d.differing = 1;
    d.first = "instance count " + std::to_string(a.instances.size()) + " vs " + std::to_string(b.instances.size());
    return d;
  }
  // the per-instance scalars and final readouts that take part, by name
  struct Field {
    const char* name;
    double (*get)(const Outcome&);
  };
  static const Field kFields[] = {
      {"foundered", [](const Outcome& o) { return o.foundered ? 1.0 : 0.0; }},
      {"founderT", [](const Outcome& o) { return o.founderT; }},
      {"tEnd", [](const Outcome& o) { return o.tEnd; }},
      {"steps", [](const Outcome& o) { return static_cast<double>(o.steps); }},
      {"trim", [](const Outcome& o) { return o.final[1]; }},
      {"list", [](const Outcome& o) { return o.final[2]; }},
      {"water", [](const Outcome& o) { return o.final[3]; }},
      {"draftF", [](const Outcome& o) { return o.final[5]; }},
      {"draftA", [](const Outcome& o) { return o.final[6]; }},
  };
  for (size_t i = 0; i < a.instances.size(); ++i) {
    const Outcome& x = a.instances[i];
    const Outcome& y = b.instances[i];
    const std::string who = "instance " + x.id;
    for (const Field& f : kFields) check([&] { return who + " " + f.name; }, f.get(x), f.get(y));
    check([&] { return who + " events"; }, static_cast<double>(x.events.size()), static_cast<double>(y.events.size()));
    for (size_t k = 0; k < std::min(x.events.size(), y.events.size()); ++k) {
      check([&] { return who + " event " + x.events[k].id + " t"; }, x.events[k].t, y.events[k].t);
      if (x.events[k].id != y.events[k].id) {
        ++d.differing;
        if (d.first.empty()) d.first = who + " event id " + x.events[k].id + " vs " + y.events[k].id;
      }
    }
    check([&] { return who + " curve length"; }, static_cast<double>(x.curve.size()), static_cast<double>(y.curve.size()));
    for (size_t k = 0; k < std::min(x.curve.size(), y.curve.size()); ++k)
      for (size_t f = 0; f < 7; ++f)
        check([&] { return who + " curve[" + std::to_string(k) + "] " + kCurveKeys[f]; }, x.curve[k][f], y.curve[k][f]);
  }
  return d;
}

// ------------------------------------------------------------------------------------------- traces

void write_trace(const TraceFile& tr, const std::string& path) {
  // a record's fields as the file's columns: ten scalars, then five per-node blocks
  static const char* const kScalars[10] = {"t", "zO", "pitch", "roll", "vz", "wth", "wph", "inflow", "Vb", "hullTop"};
  static const char* const kPerNode[5] = {"vol", "level", "cx", "cy", "cz"};
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
  for (const Outcome::Event& e : tr.events) {
    Value x = Value::object();
    x["t"] = e.t; x["id"] = e.id; x["label"] = e.label;
    ev.push_back(x);
  }
  doc["events"] = ev;
  const size_t nn = static_cast<size_t>(tr.nodes), rs = 10 + 5 * nn;
  Value hash = Value::object();
  Fnv1a64 chain;
  for (const auto& [name, recs] : {std::pair<const char*, const std::vector<Real>*>{"dense", &tr.dense}, {"samples", &tr.samples}}) {
    const size_t n = recs->size() / rs;
    Value cols = Value::object();
    cols["n"] = static_cast<long long>(n);
    const auto column = [&](const std::string& field, size_t first, size_t width) {
      std::vector<double> v;
      v.reserve(n * width);
      for (size_t i = 0; i < n; ++i) v.insert(v.end(), recs->begin() + static_cast<std::ptrdiff_t>(i * rs + first),
                                              recs->begin() + static_cast<std::ptrdiff_t>(i * rs + first + width));
      Fnv1a64 one;
      one.update_f64(v.data(), v.size());
      chain.update_f64(v.data(), v.size());
      hash[std::string(name) + "." + field] = one.hex();
      cols[field] = Value(v);
    };
    for (size_t k = 0; k < 10; ++k) column(kScalars[k], k, 1);
    for (size_t k = 0; k < 5; ++k) column(kPerNode[k], 10 + k * nn, nn);
    doc[name] = cols;
  }
  hash["all"] = chain.hex();
  doc["hash"] = hash;
  json::write_file(path, doc, 2);
}

}  // namespace sinksim::files
