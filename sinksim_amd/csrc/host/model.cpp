// Compiled ship and simulation files (docs/FORMATS.md "sinksim.compiled-ship" / "sinksim.compiled-sim", version 1).
//
// A FieldReader pulls typed arrays and scalars out of one parsed document, checks lengths and integer ranges, and keeps
// the little-endian byte image of every array the format hashes. The format's hash order is data here (kShipHashed,
// kSimHashed): once a file is read, each listed array's FNV-1a 64 must equal the file's "hash" entry for it, and the
// chain over all of them in that order its "all" entry. So the loaders may read fields in whatever order their range
// checks need; a file whose numbers did not survive the JSON round trip bit for bit is refused.
#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>

#include "json.hpp"

namespace sinksim {

// ------------------------------------------------------------------------------------------- FNV-1a 64

void Fnv1a64::update(const void* data, size_t n) {
  const auto* p = static_cast<const unsigned char*>(data);
  for (size_t i = 0; i < n; ++i) h_ = (h_ ^ p[i]) * 1099511628211ULL;
}

// numbers enter as their little-endian bytes; -0 as +0
void Fnv1a64::update_f64(const double* v, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    uint64_t bits = 0;
    const double d = v[i] == 0 ? 0.0 : v[i];
    std::memcpy(&bits, &d, sizeof bits);
    unsigned char b[8];
    for (int k = 0; k < 8; ++k) b[k] = static_cast<unsigned char>(bits >> (8 * k));
    update(b, 8);
  }
}

void Fnv1a64::update_i32(const int32_t* v, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    unsigned char b[4];
    for (int k = 0; k < 4; ++k) b[k] = static_cast<unsigned char>(static_cast<uint32_t>(v[i]) >> (8 * k));
    update(b, 4);
  }
}

std::string Fnv1a64::hex() const {
  char buf[17];
  std::snprintf(buf, sizeof buf, "%016llx", static_cast<unsigned long long>(h_));
  return buf;
}

// ------------------------------------------------------------------------------------------- views for the C ABI

ss_ship CompiledShip::view() const {
  ss_ship s{};
  s.ncol = columns(); s.x = x.data(); s.y = y.data(); s.dxdy = dxdy.data(); s.zlo = zlo.data(); s.zhi = zhi.data();
  s.nseg = segments(); s.segCol = segCol.data(); s.segA = segA.data(); s.segE = segE.data();
  s.nnode = nodes(); s.mu = mu.data(); s.vmax = vmax.data(); s.amin = amin.data(); s.aov = aov.data();
  s.segStart = segStart.data(); s.segCount = segCount.data();
  return s;
}

ss_sim CompiledSim::view() const {
  ss_sim m{};
  m.nnode = nodes(); m.rho = rho; m.g = g; m.dt = dt;
  m.ngroup = groups(); m.gNodeStart = gNodeStart.data(); m.gNodeCount = gNodeCount.data(); m.gNodes = gNodes.data(); m.gMode = gMode.data();
  m.gpx = gpx.data(); m.gpy = gpy.data(); m.gpz = gpz.data(); m.gMargin = gMargin;
  m.nconn = connections(); m.ca = ca.data(); m.cb = cb.data(); m.law = cLaw.data(); m.cx = cx.data(); m.cy = cy.data(); m.cz = cz.data();
  m.area = cArea.data(); m.coef = cCoef.data(); m.kind = cKind.data(); m.idx = cIdx.data(); m.en = cEn.data(); m.tOn = cTOn.data();
  m.skipGroup = cSkipGroup.data(); m.monitor = cMonitor.data();
  m.nmonitor = monitors(); m.monitorThreshold = monitorThreshold;
  m.nmark = marks(); m.mkx = mkx.data(); m.mky = mky.data(); m.mkz = mkz.data(); m.mkWhen = mkWhen.data();
  std::copy(founder, founder + 4, m.founder);
  std::copy(body, body + 16, m.body);
  std::copy(readout, readout + 2, m.readout);
  std::copy(state0, state0 + 7, m.state0);
  m.vol0 = vol0.data(); m.level0 = level0.data(); m.cx0 = cx0.data(); m.cy0 = cy0.data(); m.cz0 = cz0.data();
  return m;
}

Index CompiledSim::kindIndex(const std::string& name) const {
  const auto it = std::find(kindNames.begin(), kindNames.end(), name);
  return it == kindNames.end() ? -1 : static_cast<Index>(it - kindNames.begin());
}

// engine/src/simulation.cpp:225-241: monitors and marks by their ids and labels, the founder event by fixed names
std::string event_id(const CompiledSim& sim, int kind, int id) {
  switch (kind) {
    case 0: return sim.monitorId.at(static_cast<size_t>(id));
    case 1: return sim.markId.at(static_cast<size_t>(id));
    default: return "founder";
  }
}

std::string event_label(const CompiledSim& sim, int kind, int id) {
  switch (kind) {
    case 0: return sim.monitorLabel.at(static_cast<size_t>(id));
    case 1: return sim.markLabel.at(static_cast<size_t>(id));
    default: return "Foundered";
  }
}

namespace io {
namespace {

using json::Value;

// The bytes the format hashes for an array: little-endian doubles (-0 as +0), int32 and uint8 values.
std::vector<unsigned char> image(const std::vector<double>& v) {
  std::vector<unsigned char> b(v.size() * 8);
  for (size_t i = 0; i < v.size(); ++i) {
    uint64_t bits = 0;
    const double d = v[i] == 0 ? 0.0 : v[i];
    std::memcpy(&bits, &d, sizeof bits);
    for (int k = 0; k < 8; ++k) b[8 * i + static_cast<size_t>(k)] = static_cast<unsigned char>(bits >> (8 * k));
  }
  return b;
}
std::vector<unsigned char> image(const std::vector<int32_t>& v) {
  std::vector<unsigned char> b(v.size() * 4);
  for (size_t i = 0; i < v.size(); ++i)
    for (int k = 0; k < 4; ++k) b[4 * i + static_cast<size_t>(k)] = static_cast<unsigned char>(static_cast<uint32_t>(v[i]) >> (8 * k));
  return b;
}
std::vector<unsigned char> image(const std::vector<uint8_t>& v) { return std::vector<unsigned char>(v.begin(), v.end()); }

// The arrays each format hashes, in the order of its chained hash (docs/FORMATS.md).
const char* const kShipHashed[] = {"columns.x",   "columns.y",   "columns.dxdy",   "columns.zlo",    "columns.zhi",  "columns.zone",
                                   "columns.side", "nodes.mu",    "nodes.vmax",     "nodes.amin",     "nodes.aov",    "nodes.segStart",
                                   "nodes.segCount", "segments.col", "segments.a", "segments.e"};
const char* const kSimHashed[] = {
    "groups.nodeStart", "groups.nodeCount", "groups.nodes", "groups.mode", "groups.px", "groups.py", "groups.pz",
    "connections.a", "connections.b", "connections.type", "connections.x", "connections.y", "connections.z", "connections.area",
    "connections.coef", "connections.kind", "connections.idx", "connections.en", "connections.tOn", "connections.skipGroup",
    "connections.monitor", "marks.x", "marks.y", "marks.z", "marks.when", "state0.vol", "state0.level", "state0.cx", "state0.cy",
    "state0.cz"};

class FieldReader {
 public:
  FieldReader(const std::string& file, const char* format, long long version) : file_(file), doc_(json::read_file(file)) {
    const Value* f = doc_.find("format");
    if (!f || !f->is_string() || f->as_string() != format) fail(std::string("format is not ") + format);
    const long long v = json::integer(doc_, "formatVersion", file_);
    if (v != version) fail("formatVersion " + std::to_string(v) + " is not supported (this engine reads " + std::to_string(version) + ")");
  }

  [[noreturn]] void fail(const std::string& what) const { throw std::runtime_error(file_ + ": " + what); }

  const Value& doc() const { return doc_; }
  const Value& node(const std::string& path) const {
    const size_t dot = path.find('.');
    if (dot == std::string::npos) return json::member(doc_, path.c_str(), file_);
    return json::member(json::member(doc_, path.substr(0, dot).c_str(), file_), path.substr(dot + 1).c_str(), file_);
  }
  std::string text(const char* key) const { return json::string(doc_, key, file_); }
  double real(const std::string& path) const {
    const Value& v = node(path);
    if (!v.is_number()) fail(path + " is not a number");
    return v.as_double();
  }
  size_t count(const char* section) const {
    const long long n = json::integer(json::member(doc_, section, file_), "n", file_);
    if (n < 0) fail(std::string(section) + ".n is negative");
    return static_cast<size_t>(n);
  }

  // Arrays: `want` is the length the file must give (npos: any). Integers must be integral and in [lo, hi].
  static constexpr size_t kAny = static_cast<size_t>(-1);
  std::vector<double> reals(const std::string& path, size_t want = kAny) {
    std::vector<double> v = numbers(path, want);
    ledger_[path] = image(v);
    return v;
  }
  std::vector<int32_t> ints(const std::string& path, size_t want, long long lo, long long hi) {
    const std::vector<double> d = numbers(path, want);
    std::vector<int32_t> v(d.size());
    for (size_t i = 0; i < d.size(); ++i) {
      if (!(d[i] == std::floor(d[i]) && d[i] >= static_cast<double>(lo) && d[i] <= static_cast<double>(hi)))
        fail(path + "[" + std::to_string(i) + "] = " + json::format_number(d[i]) + " is not an integer in [" + std::to_string(lo) + ", " +
             std::to_string(hi) + "]");
      v[i] = static_cast<int32_t>(d[i]);
    }
    ledger_[path] = image(v);
    return v;
  }
  std::vector<uint8_t> small(const std::string& path, size_t want) {
    const std::vector<int32_t> w = ints(path, want, 0, 255);
    std::vector<uint8_t> v(w.begin(), w.end());
    ledger_[path] = image(v);
    return v;
  }
  std::vector<std::string> texts(const std::string& path, size_t want = kAny) const {
    const Value& a = node(path);
    if (!a.is_array()) fail(path + " must be an array");
    if (want != kAny && a.size() != want) fail(path + " has " + std::to_string(a.size()) + " entries, expected " + std::to_string(want));
    std::vector<std::string> v;
    for (const Value& s : a.items()) {
      if (!s.is_string()) fail(path + " must contain strings");
      v.push_back(s.as_string());
    }
    return v;
  }

  // Every array of `order` against the file's hashes, then the chain (the hash of their bytes one after the other);
  // returns the chain's value.
  template <size_t N>
  std::string verify(const char* const (&order)[N]) const {
    const Value& H = json::member(doc_, "hash", file_);
    Fnv1a64 chain;
    for (const char* name : order) {
      const auto it = ledger_.find(name);
      if (it == ledger_.end()) fail(std::string("hashed array ") + name + " was not read");
      Fnv1a64 one;
      one.update(it->second.data(), it->second.size());
      chain.update(it->second.data(), it->second.size());
      const std::string want = json::string(H, name, file_ + " hash");
      if (one.hex() != want)
        fail(std::string("hash mismatch on ") + name + " (" + one.hex() + " vs " + want + "): the JSON round trip was not exact");
    }
    const std::string all = json::string(H, "all", file_ + " hash");
    if (chain.hex() != all) fail("chained hash mismatch (" + chain.hex() + " vs " + all + ")");
    return chain.hex();
  }

 private:
  std::vector<double> numbers(const std::string& path, size_t want) const {
    const Value& a = node(path);
    if (!a.is_array()) fail(path + " must be an array");
    if (want != kAny && a.size() != want) fail(path + " has " + std::to_string(a.size()) + " entries, expected " + std::to_string(want));
    std::vector<double> v(a.size());
    for (size_t i = 0; i < a.size(); ++i) {
      if (!a[i].is_number()) fail(path + "[" + std::to_string(i) + "] is not a number");
      v[i] = a[i].as_double();
    }
    return v;
  }

  std::string file_;
  Value doc_;
  std::map<std::string, std::vector<unsigned char>> ledger_;  // the bytes of every hashed array read
};

}  // namespace

std::shared_ptr<CompiledShip> load_ship(const std::string& file) {
  FieldReader r(file, "sinksim.compiled-ship", 1);
  auto S = std::make_shared<CompiledShip>();
  S->id = r.text("id");
  if (const Value* meta = r.doc().find("meta")) {
    S->title = json::string_or(*meta, "title", "");
    S->producer = json::string_or(*meta, "generator", "") + " on " + json::string_or(*meta, "engine", "");
  }
  const size_t nc = r.count("columns"), nn = r.count("nodes"), ns = r.count("segments");
  const long long ncl = static_cast<long long>(nc), nsl = static_cast<long long>(ns);
  S->x = r.reals("columns.x", nc); S->y = r.reals("columns.y", nc); S->dxdy = r.reals("columns.dxdy", nc);
  S->zlo = r.reals("columns.zlo", nc); S->zhi = r.reals("columns.zhi", nc);
  S->zone = r.ints("columns.zone", nc, 0, 1 << 30); S->side = r.ints("columns.side", nc, 0, 1);
  S->nodeLabel = r.texts("nodes.label", nn);
  S->mu = r.reals("nodes.mu", nn); S->vmax = r.reals("nodes.vmax", nn); S->amin = r.reals("nodes.amin", nn); S->aov = r.reals("nodes.aov", nn);
  S->segStart = r.ints("nodes.segStart", nn, 0, nsl); S->segCount = r.ints("nodes.segCount", nn, 0, nsl);
  S->segCol = r.ints("segments.col", ns, 0, ncl - 1); S->segA = r.reals("segments.a", ns); S->segE = r.reals("segments.e", ns);
  for (size_t n = 0; n < nn; ++n)
    if (static_cast<size_t>(S->segStart[n]) + static_cast<size_t>(S->segCount[n]) > ns)
      r.fail("node " + std::to_string(n) + " segment range exceeds the segment table");
  S->hashAll = r.verify(kShipHashed);
  return S;
}

CompiledSim load_sim(const std::string& file) {
  FieldReader r(file, "sinksim.compiled-sim", 1);
  CompiledSim M;
  M.id = r.text("id");
  M.shipId = r.text("ship");
  M.shipHash = r.text("shipHash");
  if (const Value* meta = r.doc().find("meta")) {
    M.title = json::string_or(*meta, "title", "");
    M.producer = json::string_or(*meta, "generator", "") + " on " + json::string_or(*meta, "engine", "");
  }
  M.rho = r.real("constants.rho");
  M.g = r.real("constants.g");
  M.dt = r.real("scheme.dt");
  const char* stateKeys[7] = {"t", "zO", "pitch", "roll", "vz", "wth", "wph"};
  for (int k = 0; k < 7; ++k) M.state0[k] = r.real(std::string("state0.") + stateKeys[k]);
  M.vol0 = r.reals("state0.vol");
  const size_t nn = M.vol0.size();
  const long long nnl = static_cast<long long>(nn);
  M.level0 = r.reals("state0.level", nn); M.cx0 = r.reals("state0.cx", nn); M.cy0 = r.reals("state0.cy", nn); M.cz0 = r.reals("state0.cz", nn);

  const size_t ng = r.count("groups");
  M.gNodeCount = r.ints("groups.nodeCount", ng, 1, 1 << 20);
  M.gNodes = r.ints("groups.nodes", FieldReader::kAny, 0, nnl - 1);
  M.gNodeStart = r.ints("groups.nodeStart", ng, 0, static_cast<long long>(M.gNodes.size()));
  M.gMode = r.small("groups.mode", ng);
  M.gpx = r.reals("groups.px", ng); M.gpy = r.reals("groups.py", ng); M.gpz = r.reals("groups.pz", ng);
  M.gMargin = r.real("groups.margin");
  for (size_t g = 0; g < ng; ++g)
    if (static_cast<size_t>(M.gNodeStart[g]) + static_cast<size_t>(M.gNodeCount[g]) > M.gNodes.size())
      r.fail("group " + std::to_string(g) + " member range exceeds the member table");

  const size_t nc = r.count("connections");
  M.ca = r.ints("connections.a", nc, -1, nnl - 1); M.cb = r.ints("connections.b", nc, 0, nnl - 1);
  M.cLaw = r.small("connections.type", nc);
  M.cx = r.reals("connections.x", nc); M.cy = r.reals("connections.y", nc); M.cz = r.reals("connections.z", nc);
  M.cArea = r.reals("connections.area", nc); M.cCoef = r.reals("connections.coef", nc);
  M.cKind = r.small("connections.kind", nc);
  M.cIdx = r.ints("connections.idx", nc, -1, 1 << 30);
  M.cEn = r.small("connections.en", nc);
  M.cTOn = r.reals("connections.tOn", nc);
  M.cSkipGroup = r.ints("connections.skipGroup", nc, -1, static_cast<long long>(ng) - 1);
  M.cMonitor = r.ints("connections.monitor", nc, -1, 1 << 30);
  if (json::member(r.doc(), "connections", file).contains("kindNames")) M.kindNames = r.texts("connections.kindNames");

  M.monitorThreshold = r.real("monitors.threshold");
  M.monitorId = r.texts("monitors.id");
  M.monitorLabel = r.texts("monitors.label", M.monitorId.size());
  for (size_t c = 0; c < nc; ++c)
    if (M.cMonitor[c] >= static_cast<Index>(M.monitorId.size())) r.fail("connection " + std::to_string(c) + " refers to a monitor that does not exist");

  M.markId = r.texts("marks.id");
  const size_t nm = M.markId.size();
  M.markLabel = r.texts("marks.label", nm);
  M.mkx = r.reals("marks.x", nm); M.mky = r.reals("marks.y", nm); M.mkz = r.reals("marks.z", nm);
  M.mkWhen = r.small("marks.when", nm);

  const char* founderKeys[4] = {"hullTopBelow", "pitchAbsAbove", "buoyancyBelow", "minT"};
  for (int k = 0; k < 4; ++k) M.founder[k] = r.real(std::string("founder.") + founderKeys[k]);
  // body: ms KG g0[3] XO YO mh0 Ith0 Iph0 cDz cDth cDph qz qth qph
  M.body[0] = r.real("body.ms"); M.body[1] = r.real("body.KG");
  const Value& g0 = r.node("body.g0");
  if (!g0.is_array() || g0.size() != 3 || !g0[0].is_number() || !g0[1].is_number() || !g0[2].is_number()) r.fail("body.g0 must hold 3 numbers");
  for (int k = 0; k < 3; ++k) M.body[2 + k] = g0[static_cast<size_t>(k)].as_double();
  const char* bodyKeys[11] = {"XO", "YO", "mh0", "Ith0", "Iph0", "cDz", "cDth", "cDph", "qz", "qth", "qph"};
  for (int k = 0; k < 11; ++k) M.body[5 + k] = r.real(std::string("body.") + bodyKeys[k]);
  M.readout[0] = r.real("readout.xFP");
  M.readout[1] = r.real("readout.xAP");
  M.hashAll = r.verify(kSimHashed);
  return M;
}

void check_pairing(const CompiledShip& ship, const CompiledSim& sim) {
  if (sim.shipId != ship.id) throw std::runtime_error("simulation " + sim.id + " was compiled for ship " + sim.shipId + ", not " + ship.id);
  if (sim.shipHash != ship.hashAll)
    throw std::runtime_error("simulation " + sim.id + " was compiled against a different build of ship " + ship.id + " (hash " + sim.shipHash +
                             " vs " + ship.hashAll + ")");
  if (sim.nodes() != ship.nodes())
    throw std::runtime_error("simulation " + sim.id + " has " + std::to_string(sim.nodes()) + " nodes, ship " + ship.id + " has " +
                             std::to_string(ship.nodes()));
}

}  // namespace io
}  // namespace sinksim
