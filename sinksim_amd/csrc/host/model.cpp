#include "model.hpp"

#include <cmath>
#include <cstring>
#include <stdexcept>

#include "json.hpp"

namespace sinksim {

void Fnv1a64::update(const void* data, size_t n) {
  const auto* p = static_cast<const unsigned char*>(data);
  for (size_t i = 0; i < n; ++i) {
    h_ ^= static_cast<uint64_t>(p[i]);
    h_ *= 1099511628211ULL;
  }
}

void Fnv1a64::update_f64(const double* v, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    double d = v[i];
    if (d == 0) d = 0.0;
    uint64_t bits;
    std::memcpy(&bits, &d, 8);
    unsigned char b[8];
    for (int k = 0; k < 8; ++k) b[k] = static_cast<unsigned char>(bits >> (8 * k));
    update(b, 8);
  }
}

void Fnv1a64::update_i32(const int32_t* v, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    const uint32_t bits = static_cast<uint32_t>(v[i]);
    unsigned char b[4];
    for (int k = 0; k < 4; ++k) b[k] = static_cast<unsigned char>(bits >> (8 * k));
    update(b, 4);
  }
}

std::string Fnv1a64::hex() const {
  static const char* digits = "0123456789abcdef";
  std::string s(16, '0');
  for (int i = 15; i >= 0; --i) s[static_cast<size_t>(i)] = digits[(h_ >> (4 * (15 - i))) & 0xfu];
  return s;
}

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
  for (int k = 0; k < 4; ++k) m.founder[k] = founder[k];
  for (int k = 0; k < 16; ++k) m.body[k] = body[k];
  m.readout[0] = readout[0]; m.readout[1] = readout[1];
  for (int k = 0; k < 7; ++k) m.state0[k] = state0[k];
  m.vol0 = vol0.data(); m.level0 = level0.data(); m.cx0 = cx0.data(); m.cy0 = cy0.data(); m.cz0 = cz0.data();
  return m;
}

Index CompiledSim::kindIndex(const std::string& name) const {
  for (size_t i = 0; i < kindNames.size(); ++i)
    if (kindNames[i] == name) return static_cast<Index>(i);
  return -1;
}

std::string event_id(const CompiledSim& sim, int kind, int id) {
  if (kind == 0) return sim.monitorId.at(static_cast<size_t>(id));
  if (kind == 1) return sim.markId.at(static_cast<size_t>(id));
  return "founder";
}

std::string event_label(const CompiledSim& sim, int kind, int id) {
  if (kind == 0) return sim.monitorLabel.at(static_cast<size_t>(id));
  if (kind == 1) return sim.markLabel.at(static_cast<size_t>(id));
  return "Foundered";
}

namespace io {
namespace {

using json::Value;

[[noreturn]] void bad(const std::string& file, const std::string& what) { throw std::runtime_error(file + ": " + what); }

const Value& at(const Value& obj, const char* key, const std::string& file) { return json::member(obj, key, file); }

std::vector<Real> numbers(const Value& v, const std::string& file, const std::string& what) {
  if (!v.is_array()) bad(file, what + " must be an array");
  std::vector<Real> out;
  out.reserve(v.size());
  for (size_t i = 0; i < v.size(); ++i) {
    if (!v[i].is_number()) bad(file, what + "[" + std::to_string(i) + "] is not a number");
    out.push_back(v[i].as_double());
  }
  return out;
}

std::vector<Index> indices(const Value& v, const std::string& file, const std::string& what, Index lo, Index hi) {
  const std::vector<Real> d = numbers(v, file, what);
  std::vector<Index> out;
  out.reserve(d.size());
  for (size_t i = 0; i < d.size(); ++i) {
    if (d[i] != std::floor(d[i]) || d[i] < lo || d[i] > hi)
      bad(file, what + "[" + std::to_string(i) + "] = " + json::format_number(d[i]) + " is not an integer in [" + std::to_string(lo) + ", " +
                    std::to_string(hi) + "]");
    out.push_back(static_cast<Index>(d[i]));
  }
  return out;
}

std::vector<Byte> bytes(const Value& v, const std::string& file, const std::string& what) {
  const std::vector<Index> d = indices(v, file, what, 0, 255);
  return std::vector<Byte>(d.begin(), d.end());
}

std::vector<std::string> strings(const Value& v, const std::string& file, const std::string& what) {
  if (!v.is_array()) bad(file, what + " must be an array");
  std::vector<std::string> out;
  for (const Value& s : v.items()) {
    if (!s.is_string()) bad(file, what + " must contain strings");
    out.push_back(s.as_string());
  }
  return out;
}

void same_size(const std::string& file, const std::string& what, size_t got, size_t want) {
  if (got != want) bad(file, what + " has " + std::to_string(got) + " entries, expected " + std::to_string(want));
}

struct Hashed {
  const char* name;
  char kind;  // 'f' f64, 'i' i32, 'b' u8
  const void* data;
  size_t n;
};

void feed(Fnv1a64& h, const Hashed& a) {
  if (a.kind == 'f') h.update_f64(static_cast<const double*>(a.data), a.n);
  else if (a.kind == 'i') h.update_i32(static_cast<const int32_t*>(a.data), a.n);
  else h.update_u8(static_cast<const uint8_t*>(a.data), a.n);
}

std::string verify_hashes(const Value& doc, const std::string& file, const std::vector<Hashed>& arrays) {
  const Value& H = at(doc, "hash", file);
  Fnv1a64 chain;
  for (const Hashed& a : arrays) {
    Fnv1a64 one;
    feed(one, a);
    feed(chain, a);
    const std::string want = json::string(H, a.name, file + " hash");
    if (one.hex() != want) bad(file, std::string("hash mismatch on ") + a.name + " (" + one.hex() + " vs " + want + "): the JSON round trip was not exact");
  }
  const std::string want = json::string(H, "all", file + " hash");
  if (chain.hex() != want) bad(file, "chained hash mismatch (" + chain.hex() + " vs " + want + ")");
  return chain.hex();
}

void check_format(const Value& doc, const std::string& file, const char* format, int version) {
  if (json::string_or(doc, "format", "") != format) bad(file, std::string("format is not ") + format);
  const long long v = json::integer(doc, "formatVersion", file);
  if (v != version) bad(file, "formatVersion " + std::to_string(v) + " is not supported (engine reads " + std::to_string(version) + ")");
}

Real real_at(const Value& obj, const char* key, const std::string& file) { return json::number(obj, key, file); }

}  // namespace

std::shared_ptr<CompiledShip> load_ship(const std::string& file) {
  const Value doc = json::read_file(file);
  check_format(doc, file, "sinksim.compiled-ship", 1);
  auto S = std::make_shared<CompiledShip>();
  S->id = json::string(doc, "id", file);
  if (const Value* meta = doc.find("meta")) {
    S->title = json::string_or(*meta, "title", "");
    S->producer = json::string_or(*meta, "generator", "") + " on " + json::string_or(*meta, "engine", "");
  }
  const Value& C = at(doc, "columns", file);
  const auto nc = static_cast<size_t>(json::integer(C, "n", file));
  auto col = [&](const char* k, std::vector<Real>& dst) {
    dst = numbers(at(C, k, file), file, std::string("columns.") + k);
    same_size(file, std::string("columns.") + k, dst.size(), nc);
  };
  col("x", S->x); col("y", S->y); col("dxdy", S->dxdy); col("zlo", S->zlo); col("zhi", S->zhi);
  S->zone = indices(at(C, "zone", file), file, "columns.zone", 0, 1 << 30); same_size(file, "columns.zone", S->zone.size(), nc);
  S->side = indices(at(C, "side", file), file, "columns.side", 0, 1); same_size(file, "columns.side", S->side.size(), nc);
  const Value& N = at(doc, "nodes", file);
  const auto nn = static_cast<size_t>(json::integer(N, "n", file));
  S->nodeLabel = strings(at(N, "label", file), file, "nodes.label"); same_size(file, "nodes.label", S->nodeLabel.size(), nn);
  auto nod = [&](const char* k, std::vector<Real>& dst) {
    dst = numbers(at(N, k, file), file, std::string("nodes.") + k);
    same_size(file, std::string("nodes.") + k, dst.size(), nn);
  };
  nod("mu", S->mu); nod("vmax", S->vmax); nod("amin", S->amin); nod("aov", S->aov);
  const Value& G = at(doc, "segments", file);
  const auto ns = static_cast<size_t>(json::integer(G, "n", file));
  S->segStart = indices(at(N, "segStart", file), file, "nodes.segStart", 0, static_cast<Index>(ns)); same_size(file, "nodes.segStart", S->segStart.size(), nn);
  S->segCount = indices(at(N, "segCount", file), file, "nodes.segCount", 0, static_cast<Index>(ns)); same_size(file, "nodes.segCount", S->segCount.size(), nn);
  S->segCol = indices(at(G, "col", file), file, "segments.col", 0, static_cast<Index>(nc) - 1); same_size(file, "segments.col", S->segCol.size(), ns);
  S->segA = numbers(at(G, "a", file), file, "segments.a"); same_size(file, "segments.a", S->segA.size(), ns);
  S->segE = numbers(at(G, "e", file), file, "segments.e"); same_size(file, "segments.e", S->segE.size(), ns);
  for (size_t n = 0; n < nn; ++n)
    if (static_cast<size_t>(S->segStart[n]) + static_cast<size_t>(S->segCount[n]) > ns) bad(file, "node " + std::to_string(n) + " segment range exceeds the segment table");
  S->hashAll = verify_hashes(doc, file, {
      {"columns.x", 'f', S->x.data(), nc}, {"columns.y", 'f', S->y.data(), nc}, {"columns.dxdy", 'f', S->dxdy.data(), nc},
      {"columns.zlo", 'f', S->zlo.data(), nc}, {"columns.zhi", 'f', S->zhi.data(), nc},
      {"columns.zone", 'i', S->zone.data(), nc}, {"columns.side", 'i', S->side.data(), nc},
      {"nodes.mu", 'f', S->mu.data(), nn}, {"nodes.vmax", 'f', S->vmax.data(), nn}, {"nodes.amin", 'f', S->amin.data(), nn},
      {"nodes.aov", 'f', S->aov.data(), nn}, {"nodes.segStart", 'i', S->segStart.data(), nn}, {"nodes.segCount", 'i', S->segCount.data(), nn},
      {"segments.col", 'i', S->segCol.data(), ns}, {"segments.a", 'f', S->segA.data(), ns}, {"segments.e", 'f', S->segE.data(), ns},
  });
  return S;
}

CompiledSim load_sim(const std::string& file) {
  const Value doc = json::read_file(file);
  check_format(doc, file, "sinksim.compiled-sim", 1);
  CompiledSim M;
  M.id = json::string(doc, "id", file);
  M.shipId = json::string(doc, "ship", file);
  M.shipHash = json::string(doc, "shipHash", file);
  if (const Value* meta = doc.find("meta")) {
    M.title = json::string_or(*meta, "title", "");
    M.producer = json::string_or(*meta, "generator", "") + " on " + json::string_or(*meta, "engine", "");
  }
  const Value& K = at(doc, "constants", file);
  M.rho = real_at(K, "rho", file);
  M.g = real_at(K, "g", file);
  M.dt = real_at(at(doc, "scheme", file), "dt", file);
  const Value& st = at(doc, "state0", file);
  const char* sk[7] = {"t", "zO", "pitch", "roll", "vz", "wth", "wph"};
  for (int k = 0; k < 7; ++k) M.state0[k] = real_at(st, sk[k], file);
  M.vol0 = numbers(at(st, "vol", file), file, "state0.vol");
  const size_t nn = M.vol0.size();
  M.level0 = numbers(at(st, "level", file), file, "state0.level"); same_size(file, "state0.level", M.level0.size(), nn);
  M.cx0 = numbers(at(st, "cx", file), file, "state0.cx"); same_size(file, "state0.cx", M.cx0.size(), nn);
  M.cy0 = numbers(at(st, "cy", file), file, "state0.cy"); same_size(file, "state0.cy", M.cy0.size(), nn);
  M.cz0 = numbers(at(st, "cz", file), file, "state0.cz"); same_size(file, "state0.cz", M.cz0.size(), nn);

  const Value& G = at(doc, "groups", file);
  const auto ng = static_cast<size_t>(json::integer(G, "n", file));
  M.gNodeCount = indices(at(G, "nodeCount", file), file, "groups.nodeCount", 1, 1 << 20); same_size(file, "groups.nodeCount", M.gNodeCount.size(), ng);
  M.gNodes = indices(at(G, "nodes", file), file, "groups.nodes", 0, static_cast<Index>(nn) - 1);
  M.gNodeStart = indices(at(G, "nodeStart", file), file, "groups.nodeStart", 0, static_cast<Index>(M.gNodes.size()));
  same_size(file, "groups.nodeStart", M.gNodeStart.size(), ng);
  M.gMode = bytes(at(G, "mode", file), file, "groups.mode"); same_size(file, "groups.mode", M.gMode.size(), ng);
  M.gpx = numbers(at(G, "px", file), file, "groups.px"); same_size(file, "groups.px", M.gpx.size(), ng);
  M.gpy = numbers(at(G, "py", file), file, "groups.py"); same_size(file, "groups.py", M.gpy.size(), ng);
  M.gpz = numbers(at(G, "pz", file), file, "groups.pz"); same_size(file, "groups.pz", M.gpz.size(), ng);
  M.gMargin = real_at(G, "margin", file);
  for (size_t g = 0; g < ng; ++g)
    if (static_cast<size_t>(M.gNodeStart[g]) + static_cast<size_t>(M.gNodeCount[g]) > M.gNodes.size())
      bad(file, "group " + std::to_string(g) + " member range exceeds the member table");

  const Value& C = at(doc, "connections", file);
  const auto ncn = static_cast<size_t>(json::integer(C, "n", file));
  auto ci = [&](const char* k, Index lo, Index hi) {
    std::vector<Index> v = indices(at(C, k, file), file, std::string("connections.") + k, lo, hi);
    same_size(file, std::string("connections.") + k, v.size(), ncn);
    return v;
  };
  auto cf = [&](const char* k) {
    std::vector<Real> v = numbers(at(C, k, file), file, std::string("connections.") + k);
    same_size(file, std::string("connections.") + k, v.size(), ncn);
    return v;
  };
  auto cb8 = [&](const char* k) {
    std::vector<Byte> v = bytes(at(C, k, file), file, std::string("connections.") + k);
    same_size(file, std::string("connections.") + k, v.size(), ncn);
    return v;
  };
  M.ca = ci("a", -1, static_cast<Index>(nn) - 1);
  M.cb = ci("b", 0, static_cast<Index>(nn) - 1);
  M.cLaw = cb8("type");
  M.cx = cf("x"); M.cy = cf("y"); M.cz = cf("z"); M.cArea = cf("area"); M.cCoef = cf("coef");
  M.cKind = cb8("kind");
  M.cIdx = ci("idx", -1, 1 << 30);
  M.cEn = cb8("en");
  M.cTOn = cf("tOn");
  M.cSkipGroup = ci("skipGroup", -1, static_cast<Index>(ng) - 1);
  M.cMonitor = ci("monitor", -1, 1 << 30);
  if (const Value* kn = C.find("kindNames")) M.kindNames = strings(*kn, file, "connections.kindNames");

  const Value& Mo = at(doc, "monitors", file);
  M.monitorThreshold = real_at(Mo, "threshold", file);
  M.monitorId = strings(at(Mo, "id", file), file, "monitors.id");
  M.monitorLabel = strings(at(Mo, "label", file), file, "monitors.label");
  same_size(file, "monitors.label", M.monitorLabel.size(), M.monitorId.size());
  for (size_t c = 0; c < ncn; ++c)
    if (M.cMonitor[c] >= static_cast<Index>(M.monitorId.size())) bad(file, "connection " + std::to_string(c) + " refers to a monitor that does not exist");

  const Value& Mk = at(doc, "marks", file);
  M.markId = strings(at(Mk, "id", file), file, "marks.id");
  M.markLabel = strings(at(Mk, "label", file), file, "marks.label");
  M.mkx = numbers(at(Mk, "x", file), file, "marks.x");
  M.mky = numbers(at(Mk, "y", file), file, "marks.y");
  M.mkz = numbers(at(Mk, "z", file), file, "marks.z");
  M.mkWhen = bytes(at(Mk, "when", file), file, "marks.when");
  same_size(file, "marks.label", M.markLabel.size(), M.markId.size());
  same_size(file, "marks.x", M.mkx.size(), M.markId.size());
  same_size(file, "marks.y", M.mky.size(), M.markId.size());
  same_size(file, "marks.z", M.mkz.size(), M.markId.size());
  same_size(file, "marks.when", M.mkWhen.size(), M.markId.size());

  const Value& Fo = at(doc, "founder", file);
  M.founder[0] = real_at(Fo, "hullTopBelow", file);
  M.founder[1] = real_at(Fo, "pitchAbsAbove", file);
  M.founder[2] = real_at(Fo, "buoyancyBelow", file);
  M.founder[3] = real_at(Fo, "minT", file);
  const Value& B = at(doc, "body", file);
  const std::vector<Real> g0 = numbers(at(B, "g0", file), file, "body.g0");
  same_size(file, "body.g0", g0.size(), 3);
  const char* bk[9] = {"mh0", "Ith0", "Iph0", "cDz", "cDth", "cDph", "qz", "qth", "qph"};
  M.body[0] = real_at(B, "ms", file); M.body[1] = real_at(B, "KG", file);
  M.body[2] = g0[0]; M.body[3] = g0[1]; M.body[4] = g0[2];
  M.body[5] = real_at(B, "XO", file); M.body[6] = real_at(B, "YO", file);
  for (int k = 0; k < 9; ++k) M.body[7 + k] = real_at(B, bk[k], file);
  const Value& R = at(doc, "readout", file);
  M.readout[0] = real_at(R, "xFP", file);
  M.readout[1] = real_at(R, "xAP", file);

  M.hashAll = verify_hashes(doc, file, {
      {"groups.nodeStart", 'i', M.gNodeStart.data(), ng}, {"groups.nodeCount", 'i', M.gNodeCount.data(), ng},
      {"groups.nodes", 'i', M.gNodes.data(), M.gNodes.size()}, {"groups.mode", 'b', M.gMode.data(), ng},
      {"groups.px", 'f', M.gpx.data(), ng}, {"groups.py", 'f', M.gpy.data(), ng}, {"groups.pz", 'f', M.gpz.data(), ng},
      {"connections.a", 'i', M.ca.data(), ncn}, {"connections.b", 'i', M.cb.data(), ncn}, {"connections.type", 'b', M.cLaw.data(), ncn},
      {"connections.x", 'f', M.cx.data(), ncn}, {"connections.y", 'f', M.cy.data(), ncn}, {"connections.z", 'f', M.cz.data(), ncn},
      {"connections.area", 'f', M.cArea.data(), ncn}, {"connections.coef", 'f', M.cCoef.data(), ncn},
      {"connections.kind", 'b', M.cKind.data(), ncn}, {"connections.idx", 'i', M.cIdx.data(), ncn},
      {"connections.en", 'b', M.cEn.data(), ncn}, {"connections.tOn", 'f', M.cTOn.data(), ncn},
      {"connections.skipGroup", 'i', M.cSkipGroup.data(), ncn}, {"connections.monitor", 'i', M.cMonitor.data(), ncn},
      {"marks.x", 'f', M.mkx.data(), M.mkx.size()}, {"marks.y", 'f', M.mky.data(), M.mky.size()}, {"marks.z", 'f', M.mkz.data(), M.mkz.size()},
      {"marks.when", 'b', M.mkWhen.data(), M.mkWhen.size()},
      {"state0.vol", 'f', M.vol0.data(), nn}, {"state0.level", 'f', M.level0.data(), nn},
      {"state0.cx", 'f', M.cx0.data(), nn}, {"state0.cy", 'f', M.cy0.data(), nn}, {"state0.cz", 'f', M.cz0.data(), nn},
  });
  return M;
}

void check_pairing(const CompiledShip& ship, const CompiledSim& sim) {
  if (sim.shipId != ship.id) throw std::runtime_error("simulation " + sim.id + " was compiled for ship " + sim.shipId + ", not " + ship.id);
  if (sim.shipHash != ship.hashAll)
    throw std::runtime_error("simulation " + sim.id + " was compiled against a different build of ship " + ship.id + " (hash " + sim.shipHash + " vs " +
                             ship.hashAll + ")");
  if (sim.nodes() != ship.nodes())
    throw std::runtime_error("simulation " + sim.id + " has " + std::to_string(sim.nodes()) + " nodes, ship " + ship.id + " has " +
                             std::to_string(ship.nodes()));
}

}  // namespace io
}  // namespace sinksim
