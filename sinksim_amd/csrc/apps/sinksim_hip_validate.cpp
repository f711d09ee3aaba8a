// sinksim_hip_validate: run every case of a catalog on the GPU with the oracle's run options (curves every 30 s,
// stopWhenStable, the case's tMaxH) and print the validation table; optionally compare with a reference table and
// write the engine's own. The drop-in for the reference's sinksim_validate (apps/sinksim_validate.cpp:20-120): the
// same arguments, output lines, comparison tolerances, table format and exit codes. All cases of one tMax run as one
// fleet batch (one instance per case, each of its own scenario), so the catalog costs one launch's latency.
//   sinksim_hip_validate <catalog.json> [--numerics portable|portable32] [--compare validation.oracle.json] [--out validation.engine.json]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <exception>
#include <filesystem>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../host/batch.hpp"
#include "../host/json.hpp"
#include "options.hpp"

using namespace sinksim;
namespace fs = std::filesystem;

namespace {

struct Case {
  std::string id, label;
  CompiledSim sim;
  double tMax;
};

struct Outcome {
  bool foundered = false;
  double founderT = -1, tEnd = 0, trimDeg = 0, listDeg = 0, waterT = 0;
  std::vector<std::pair<double, std::pair<std::string, std::string>>> events;  // t, (id, label)
};

// One fleet batch over the cases (one instance each), stopWhenStable on; the outcomes in case order.
std::vector<Outcome> run_cases(const std::shared_ptr<const CompiledShip>& ship, const std::vector<const Case*>& cases, double tMax, bool mixed) {
  const int32_t n = static_cast<int32_t>(cases.size());
  std::vector<ss_ship> ships(static_cast<size_t>(n), ship->view());
  std::vector<ss_sim> sims;
  std::vector<int32_t> instModel;
  for (int32_t i = 0; i < n; ++i) {
    sims.push_back(cases[static_cast<size_t>(i)]->sim.view());
    instModel.push_back(i);
  }
  ss_batch_options o{};
  o.count = n;
  o.numerics = mixed ? SS_NUMERICS_PORTABLE32 : SS_NUMERICS_PORTABLE;
  o.traced = 0;
  o.denseSteps = 0;
  o.sampleEvery = 60;
  o.curveEvery = 30;  // the oracle's run option (apps/sinksim_validate.cpp:50)
  o.tMax = tMax;
  o.stepsPerLaunch = 0;
  ss_batch* raw = nullptr;
  hip::check(ss_batch_create_fleet(ships.data(), sims.data(), n, instModel.data(), &o, &raw));
  std::unique_ptr<ss_batch, void (*)(ss_batch*)> b(raw, [](ss_batch* p) { ss_batch_destroy(p); });
  hip::check(ss_batch_set_stop_when_stable(b.get(), 1));
  double secs = 0;
  hip::check(ss_batch_run(b.get(), &secs));
  std::vector<Outcome> out;
  for (int32_t i = 0; i < n; ++i) {
    const CompiledSim& sim = cases[static_cast<size_t>(i)]->sim;
    const size_t nn = static_cast<size_t>(ship->nodes());
    std::vector<double> vol(nn), level(nn), cx(nn), cy(nn), cz(nn);
    const int32_t cap = sim.monitors() + sim.marks() + 1;
    std::vector<ss_event> ev(static_cast<size_t>(cap));
    ss_instance_result x{};
    hip::check(ss_batch_result(b.get(), i, &x, vol.data(), level.data(), cx.data(), cy.data(), cz.data(), ev.data(), cap));
    double r7[7];
    hip::check(ss_batch_final_readouts(b.get(), i, r7));
    Outcome oc;
    oc.foundered = x.foundered != 0;
    oc.founderT = x.founderT;
    oc.tEnd = x.tEnd;
    oc.trimDeg = r7[1];
    oc.listDeg = r7[2];
    oc.waterT = r7[3];
    for (int32_t k = 0; k < std::min(x.events, cap); ++k) {
      const ss_event& e = ev[static_cast<size_t>(k)];
      oc.events.push_back({e.t, {event_id(sim, e.kind, e.id), event_label(sim, e.kind, e.id)}});
    }
    out.push_back(std::move(oc));
  }
  return out;
}

}  // namespace

int main(int argc, char** argv) {
  const tools::CommandLine args(argc, argv, {});
  if (!args.error().empty()) {
    std::fprintf(stderr, "%s\n", args.error().c_str());
    return 2;
  }
  if (args.positional().empty()) {
    std::fprintf(stderr, "usage: sinksim_hip_validate <catalog.json> [--numerics portable|portable32] [--compare validation.oracle.json] [--out validation.engine.json]\n");
    return 2;
  }
  bool mixed = false;
  if (!tools::device_numerics(args, "sinksim_hip_validate", mixed, "portable")) return 2;
  try {
    std::string why;
    if (!hip::available(&why)) {
      std::fprintf(stderr, "sinksim_hip_validate: no usable HIP device (%s)\n", why.c_str());
      return 1;
    }
    const fs::path catalogPath = args.positional()[0];
    const std::string cat = catalogPath.string();
    const json::Value catalog = json::read_file(cat);
    if (json::string_or(catalog, "format", "") != "sinksim.catalog") throw std::runtime_error(cat + ": not a sinksim.catalog file");
    const fs::path dir = catalogPath.parent_path();
    const std::shared_ptr<const CompiledShip> ship = io::load_ship((dir / json::string(catalog, "shipFile", cat)).string());
    json::Value reference;
    if (args.has("compare")) reference = json::read_file(args.text("compare", ""));
    const json::Value* refCases = reference.is_object() ? &json::member(reference, "cases", "reference table") : nullptr;

    std::vector<Case> cases;
    for (const json::Value& cs : json::member(catalog, "cases", cat).items()) {
      Case c{json::string(cs, "id", cat), json::string(cs, "label", cat), io::load_sim((dir / json::string(cs, "file", cat)).string()),
             json::number_or(cs, "tMaxH", 8) * 3600};
      io::check_pairing(*ship, c.sim);
      cases.push_back(std::move(c));
    }
    const auto tStart = std::chrono::steady_clock::now();
    // one fleet per distinct tMax
    std::vector<Outcome> outcome(cases.size());
    std::map<double, std::vector<size_t>> byT;
    for (size_t i = 0; i < cases.size(); ++i) byT[cases[i].tMax].push_back(i);
    for (const auto& [tMax, idx] : byT) {
      std::vector<const Case*> group;
      for (size_t i : idx) group.push_back(&cases[i]);
      const std::vector<Outcome> r = run_cases(ship, group, tMax, mixed);
      for (size_t k = 0; k < idx.size(); ++k) outcome[idx[k]] = r[k];
    }

    json::Value outCases = json::Value::object();
    int mismatches = 0;
    for (size_t i = 0; i < cases.size(); ++i) {
      const Case& c = cases[i];
      const Outcome& r = outcome[i];
      std::printf("%-44s %s  trim %.2f\xC2\xB0  list %.2f\xC2\xB0  water %.0f t\n", c.label.c_str(),
                  r.foundered ? ("FOUNDERS at " + tools::hours_minutes(r.founderT)).c_str() : ("afloat (checked to " + tools::hours_minutes(r.tEnd) + ")").c_str(), r.trimDeg,
                  r.listDeg, std::round(r.waterT));
      json::Value rec = json::Value::object();
      rec["label"] = c.label;
      rec["foundered"] = r.foundered;
      rec["founderT"] = r.foundered ? json::Value(r.founderT) : json::Value(nullptr);
      rec["founderMin"] = r.foundered ? json::Value(std::round(r.founderT / 60 * 10) / 10) : json::Value(nullptr);
      rec["endT"] = r.tEnd;
      rec["trimDeg"] = r.trimDeg;
      rec["listDeg"] = r.listDeg;
      rec["waterT"] = r.waterT;
      json::Value ev = json::Value::array();
      for (const auto& e : r.events) {
        json::Value x = json::Value::object();
        x["tSec"] = e.first;
        x["id"] = e.second.first;
        x["label"] = e.second.second;
        ev.push_back(std::move(x));
      }
      rec["events"] = std::move(ev);
      outCases[c.id] = std::move(rec);

      if (refCases) {
        const json::Value* refCase = refCases->find(c.id);
        if (!refCase) { std::printf("   (no reference for %s)\n", c.id.c_str()); continue; }
        const bool refFoundered = json::boolean(*refCase, "foundered", c.id);
        std::string problems;
        if (refFoundered != r.foundered) problems += " verdict";
        if (refFoundered && r.foundered) {
          const double dT = std::fabs(json::number(*refCase, "founderT", c.id) - r.founderT);
          if (dT > 1.0) problems += " founderT(" + std::to_string(dT) + " s)";
        }
        // the reference's tolerances (apps/sinksim_validate.cpp:87-99): equilibria are tight, a foundering plunge is
        // chaotic in its last minute and compared loosely
        const double tolTrim = refFoundered ? 0.5 : 0.01, tolWater = refFoundered ? 200 : 1, tolList = refFoundered ? 0.5 : 0.01;
        const double dTrim = std::fabs(json::number(*refCase, "trimDeg", c.id) - r.trimDeg);
        const double dList = std::fabs(json::number(*refCase, "listDeg", c.id) - r.listDeg);
        const double dWater = std::fabs(json::number(*refCase, "waterT", c.id) - r.waterT);
        if (dTrim > tolTrim) problems += " trim(" + std::to_string(dTrim) + ")";
        if (dList > tolList) problems += " list(" + std::to_string(dList) + ")";
        if (dWater > tolWater) problems += " water(" + std::to_string(dWater) + " t)";
        if (!refFoundered) {
          const double dEnd = std::fabs(json::number(*refCase, "endT", c.id) - r.tEnd);
          if (dEnd > 60) problems += " endT(" + std::to_string(dEnd) + " s)";
        }
        if (!problems.empty()) {
          ++mismatches;
          std::printf("   MISMATCH vs reference:%s\n", problems.c_str());
        }
      }
    }
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - tStart).count();
    std::printf("%zu cases in %.1f s\n", cases.size(), s);
    if (args.has("out")) {
      json::Value doc = json::Value::object();
      doc["format"] = "sinksim.validation";
      doc["formatVersion"] = 1;
      doc["ship"] = ship->id;
      doc["producer"] = std::string("sinksim_hip_validate (HIP batch engine, numerics ") + (mixed ? "portable32" : "portable") + ")";
      doc["cases"] = std::move(outCases);
      json::write_file(args.text("out", ""), doc, 2);
      std::printf("wrote %s\n", args.text("out", "").c_str());
    }
    if (refCases) std::printf("RESULT: %s (%d mismatching cases)\n", mismatches == 0 ? "PASS" : "FAIL", mismatches);
    return mismatches == 0 ? 0 : 1;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "sinksim_hip_validate: %s\n", e.what());
    return 1;
  }
}
