// Compiled ship and simulation files on the host (C++): owning containers, the loaders that range-check and re-hash
// every array, and the views the C ABI takes. Reference: engine/include/sinksim/model.hpp:15-109,
// engine/include/sinksim/hash.hpp:15-56, engine/src/compiled.cpp:82-279.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/sinksim_hip.h"

namespace sinksim {

using Real = double;
using Index = std::int32_t;
using Byte = std::uint8_t;

// FNV-1a 64 over the little-endian bytes of numeric arrays, -0 hashed as 0 (hash.hpp:15-56).
class Fnv1a64 {
 public:
  void update(const void* data, size_t n);
  void update_f64(const double* v, size_t n);
  void update_i32(const int32_t* v, size_t n);
  void update_u8(const uint8_t* v, size_t n) { update(v, n); }
  uint64_t value() const { return h_; }
  std::string hex() const;

 private:
  uint64_t h_ = 14695981039346656037ULL;
};

struct CompiledShip {
  std::string id, title, hashAll, producer;
  std::vector<Real> x, y, dxdy, zlo, zhi;
  std::vector<Index> zone, side;
  std::vector<std::string> nodeLabel;
  std::vector<Real> mu, vmax, amin, aov;
  std::vector<Index> segStart, segCount, segCol;
  std::vector<Real> segA, segE;

  Index columns() const { return static_cast<Index>(x.size()); }
  Index nodes() const { return static_cast<Index>(mu.size()); }
  Index segments() const { return static_cast<Index>(segCol.size()); }
  ss_ship view() const;
};

struct CompiledSim {
  std::string id, title, shipId, shipHash, hashAll, producer;
  Real rho = 0, g = 0, dt = 0;
  std::vector<Index> gNodeStart, gNodeCount, gNodes;
  std::vector<Byte> gMode;
  std::vector<Real> gpx, gpy, gpz;
  Real gMargin = 0;
  std::vector<Index> ca, cb;
  std::vector<Byte> cLaw;
  std::vector<Real> cx, cy, cz, cArea, cCoef;
  std::vector<Byte> cKind;
  std::vector<Index> cIdx;
  std::vector<Byte> cEn;
  std::vector<Real> cTOn;
  std::vector<Index> cSkipGroup, cMonitor;
  std::vector<std::string> kindNames;
  Real monitorThreshold = 0;
  std::vector<std::string> monitorId, monitorLabel, markId, markLabel;
  std::vector<Real> mkx, mky, mkz;
  std::vector<Byte> mkWhen;
  Real founder[4] = {0, 0, 0, 0};  // hullTopBelow pitchAbsAbove buoyancyBelow minT
  Real body[16] = {};              // ms KG g0x g0y g0z XO YO mh0 Ith0 Iph0 cDz cDth cDph qz qth qph
  Real readout[2] = {0, 0};        // xFP xAP
  Real state0[7] = {};             // t zO pitch roll vz wth wph
  std::vector<Real> vol0, level0, cx0, cy0, cz0;

  Index groups() const { return static_cast<Index>(gNodeCount.size()); }
  Index connections() const { return static_cast<Index>(ca.size()); }
  Index monitors() const { return static_cast<Index>(monitorId.size()); }
  Index marks() const { return static_cast<Index>(mkx.size()); }
  Index nodes() const { return static_cast<Index>(vol0.size()); }
  Index kindIndex(const std::string& name) const;  // -1 when the kind is not defined
  ss_sim view() const;
};

namespace io {
std::shared_ptr<CompiledShip> load_ship(const std::string& file);
CompiledSim load_sim(const std::string& file);
void check_pairing(const CompiledShip& ship, const CompiledSim& sim);
}  // namespace io

// Event ids and labels (engine/src/simulation.cpp:225-241): kind 0 overflow (monitor), 1 mark, 2 founder.
std::string event_id(const CompiledSim& sim, int kind, int id);
std::string event_label(const CompiledSim& sim, int kind, int id);

}  // namespace sinksim
