"""Compiled ship and simulation files: load, range-check and re-hash every array.

Mirror of the reference's loaders (engine/src/compiled.cpp:108-279) with the same refusals: an unknown format or
formatVersion, a missing member, a non-integer index or one out of range, an array of the wrong length, an FNV-1a
mismatch on any array or on the chained hash, and a simulation paired with the wrong ship. Arrays are numpy
arrays in the engine's types (f64, i32, u8).
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field

import numpy as np

from .fnv import Fnv1a64

SHIP_FORMAT = "sinksim.compiled-ship"
SIM_FORMAT = "sinksim.compiled-sim"
FORMAT_VERSION = 1


class FormatError(RuntimeError):
    """A compiled file the engine refuses (the reference throws std::runtime_error with the same text)."""


def _bad(path: str, what: str) -> FormatError:
    return FormatError(f"{path}: {what}")


def _member(obj, key: str, path: str):
    if not isinstance(obj, dict):
        raise _bad(path, f'expected an object holding "{key}"')
    if key not in obj:
        raise _bad(path, f'missing member "{key}"')
    return obj[key]


def _number(obj, key: str, path: str) -> float:
    v = _member(obj, key, path)
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        raise _bad(path, f'member "{key}" is not a number')
    return float(v)


def _integer(obj, key: str, path: str) -> int:
    d = _number(obj, key, path)
    if d != math.floor(d):
        raise _bad(path, f'member "{key}" is not an integer')
    return int(d)


def _string(obj, key: str, path: str) -> str:
    v = _member(obj, key, path)
    if not isinstance(v, str):
        raise _bad(path, f'member "{key}" is not a string')
    return v


def _string_or(obj, key: str, fallback: str) -> str:
    if not isinstance(obj, dict):
        return fallback
    v = obj.get(key)
    return v if isinstance(v, str) else fallback


def _numbers(v, path: str, what: str) -> np.ndarray:
    if not isinstance(v, list):
        raise _bad(path, f"{what} must be an array")
    for i, x in enumerate(v):
        if isinstance(x, bool) or not isinstance(x, (int, float)):
            raise _bad(path, f"{what}[{i}] is not a number")
    return np.asarray(v, dtype=np.float64).reshape(-1)


def _indices(v, path: str, what: str, lo: int, hi: int) -> np.ndarray:
    d = _numbers(v, path, what)
    bad = (d != np.floor(d)) | (d < lo) | (d > hi)
    if bad.any():
        i = int(np.flatnonzero(bad)[0])
        raise _bad(path, f"{what}[{i}] = {d[i]!r} is not an integer in [{lo}, {hi}]")
    return d.astype(np.int32)


def _bytes(v, path: str, what: str) -> np.ndarray:
    return _indices(v, path, what, 0, 255).astype(np.uint8)


def _strings(v, path: str, what: str) -> list[str]:
    if not isinstance(v, list):
        raise _bad(path, f"{what} must be an array")
    for s in v:
        if not isinstance(s, str):
            raise _bad(path, f"{what} must contain strings")
    return list(v)


def _same_size(path: str, what: str, got: int, want: int) -> None:
    if got != want:
        raise _bad(path, f"{what} has {got} entries, expected {want}")


def _check_format(doc, path: str, fmt: str, version: int) -> None:
    if _string_or(doc, "format", "") != fmt:
        raise _bad(path, f"format is not {fmt}")
    v = _integer(doc, "formatVersion", path)
    if v != version:
        raise _bad(path, f"formatVersion {v} is not supported (engine reads {version})")


def _verify_hashes(doc, path: str, arrays: list[tuple[str, np.ndarray]]) -> str:
    """Each array's FNV-1a against hash.<name>, then the chain over all of them against hash.all
    (engine/src/compiled.cpp:82-96)."""
    H = _member(doc, "hash", path)
    chain = Fnv1a64()
    for name, a in arrays:
        one = Fnv1a64()
        one.update_array(a)
        chain.update_array(a)
        want = _string(H, name, path + " hash")
        if one.hex() != want:
            raise _bad(path, f"hash mismatch on {name} ({one.hex()} vs {want}): the JSON round trip was not exact")
    want = _string(H, "all", path + " hash")
    if chain.hex() != want:
        raise _bad(path, f"chained hash mismatch ({chain.hex()} vs {want})")
    return chain.hex()


def _read(path: str):
    try:
        with open(path, "rb") as f:
            return json.loads(f.read())
    except OSError:
        raise FormatError(f"cannot open {path}") from None
    except json.JSONDecodeError as e:
        raise FormatError(f"{path}: {e}") from None


@dataclass
class CompiledShip:
    """engine/include/sinksim/model.hpp:15-35"""
    id: str
    title: str
    hashAll: str
    x: np.ndarray
    y: np.ndarray
    dxdy: np.ndarray
    zlo: np.ndarray
    zhi: np.ndarray
    zone: np.ndarray
    side: np.ndarray
    nodeLabel: list
    mu: np.ndarray
    vmax: np.ndarray
    amin: np.ndarray
    aov: np.ndarray
    segStart: np.ndarray
    segCount: np.ndarray
    segCol: np.ndarray
    segA: np.ndarray
    segE: np.ndarray
    producer: str = ""

    @property
    def columns(self) -> int:
        return int(self.x.size)

    @property
    def nodes(self) -> int:
        return int(self.mu.size)

    @property
    def segments(self) -> int:
        return int(self.segCol.size)


@dataclass
class CompiledSim:
    """engine/include/sinksim/model.hpp:49-96. body = ms KG g0x g0y g0z XO YO mh0 Ith0 Iph0 cDz cDth cDph qz qth
    qph; founder = hullTopBelow pitchAbsAbove buoyancyBelow minT; readout = xFP xAP; state0 = t zO pitch roll vz
    wth wph."""
    id: str
    title: str
    shipId: str
    shipHash: str
    hashAll: str
    rho: float
    g: float
    dt: float
    gNodeStart: np.ndarray
    gNodeCount: np.ndarray
    gNodes: np.ndarray
    gMode: np.ndarray
    gpx: np.ndarray
    gpy: np.ndarray
    gpz: np.ndarray
    gMargin: float
    ca: np.ndarray
    cb: np.ndarray
    cLaw: np.ndarray
    cx: np.ndarray
    cy: np.ndarray
    cz: np.ndarray
    cArea: np.ndarray
    cCoef: np.ndarray
    cKind: np.ndarray
    cIdx: np.ndarray
    cEn: np.ndarray
    cTOn: np.ndarray
    cSkipGroup: np.ndarray
    cMonitor: np.ndarray
    kindNames: list
    monitorThreshold: float
    monitorId: list
    monitorLabel: list
    markId: list
    markLabel: list
    mkx: np.ndarray
    mky: np.ndarray
    mkz: np.ndarray
    mkWhen: np.ndarray
    founder: np.ndarray
    body: np.ndarray
    readout: np.ndarray
    state0: np.ndarray
    vol0: np.ndarray
    level0: np.ndarray
    cx0: np.ndarray
    cy0: np.ndarray
    cz0: np.ndarray
    producer: str = ""
    params: dict = field(default_factory=dict)

    @property
    def groups(self) -> int:
        return int(self.gNodeCount.size)

    @property
    def connections(self) -> int:
        return int(self.ca.size)

    @property
    def monitors(self) -> int:
        return len(self.monitorId)

    @property
    def marks(self) -> int:
        return int(self.mkx.size)

    @property
    def nodes(self) -> int:
        return int(self.vol0.size)

    def kind_index(self, name: str) -> int:
        """CompiledSim::kindIndex (engine/src/model.cpp:20-24): -1 when the kind is not defined."""
        try:
            return self.kindNames.index(name)
        except ValueError:
            return -1

    def event_id(self, kind: int, id_: int) -> str:
        """Simulation::event_id (engine/src/simulation.cpp:225-232)."""
        if kind == 0:
            return self.monitorId[id_]
        if kind == 1:
            return self.markId[id_]
        return "founder"

    def event_label(self, kind: int, id_: int) -> str:
        if kind == 0:
            return self.monitorLabel[id_]
        if kind == 1:
            return self.markLabel[id_]
        return "Foundered"


def load_ship(path: str) -> CompiledShip:
    """io::load_ship (engine/src/compiled.cpp:108-156)."""
    doc = _read(path)
    _check_format(doc, path, SHIP_FORMAT, FORMAT_VERSION)
    sid = _string(doc, "id", path)
    meta = doc.get("meta", {}) if isinstance(doc, dict) else {}
    title = _string_or(meta, "title", "")
    producer = _string_or(meta, "generator", "") + " on " + _string_or(meta, "engine", "")
    C = _member(doc, "columns", path)
    nc = _integer(C, "n", path)
    cols = {}
    for k in ("x", "y", "dxdy", "zlo", "zhi"):
        cols[k] = _numbers(_member(C, k, path), path, "columns." + k)
        _same_size(path, "columns." + k, cols[k].size, nc)
    zone = _indices(_member(C, "zone", path), path, "columns.zone", 0, 1 << 30)
    _same_size(path, "columns.zone", zone.size, nc)
    side = _indices(_member(C, "side", path), path, "columns.side", 0, 1)
    _same_size(path, "columns.side", side.size, nc)
    N = _member(doc, "nodes", path)
    nn = _integer(N, "n", path)
    label = _strings(_member(N, "label", path), path, "nodes.label")
    _same_size(path, "nodes.label", len(label), nn)
    nodes = {}
    for k in ("mu", "vmax", "amin", "aov"):
        nodes[k] = _numbers(_member(N, k, path), path, "nodes." + k)
        _same_size(path, "nodes." + k, nodes[k].size, nn)
    G = _member(doc, "segments", path)
    ns = _integer(G, "n", path)
    segStart = _indices(_member(N, "segStart", path), path, "nodes.segStart", 0, ns)
    _same_size(path, "nodes.segStart", segStart.size, nn)
    segCount = _indices(_member(N, "segCount", path), path, "nodes.segCount", 0, ns)
    _same_size(path, "nodes.segCount", segCount.size, nn)
    segCol = _indices(_member(G, "col", path), path, "segments.col", 0, nc - 1)
    _same_size(path, "segments.col", segCol.size, ns)
    segA = _numbers(_member(G, "a", path), path, "segments.a")
    _same_size(path, "segments.a", segA.size, ns)
    segE = _numbers(_member(G, "e", path), path, "segments.e")
    _same_size(path, "segments.e", segE.size, ns)
    for n in range(nn):
        if int(segStart[n]) + int(segCount[n]) > ns:
            raise _bad(path, f"node {n} segment range exceeds the segment table")
    hashAll = _verify_hashes(doc, path, [
        ("columns.x", cols["x"]), ("columns.y", cols["y"]), ("columns.dxdy", cols["dxdy"]),
        ("columns.zlo", cols["zlo"]), ("columns.zhi", cols["zhi"]), ("columns.zone", zone), ("columns.side", side),
        ("nodes.mu", nodes["mu"]), ("nodes.vmax", nodes["vmax"]), ("nodes.amin", nodes["amin"]),
        ("nodes.aov", nodes["aov"]), ("nodes.segStart", segStart), ("nodes.segCount", segCount),
        ("segments.col", segCol), ("segments.a", segA), ("segments.e", segE),
    ])
    return CompiledShip(sid, title, hashAll, cols["x"], cols["y"], cols["dxdy"], cols["zlo"], cols["zhi"], zone, side,
                        label, nodes["mu"], nodes["vmax"], nodes["amin"], nodes["aov"], segStart, segCount, segCol,
                        segA, segE, producer)


def load_sim(path: str) -> CompiledSim:
    """io::load_sim (engine/src/compiled.cpp:158-273)."""
    doc = _read(path)
    _check_format(doc, path, SIM_FORMAT, FORMAT_VERSION)
    sid = _string(doc, "id", path)
    shipId = _string(doc, "ship", path)
    shipHash = _string(doc, "shipHash", path)
    meta = doc.get("meta", {}) if isinstance(doc, dict) else {}
    title = _string_or(meta, "title", "")
    producer = _string_or(meta, "generator", "") + " on " + _string_or(meta, "engine", "")
    K = _member(doc, "constants", path)
    rho = _number(K, "rho", path)
    g = _number(K, "g", path)
    dt = _number(_member(doc, "scheme", path), "dt", path)

    st = _member(doc, "state0", path)
    state0 = np.array([_number(st, k, path) for k in ("t", "zO", "pitch", "roll", "vz", "wth", "wph")], dtype=np.float64)
    vol0 = _numbers(_member(st, "vol", path), path, "state0.vol")
    nn = vol0.size
    per_node = {}
    for k in ("level", "cx", "cy", "cz"):
        per_node[k] = _numbers(_member(st, k, path), path, "state0." + k)
        _same_size(path, "state0." + k, per_node[k].size, nn)

    G = _member(doc, "groups", path)
    ng = _integer(G, "n", path)
    gNodeCount = _indices(_member(G, "nodeCount", path), path, "groups.nodeCount", 1, 1 << 20)
    _same_size(path, "groups.nodeCount", gNodeCount.size, ng)
    gNodes = _indices(_member(G, "nodes", path), path, "groups.nodes", 0, nn - 1)
    gNodeStart = _indices(_member(G, "nodeStart", path), path, "groups.nodeStart", 0, gNodes.size)
    _same_size(path, "groups.nodeStart", gNodeStart.size, ng)
    gMode = _bytes(_member(G, "mode", path), path, "groups.mode")
    _same_size(path, "groups.mode", gMode.size, ng)
    gp = {}
    for k in ("px", "py", "pz"):
        gp[k] = _numbers(_member(G, k, path), path, "groups." + k)
        _same_size(path, "groups." + k, gp[k].size, ng)
    gMargin = _number(G, "margin", path)
    for gi in range(ng):
        if int(gNodeStart[gi]) + int(gNodeCount[gi]) > gNodes.size:
            raise _bad(path, f"group {gi} member range exceeds the member table")

    C = _member(doc, "connections", path)
    ncn = _integer(C, "n", path)

    def conn_idx(k, lo, hi):
        a = _indices(_member(C, k, path), path, "connections." + k, lo, hi)
        _same_size(path, "connections." + k, a.size, ncn)
        return a

    def conn_num(k):
        a = _numbers(_member(C, k, path), path, "connections." + k)
        _same_size(path, "connections." + k, a.size, ncn)
        return a

    def conn_byte(k):
        a = _bytes(_member(C, k, path), path, "connections." + k)
        _same_size(path, "connections." + k, a.size, ncn)
        return a

    ca = conn_idx("a", -1, nn - 1)
    cb = conn_idx("b", 0, nn - 1)
    law = conn_byte("type")
    cx, cy, cz, area, coef = conn_num("x"), conn_num("y"), conn_num("z"), conn_num("area"), conn_num("coef")
    kind = conn_byte("kind")
    idx = conn_idx("idx", -1, 1 << 30)
    en = conn_byte("en")
    tOn = conn_num("tOn")
    skip = conn_idx("skipGroup", -1, ng - 1)
    mon = conn_idx("monitor", -1, 1 << 30)
    kindNames = _strings(C["kindNames"], path, "connections.kindNames") if "kindNames" in C else []

    Mo = _member(doc, "monitors", path)
    monThreshold = _number(Mo, "threshold", path)
    monId = _strings(_member(Mo, "id", path), path, "monitors.id")
    monLabel = _strings(_member(Mo, "label", path), path, "monitors.label")
    _same_size(path, "monitors.label", len(monLabel), len(monId))
    for c in range(ncn):
        if mon[c] >= len(monId):
            raise _bad(path, f"connection {c} refers to a monitor that does not exist")

    Mk = _member(doc, "marks", path)
    markId = _strings(_member(Mk, "id", path), path, "marks.id")
    markLabel = _strings(_member(Mk, "label", path), path, "marks.label")
    mk = {k: _numbers(_member(Mk, k, path), path, "marks." + k) for k in ("x", "y", "z")}
    mkWhen = _bytes(_member(Mk, "when", path), path, "marks.when")
    _same_size(path, "marks.label", len(markLabel), len(markId))
    for k in ("x", "y", "z"):
        _same_size(path, "marks." + k, mk[k].size, len(markId))
    _same_size(path, "marks.when", mkWhen.size, len(markId))

    Fo = _member(doc, "founder", path)
    founder = np.array([_number(Fo, k, path) for k in ("hullTopBelow", "pitchAbsAbove", "buoyancyBelow", "minT")])
    B = _member(doc, "body", path)
    g0 = _numbers(_member(B, "g0", path), path, "body.g0")
    _same_size(path, "body.g0", g0.size, 3)
    body = np.array([_number(B, "ms", path), _number(B, "KG", path), g0[0], g0[1], g0[2], _number(B, "XO", path),
                     _number(B, "YO", path)] + [_number(B, k, path) for k in
                                                 ("mh0", "Ith0", "Iph0", "cDz", "cDth", "cDph", "qz", "qth", "qph")],
                    dtype=np.float64)
    R = _member(doc, "readout", path)
    readout = np.array([_number(R, "xFP", path), _number(R, "xAP", path)])

    hashAll = _verify_hashes(doc, path, [
        ("groups.nodeStart", gNodeStart), ("groups.nodeCount", gNodeCount), ("groups.nodes", gNodes),
        ("groups.mode", gMode), ("groups.px", gp["px"]), ("groups.py", gp["py"]), ("groups.pz", gp["pz"]),
        ("connections.a", ca), ("connections.b", cb), ("connections.type", law), ("connections.x", cx),
        ("connections.y", cy), ("connections.z", cz), ("connections.area", area), ("connections.coef", coef),
        ("connections.kind", kind), ("connections.idx", idx), ("connections.en", en), ("connections.tOn", tOn),
        ("connections.skipGroup", skip), ("connections.monitor", mon),
        ("marks.x", mk["x"]), ("marks.y", mk["y"]), ("marks.z", mk["z"]), ("marks.when", mkWhen),
        ("state0.vol", vol0), ("state0.level", per_node["level"]), ("state0.cx", per_node["cx"]),
        ("state0.cy", per_node["cy"]), ("state0.cz", per_node["cz"]),
    ])
    params = doc.get("params", {}) if isinstance(doc.get("params", {}), dict) else {}
    return CompiledSim(sid, title, shipId, shipHash, hashAll, rho, g, dt, gNodeStart, gNodeCount, gNodes, gMode,
                       gp["px"], gp["py"], gp["pz"], gMargin, ca, cb, law, cx, cy, cz, area, coef, kind, idx, en, tOn,
                       skip, mon, kindNames, monThreshold, monId, monLabel, markId, markLabel, mk["x"], mk["y"],
                       mk["z"], mkWhen, founder, body, readout, state0, vol0, per_node["level"], per_node["cx"],
                       per_node["cy"], per_node["cz"], producer, params)


def check_pairing(ship: CompiledShip, sim: CompiledSim) -> None:
    """io::check_pairing (engine/src/compiled.cpp:275-279)."""
    if sim.shipId != ship.id:
        raise FormatError(f"simulation {sim.id} was compiled for ship {sim.shipId}, not {ship.id}")
    if sim.shipHash != ship.hashAll:
        raise FormatError(f"simulation {sim.id} was compiled against a different build of ship {ship.id} "
                          f"(hash {sim.shipHash} vs {ship.hashAll})")
    if sim.nodes != ship.nodes:
        raise FormatError(f"simulation {sim.id} has {sim.nodes} nodes, ship {ship.id} has {ship.nodes}")


def max_group_size(sim: CompiledSim) -> int:
    """CompiledSim::maxGroupSize (engine/src/model.cpp:13-18)."""
    return max(1, int(sim.gNodeCount.max()) if sim.gNodeCount.size else 1)
