"""Sweeps (docs/FORMATS.md "sinksim.sweep" and "sinksim.batch-results"; reference: engine/include/sinksim/sweep.hpp,
engine/src/sweep.cpp:30-281 and tools/py/sweep.py).

A sweep is many instances of one compiled simulation that differ by connection-kind scale factors and
per-connection overrides. `apply` turns each instance into its connection table (enabled, area, activation time)
exactly as the reference's three setters would, so a batch can take the whole table in one call.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np


@dataclass
class ConnectionOverride:
    """sweep.hpp:20-29: one connection (conn >= 0) or every connection of a kind (and idx unless -1)."""
    conn: int = -1
    kind: int = 255
    idx: int = -1
    enabled: bool | None = None
    tOn: float | None = None
    area: float | None = None


@dataclass
class Instance:
    id: str
    kindScale: list = field(default_factory=list)   # [(kind byte, factor)] in file order
    overrides: list = field(default_factory=list)


@dataclass
class Sweep:
    ship: str = ""
    sim: str = ""
    tMax: float = 5 * 3600.0
    every: float = 30.0
    stopWhenStable: bool = False
    instances: list = field(default_factory=list)


def _kind_byte(sim, name: str, where: str) -> int:
    k = sim.kind_index(name)
    if k < 0:
        raise RuntimeError(f'{where}: unknown connection kind "{name}"')
    return k


def _kind_name(sim, kind: int) -> str:
    return sim.kindNames[kind] if kind < len(sim.kindNames) else str(kind)


def _integer(o: dict, key: str, where: str) -> int:
    v = o.get(key)
    if isinstance(v, bool) or not isinstance(v, (int, float)) or v != math.floor(v):
        raise RuntimeError(f'{where}: member "{key}" is not an integer')
    return int(v)


def load_sweep_doc(doc: dict, sim, path: str = "<sweep>") -> Sweep:
    """sweep::load_sweep (engine/src/sweep.cpp:30-75) on a parsed document."""
    if doc.get("format") != "sinksim.sweep":
        raise RuntimeError(f"{path}: format is not sinksim.sweep")
    if _integer(doc, "formatVersion", path) != 1:
        raise RuntimeError(f"{path}: unsupported sweep formatVersion")
    num = lambda k, d: float(doc[k]) if isinstance(doc.get(k), (int, float)) and not isinstance(doc.get(k), bool) else d
    sw = Sweep(ship=doc.get("ship", "") if isinstance(doc.get("ship"), str) else "",
               sim=doc.get("sim", "") if isinstance(doc.get("sim"), str) else "",
               tMax=num("tMax", 5 * 3600.0), every=num("every", 30.0),
               stopWhenStable=doc.get("stopWhenStable") is True)
    if "instances" not in doc:
        raise RuntimeError(f'{path}: missing member "instances"')
    if not isinstance(doc["instances"], list):
        raise RuntimeError(f"{path}: instances must be an array")
    for k, j in enumerate(doc["instances"]):
        inst = Instance(id=j["id"] if isinstance(j.get("id"), str) else f"instance{k}")
        where = f"{path} instance {inst.id}"
        if "scale" in j:
            s = j["scale"]
            if not isinstance(s, dict):
                raise RuntimeError(f"{where}: scale must be an object of kind -> factor")
            for name, v in s.items():
                if isinstance(v, bool) or not isinstance(v, (int, float)):
                    raise RuntimeError(f"{where}: scale factor for {name} is not a number")
                inst.kindScale.append((_kind_byte(sim, name, where), float(v)))
        if "connections" in j:
            lst = j["connections"]
            if not isinstance(lst, list):
                raise RuntimeError(f"{where}: connections must be an array")
            for o in lst:
                ov = ConnectionOverride()
                if "conn" in o:
                    ov.conn = _integer(o, "conn", where)
                else:
                    if not isinstance(o.get("kind"), str):
                        raise RuntimeError(f'{where}: member "kind" is not a string')
                    ov.kind = _kind_byte(sim, o["kind"], where)
                    if "idx" in o:
                        ov.idx = _integer(o, "idx", where)
                if "enabled" in o:
                    if not isinstance(o["enabled"], bool):
                        raise RuntimeError(f'{where}: member "enabled" is not a boolean')
                    ov.enabled = o["enabled"]
                if "tOn" in o:
                    ov.tOn = float(o["tOn"])
                if "area" in o:
                    ov.area = float(o["area"])
                inst.overrides.append(ov)
        sw.instances.append(inst)
    return sw


def load_sweep(path: str, sim) -> Sweep:
    with open(path, "r", encoding="utf-8") as f:
        return load_sweep_doc(json.load(f), sim, path)


def sweep_document(sw: Sweep, sim) -> dict:
    """The document save_sweep writes (engine/src/sweep.cpp:77-112)."""
    insts = []
    for inst in sw.instances:
        j = {"id": inst.id}
        if inst.kindScale:
            j["scale"] = {_kind_name(sim, k): f for k, f in inst.kindScale}
        if inst.overrides:
            lst = []
            for o in inst.overrides:
                x = {"conn": o.conn} if o.conn >= 0 else {"kind": _kind_name(sim, o.kind), **({"idx": o.idx} if o.idx >= 0 else {})}
                if o.enabled is not None:
                    x["enabled"] = o.enabled
                if o.tOn is not None:
                    x["tOn"] = o.tOn
                if o.area is not None:
                    x["area"] = o.area
                lst.append(x)
            j["connections"] = lst
        insts.append(j)
    return {"format": "sinksim.sweep", "formatVersion": 1, "ship": sw.ship, "sim": sw.sim, "tMax": sw.tMax,
            "every": sw.every, "stopWhenStable": sw.stopWhenStable, "instances": insts}


def save_sweep(sw: Sweep, path: str, sim) -> None:
    with open(path, "w", encoding="utf-8") as f:
        json.dump(sweep_document(sw, sim), f, indent=2)
        f.write("\n")


def apply(inst: Instance, sim, en: np.ndarray, area: np.ndarray, tOn: np.ndarray) -> None:
    """sweep::apply (engine/include/sinksim/sweep.hpp:49-65) into one instance's connection table: kind scales
    multiply the compiled area, then the overrides in order."""
    kinds = sim.cKind
    for kind, factor in inst.kindScale:
        m = kinds == kind
        area[m] = sim.cArea[m] * factor
    for o in inst.overrides:
        if o.conn >= 0:
            m = np.zeros(sim.connections, bool)
            if o.conn < sim.connections:
                m[o.conn] = True
        else:
            m = (kinds == o.kind) & ((o.idx < 0) | (sim.cIdx == o.idx))
        if o.enabled is not None:
            en[m] = 1 if o.enabled else 0
        if o.area is not None:
            area[m] = o.area
        if o.tOn is not None:
            tOn[m] = o.tOn


def connection_tables(sw: Sweep, sim) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """[count, nconn] en / area / tOn tables of every instance of the sweep."""
    n = len(sw.instances)
    en = np.tile(sim.cEn.astype(np.uint8), (n, 1))
    area = np.tile(sim.cArea.astype(np.float64), (n, 1))
    tOn = np.tile(sim.cTOn.astype(np.float64), (n, 1))
    for i, inst in enumerate(sw.instances):
        apply(inst, sim, en[i], area[i], tOn[i])
    return en, area, tOn


# ------------------------------------------------------------------------------------------- results

@dataclass
class InstanceResult:
    """sweep.hpp:78-87; final = trimDeg listDeg waterT draftF draftA inflowTpm; curve [n, 7] = t trim list water
    inflow draftF draftA; events [(t, id, label)]."""
    id: str
    foundered: bool
    founderT: float
    tEnd: float
    steps: int
    final: dict
    events: list
    curve: np.ndarray


@dataclass
class Results:
    ship: str = ""
    sim: str = ""
    numerics: str = ""
    engine: str = ""
    tMax: float = 0.0
    every: float = 0.0
    instances: list = field(default_factory=list)


CURVE_KEYS = ("t", "trim", "list", "water", "inflow", "draftF", "draftA")


def results_document(r: Results) -> dict:
    """save_results (engine/src/sweep.cpp:163-201)."""
    insts = []
    for x in r.instances:
        c = np.asarray(x.curve).reshape(-1, 7)
        insts.append({"id": x.id, "foundered": bool(x.foundered), "founderT": x.founderT, "tEnd": x.tEnd,
                      "steps": int(x.steps), "final": dict(x.final),
                      "events": [{"t": t, "id": i, "label": lab} for t, i, lab in x.events],
                      "curve": {k: [float(v) for v in c[:, j]] for j, k in enumerate(CURVE_KEYS)}})
    return {"format": "sinksim.batch-results", "formatVersion": 1, "ship": r.ship, "sim": r.sim, "numerics": r.numerics,
            "engine": r.engine, "tMax": r.tMax, "every": r.every, "instances": insts}


def save_results(r: Results, path: str) -> None:
    with open(path, "w", encoding="utf-8") as f:
        json.dump(results_document(r), f, indent=2)
        f.write("\n")


def load_results(path: str) -> Results:
    """load_results (engine/src/sweep.cpp:203-232)."""
    with open(path, "r", encoding="utf-8") as f:
        doc = json.load(f)
    if doc.get("format") != "sinksim.batch-results":
        raise RuntimeError(f"{path}: format is not sinksim.batch-results")
    r = Results(doc.get("ship", ""), doc.get("sim", ""), doc.get("numerics", ""), doc.get("engine", ""),
                float(doc.get("tMax", 0)), float(doc.get("every", 0)))
    for j in doc["instances"]:
        c = j["curve"]
        curve = np.stack([np.asarray(c[k], dtype=np.float64) for k in CURVE_KEYS], axis=1) if c["t"] else np.zeros((0, 7))
        r.instances.append(InstanceResult(j.get("id", ""), bool(j["foundered"]), float(j["founderT"]), float(j["tEnd"]),
                                          int(j["steps"]), {k: float(v) for k, v in j["final"].items()},
                                          [(float(e["t"]), e["id"], e.get("label", "")) for e in j["events"]], curve))
    return r


@dataclass
class Diff:
    differing: int = 0
    compared: int = 0
    maxAbs: float = 0.0
    first: str = ""


def compare(a: Results, b: Results) -> Diff:
    """sweep::compare (engine/src/sweep.cpp:234-281): exact, value for value."""
    d = Diff()

    def note(where, x, y):
        d.compared += 1
        if x != y and not (isinstance(x, float) and isinstance(y, float) and math.isnan(x) and math.isnan(y)):
            d.differing += 1
            ad = abs(float(x) - float(y))
            d.maxAbs = max(d.maxAbs, ad)
            if not d.first:
                d.first = f"{where}: {x!r} vs {y!r}"

    if len(a.instances) != len(b.instances):
        d.differing += 1
        d.first = f"instance count {len(a.instances)} vs {len(b.instances)}"
        return d
    for x, y in zip(a.instances, b.instances):
        w = f"instance {x.id}"
        note(w + " foundered", float(x.foundered), float(y.foundered))
        note(w + " founderT", x.founderT, y.founderT)
        note(w + " tEnd", x.tEnd, y.tEnd)
        note(w + " steps", float(x.steps), float(y.steps))
        for k, name in (("trimDeg", "trim"), ("listDeg", "list"), ("waterT", "water"), ("draftF", "draftF"), ("draftA", "draftA")):
            note(f"{w} {name}", x.final[k], y.final[k])
        note(w + " events", float(len(x.events)), float(len(y.events)))
        for ex, ey in zip(x.events, y.events):
            note(f"{w} event {ex[1]} t", ex[0], ey[0])
            if ex[1] != ey[1]:
                d.differing += 1
                if not d.first:
                    d.first = f"{w} event id {ex[1]} vs {ey[1]}"
        note(w + " curve length", float(len(x.curve)), float(len(y.curve)))
        for k in range(min(len(x.curve), len(y.curve))):
            for j, name in enumerate(CURVE_KEYS):
                note(f"{w} curve[{k}] {name}", float(x.curve[k][j]), float(y.curve[k][j]))
    return d


def final_dict(rec7) -> dict:
    """A curve-layout record (t trim list water inflow draftF draftA) as the results' "final" object."""
    return {"trimDeg": float(rec7[1]), "listDeg": float(rec7[2]), "waterT": float(rec7[3]), "draftF": float(rec7[5]),
            "draftA": float(rec7[6]), "inflowTpm": float(rec7[4])}


def data_dir() -> str:
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
