"""sinksim.trace files (docs/FORMATS.md; reference: engine/src/trace.cpp:11-179).

A trace holds the complete state at every record, dense for the first steps and sampled afterwards. Here the
records are [n, 10 + 5 * nodes] arrays in the engine's record layout: t zO pitch roll vz wth wph inflow Vb hullTop,
then vol, level, cx, cy, cz per node. Reading re-hashes every column (and the chain) when the file carries hashes.
"""
from __future__ import annotations

import gzip
import json
from dataclasses import dataclass, field

import numpy as np

from .fnv import Fnv1a64

SCALARS = ("t", "zO", "pitch", "roll", "vz", "wth", "wph", "inflow", "Vb", "hullTop")
PER_NODE = ("vol", "level", "cx", "cy", "cz")
FIELDS = SCALARS + PER_NODE
TRACE_FORMAT_VERSION = 1


@dataclass
class Trace:
    role: str = "trace"
    ship: str = ""
    sim: str = ""
    title: str = ""
    producer: str = ""
    engine: str = ""
    dt: float = 0.0
    nodes: int = 0
    denseSteps: int = 0
    sampleEvery: float = 0.0
    steps: int = 0
    tEnd: float = 0.0
    foundered: bool = False
    founderT: float = -1.0
    events: list = field(default_factory=list)      # (t, id, label)
    dense: np.ndarray = None                         # [n, 10 + 5 * nodes]
    samples: np.ndarray = None


def _columns_to_records(v: dict, nodes: int, path: str, name: str) -> np.ndarray:
    n = int(v["n"])
    rec = np.zeros((n, 10 + 5 * nodes))
    for k, f in enumerate(SCALARS):
        arr = np.asarray(v[f], dtype=np.float64)
        if arr.size != n:
            raise RuntimeError(f"{path}: {name}.{f} has {arr.size} values, expected {n}")
        rec[:, k] = arr
    for k, f in enumerate(PER_NODE):
        arr = np.asarray(v[f], dtype=np.float64)
        if arr.size != n * nodes:
            raise RuntimeError(f"{path}: {name}.{f} has {arr.size} values, expected {n * nodes}")
        rec[:, 10 + k * nodes:10 + (k + 1) * nodes] = arr.reshape(n, nodes)
    return rec


def _column(rec: np.ndarray, f: str, nodes: int) -> np.ndarray:
    if f in SCALARS:
        return rec[:, SCALARS.index(f)]
    k = PER_NODE.index(f)
    return rec[:, 10 + k * nodes:10 + (k + 1) * nodes].reshape(-1)


def load_trace(path: str, verify: bool = True) -> Trace:
    """io::load_trace (engine/src/trace.cpp:106-141); reads .json or .json.gz."""
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rb") as fh:
        doc = json.loads(fh.read())
    if doc.get("format") != "sinksim.trace":
        raise RuntimeError(f"{path}: format is not sinksim.trace")
    if doc.get("formatVersion") != TRACE_FORMAT_VERSION:
        raise RuntimeError(f"{path}: unsupported trace formatVersion")
    meta = doc.get("meta", {})
    tr = Trace(role=doc.get("role", "trace"), ship=doc.get("ship", ""), sim=doc.get("sim", ""),
               title=meta.get("title", ""), producer=meta.get("producer", ""), engine=meta.get("engine", ""),
               dt=float(doc["dt"]), nodes=int(doc["nodes"]), denseSteps=int(doc["denseSteps"]),
               sampleEvery=float(doc["sampleEvery"]), steps=int(doc["steps"]), tEnd=float(doc["tEnd"]),
               foundered=bool(doc["foundered"]), founderT=float(doc["founderT"]))
    tr.events = [(float(e["t"]), e["id"], e["label"]) for e in doc["events"]]
    tr.dense = _columns_to_records(doc["dense"], tr.nodes, path, "dense")
    tr.samples = _columns_to_records(doc["samples"], tr.nodes, path, "samples")
    if verify and "hash" in doc:
        H = doc["hash"]
        chain = Fnv1a64()
        for name, rec in (("dense", tr.dense), ("samples", tr.samples)):
            for f in FIELDS:
                col = _column(rec, f, tr.nodes)
                one = Fnv1a64()
                one.update_array(col)
                chain.update_array(col)
                want = H.get(f"{name}.{f}")
                if isinstance(want, str) and want != one.hex():
                    raise RuntimeError(f"{path}: hash mismatch on {name}.{f}: the JSON round trip was not exact")
        if isinstance(H.get("all"), str) and H["all"] != chain.hex():
            raise RuntimeError(f"{path}: chained hash mismatch")
    return tr


def _num(x: float):
    x = float(x)
    if x != x or x in (float("inf"), float("-inf")):
        return None
    return 0.0 if x == 0 else x


def save_trace(tr: Trace, path: str) -> None:
    """io::save_trace (engine/src/trace.cpp:143-179): columnar records with per-column and chained FNV hashes."""
    doc = {"format": "sinksim.trace", "formatVersion": TRACE_FORMAT_VERSION, "role": tr.role, "ship": tr.ship,
           "sim": tr.sim, "meta": {"title": tr.title, "producer": tr.producer, "engine": tr.engine},
           "dt": tr.dt, "nodes": tr.nodes, "denseSteps": tr.denseSteps, "sampleEvery": tr.sampleEvery,
           "steps": tr.steps, "tEnd": tr.tEnd, "foundered": tr.foundered, "founderT": tr.founderT,
           "events": [{"t": t, "id": i, "label": lab} for t, i, lab in tr.events]}
    H = {}
    chain = Fnv1a64()
    for name, rec in (("dense", tr.dense), ("samples", tr.samples)):
        v = {"n": int(rec.shape[0])}
        for f in FIELDS:
            col = _column(rec, f, tr.nodes)
            v[f] = [_num(x) for x in col]
            one = Fnv1a64()
            one.update_array(col)
            chain.update_array(col)
            H[f"{name}.{f}"] = one.hex()
        doc[name] = v
    H["all"] = chain.hex()
    doc["hash"] = H
    with open(path, "w", encoding="utf-8") as fh:
        json.dump(doc, fh, indent=2)
        fh.write("\n")
