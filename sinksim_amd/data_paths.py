"""Locations of the model files shipped with the package (the reference's titanic64 ship and its scenarios,
MIT-licensed, see data/NOTICE)."""
from __future__ import annotations

import os

from .model import check_pairing, load_ship, load_sim

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
SHIP = os.path.join(DATA, "ships", "titanic", "titanic64.ship.json")
SIMS = os.path.join(DATA, "ships", "titanic", "sims")
OBSERVATIONS = os.path.join(DATA, "observations", "titanic-1912.json")
CATALOG = os.path.join(DATA, "ships", "titanic", "catalog.json")  # the validation catalog (ships/titanic/catalog.json)

_cache: dict = {}


def sim_path(sim_id: str) -> str:
    return os.path.join(SIMS, f"{sim_id}.sim.json")


def titanic(sim_id: str = "titanic"):
    """(ship, sim) of the titanic64 model and one of its scenarios, loaded and hash-checked once per process."""
    if "ship" not in _cache:
        _cache["ship"] = load_ship(SHIP)
    key = "sim:" + sim_id
    if key not in _cache:
        _cache[key] = load_sim(sim_path(sim_id))
        check_pairing(_cache["ship"], _cache[key])
    return _cache["ship"], _cache[key]


def catalog_cases() -> list[dict]:
    """The validation catalog's cases: id, label, the scenario's sim id and tMax (s)."""
    import json
    with open(CATALOG) as f:
        doc = json.load(f)
    return [{"id": c["id"], "label": c["label"], "sim": os.path.basename(c["file"]).replace(".sim.json", ""),
             "t_max": float(c.get("tMaxH", 8)) * 3600.0} for c in doc["cases"]]
