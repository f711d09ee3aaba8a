"""Writes tests/golden/titanic_batch1024.npz: the reference CPU engine's results for every instance of the 1,024-
instance titanic64 batch (sinksim_amd.workloads.titanic_batch_sweep, seed 20261008, 4 h cap, curves every 30 s), in
the portable and portable32 numerics.

The results come from the reference's own sweep runner (apps/sinksim_sweep.cpp), built from the reference's
sources by `make -C oracle ref` into oracle/_ref/sinksim_sweep. Per numerics and instance the fixture keeps the
step count, the founder flag and time, the end time, the final readouts, the events (time and an index into a
list of event ids), the curve length and a 64-bit BLAKE2b digest of the curve records (8-byte little-endian
doubles, negative zero written as zero), so that the whole batch is pinned in about 200 KB.

    python tests/golden/make_batch1024.py [--threads 8]
"""
from __future__ import annotations

import argparse
import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from sinksim_amd import sweep as S  # noqa: E402
from sinksim_amd.data_paths import SHIP, sim_path, titanic  # noqa: E402
from sinksim_amd.workloads import titanic_batch_sweep  # noqa: E402

OUT = os.path.join(HERE, "titanic_batch1024.npz")
REF = os.path.join(ROOT, "oracle", "_ref", "sinksim_sweep")


def curve_digest(curve: np.ndarray) -> str:
    a = np.where(curve == 0.0, 0.0, np.asarray(curve, dtype=np.float64)).astype("<f8")
    return hashlib.blake2b(np.ascontiguousarray(a).tobytes(), digest_size=8).hexdigest()


def pack(res: S.Results, names: list[str], prefix: str) -> dict:
    n = len(res.instances)
    out = {f"{prefix}steps": np.zeros(n, np.int64), f"{prefix}foundered": np.zeros(n, np.uint8),
           f"{prefix}founderT": np.zeros(n), f"{prefix}tEnd": np.zeros(n), f"{prefix}final": np.zeros((n, 7)),
           f"{prefix}curveLen": np.zeros(n, np.int32), f"{prefix}curveDigest": np.zeros(n, "<U16"),
           f"{prefix}evCount": np.zeros(n, np.int32)}
    evT, evName = [], []
    for i, r in enumerate(res.instances):
        out[f"{prefix}steps"][i] = r.steps
        out[f"{prefix}foundered"][i] = r.foundered
        out[f"{prefix}founderT"][i] = r.founderT
        out[f"{prefix}tEnd"][i] = r.tEnd
        out[f"{prefix}final"][i] = r.curve[-1]
        out[f"{prefix}curveLen"][i] = r.curve.shape[0]
        out[f"{prefix}curveDigest"][i] = curve_digest(r.curve)
        out[f"{prefix}evCount"][i] = len(r.events)
        for t, eid, _ in r.events:
            if eid not in names:
                names.append(eid)
            evT.append(t)
            evName.append(names.index(eid))
    out[f"{prefix}evT"] = np.asarray(evT, np.float64)
    out[f"{prefix}evName"] = np.asarray(evName, np.int32)
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    if not os.path.exists(REF):
        sys.exit(f"{REF} is missing: make -C oracle ref")
    _, sim = titanic("titanic")
    sw = titanic_batch_sweep(sim, count=1024)
    names: list[str] = []
    arrays = {}
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "batch.sweep.json")
        S.save_sweep(sw, path, sim)
        for numerics in ("portable", "portable32"):
            res_path = os.path.join(d, f"{numerics}.json")
            subprocess.run([REF, SHIP, sim_path("titanic"), path, "--numerics", numerics, "--threads", str(args.threads),
                            "--results", res_path], check=True)
            res = S.load_results(res_path)
            assert [r.id for r in res.instances] == [x.id for x in sw.instances]
            arrays.update(pack(res, names, f"{numerics}_"))
    arrays["ids"] = np.asarray([x.id for x in sw.instances], "<U8")
    arrays["evNames"] = np.asarray(names)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT}: {len(sw.instances)} instances, {len(names)} distinct event ids")


if __name__ == "__main__":
    main()
