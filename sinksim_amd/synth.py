"""Synthetic compiled models built in memory.

box_barge: the reference's kernel-test rig (tests/test_kernel.cpp:15-138), a box L x B x H split into `parts`
nodes along its length, columns dx by dy, floating at draft T, whose hydrostatics are known in closed form.
"""
from __future__ import annotations

import math

import numpy as np

from .model import CompiledShip, CompiledSim

RHO, G, DT = 1025.0, 9.81, 0.25


def box_barge(L=100.0, B=20.0, H=10.0, T=4.0, dx=5.0, dy=1.0, parts=2, shared_surface=False, connections=()):
    """(ship, sim) of the rig. connections: (a, b, law, x, y, z, area_or_width, coef) tuples, a = -1 for the sea."""
    nx, ny = int(round(L / dx)), int(round(B / dy))
    x, y, dxdy, zlo, zhi, segCol, segA, segE, segStart, segCount = [], [], [], [], [], [], [], [], [], []
    mu, vmax, amin, aov = [], [], [], []
    for k in range(parts):
        segStart.append(len(segCol))
        footprint = 0.0
        for i in range(nx):
            if i * parts // nx != k:
                continue
            for j in range(ny):
                col = len(x)
                x.append(-L / 2 + (i + 0.5) * dx)
                y.append(-B / 2 + (j + 0.5) * dy)
                dxdy.append(dx * dy)
                zlo.append(0.0)
                zhi.append(H)
                segCol.append(col)
                segA.append(0.0)
                segE.append(H)
                footprint += dx * dy
        segCount.append(len(segCol) - segStart[-1])
        mu.append(1.0)
        vmax.append(footprint * H)
        amin.append(4.0)
        aov.append(max(1.0, 0.03 * footprint))
    f64 = lambda v: np.asarray(v, dtype=np.float64)
    i32 = lambda v: np.asarray(v, dtype=np.int32)
    u8 = lambda v: np.asarray(v, dtype=np.uint8)
    ncol = len(x)
    ship = CompiledShip("box", "box barge", "", f64(x), f64(y), f64(dxdy), f64(zlo), f64(zhi), i32([0] * ncol),
                        i32([0] * ncol), [f"part{k}" for k in range(parts)], f64(mu), f64(vmax), f64(amin), f64(aov),
                        i32(segStart), i32(segCount), i32(segCol), f64(segA), f64(segE))
    if shared_surface:
        gS, gC, gN, gM = [0], [parts], list(range(parts)), [1]
    else:
        gS, gC, gN, gM = list(range(parts)), [1] * parts, list(range(parts)), [0] * parts
    ng = len(gS)
    conns = list(connections)
    nc = len(conns)
    col = lambda k: [c[k] for c in conns]
    ms = RHO * L * B * T
    mh0, Ith0, Iph0 = 2 * ms, 2 * ms * (0.25 * L) ** 2, 1.2 * ms * (0.38 * B) ** 2
    Awp, V = L * B, L * B * T
    BML, BMT = (L ** 3 * B / 12) / V, (L * B ** 3 / 12) / V
    GM = T / 2 + BMT
    Kz, Kth, Kph = RHO * G * Awp, RHO * G * V * BML, RHO * G * V * GM
    cDz, cDth, cDph = 2 * 0.6 * math.sqrt(Kz * mh0), 2 * 0.6 * math.sqrt(Kth * Ith0), 2 * 0.6 * math.sqrt(Kph * Iph0)
    body = f64([ms, 0, 0, 0, 0, 0, 0, mh0, Ith0, Iph0, cDz, cDth, cDph, 0, 0, 0])
    zeros = f64([0.0] * parts)
    sim = CompiledSim(
        "box", "box barge", "box", "", "", RHO, G, DT, i32(gS), i32(gC), i32(gN), u8(gM), f64([0] * ng), f64([0] * ng),
        f64([0] * ng), 0.05, i32(col(0)), i32(col(1)), u8(col(2)), f64(col(3)), f64(col(4)), f64(col(5)), f64(col(6)),
        f64(col(7)), u8([0] * nc), i32([-1] * nc), u8([1] * nc), f64([0] * nc), i32([-1] * nc), i32([-1] * nc), ["k0"],
        0.05, [], [], [], [], f64([]), f64([]), f64([]), u8([]), f64([-1e9, 10, -1, 1e9]), body, f64([L / 2, -L / 2]),
        f64([0, -T, 0, 0, 0, 0, 0]), zeros.copy(), zeros.copy(), zeros.copy(), zeros.copy(), zeros.copy())
    return ship, sim


# ------------------------------------------------------------------------------------------- uneven synthetic hulls

KIND_NAMES = ["door", "over", "down", "xlow", "xmerge", "xup", "top", "breach", "open"]


def synthetic_hull(seed: int, columns: int | None = None, spaces: int | None = None, connections: int | None = None,
                   rho: float = 1025.0, g: float = 9.81, dt: float = 0.25, hashed: bool = False):
    """A seeded ship and flooding scenario of the kind the uneven batch runs: length 100-300 m, parabolic sections,
    `columns` (1,000-16,000) hull columns, `spaces` (8-128) nodes, `connections` (20-600; at least the structural
    minimum) doors, bulkhead-top weirs, deck openings, cross-flow ducts, deck-edge weirs to the sea and 1-6 breaches
    of 0.01-1 m^2 below the waterline. Unset sizes are drawn from the seed (connections log-uniform).

    Compartments are bounded by transverse bulkheads; each is split port/starboard and lower/upper at a flat deck, so
    a compartment holds four nodes. Lower pairs share a surface always, above a threshold point, never, or are
    solved singly; some compartments join all four nodes in one group (a shared surface over N = 4 members). Extra
    connections beyond the structural ones are further doors between neighbouring compartments (closed with
    probability 0.7), deck openings and upper cross-flow ducts. The body floats level at its design draft: mass =
    displaced mass there, centre of gravity above the centre of buoyancy with a positive metacentric height, added
    mass and damping as the reference's box-barge rig. `hashed` fills the FNV hashes of the in-memory objects
    (write_compiled computes its own)."""
    rng = np.random.default_rng(seed)
    L = float(rng.uniform(100.0, 300.0))
    B = L / float(rng.uniform(7.5, 10.0))
    D = B * float(rng.uniform(0.55, 0.8))
    T = D * float(rng.uniform(0.45, 0.62))
    ncol_target = int(columns if columns is not None else rng.integers(1000, 16001))
    nsp_target = int(spaces if spaces is not None else rng.integers(8, 129))
    nconn_target = int(connections if connections is not None else round(math.exp(rng.uniform(math.log(20.0), math.log(600.0)))))
    ncomp = max(2, nsp_target // 4)
    # grid: the plan fills about 0.78 of the bounding box
    cells = ncol_target / 0.78
    ny = max(6, int(round(math.sqrt(cells * B / L))))
    nx = max(2 * ncomp, int(round(cells / ny)))
    dx, dy = L / nx, B / ny
    zdeck = D * float(rng.uniform(0.4, 0.55))

    def half_breadth(x):
        return 0.5 * B * np.sqrt(np.maximum(0.0, 1.0 - np.abs(2.0 * np.asarray(x) / L) ** 2.4))

    # columns, row by row along the length (x outer, y inner)
    X, Y = np.meshgrid(-L / 2 + (np.arange(nx) + 0.5) * dx, -B / 2 + (np.arange(ny) + 0.5) * dy, indexing="ij")
    X, Y = X.reshape(-1), Y.reshape(-1)
    hb = half_breadth(X)
    keep = (np.abs(Y) < hb) & (hb > 0)
    rise = 0.35 * D * np.abs(2.0 * X / L) ** 5
    with np.errstate(divide="ignore", invalid="ignore"):
        zlo = rise + (D - rise) * (np.abs(Y) / np.where(hb > 0, hb, 1.0)) ** 2
    zhi = D + 0.02 * L * (2.0 * X / L) ** 2
    keep &= (zhi - zlo) >= 0.05
    bulkheads = np.linspace(-L / 2, L / 2, ncomp + 1)
    cx_, cy_, czlo, czhi = X[keep], Y[keep], zlo[keep], zhi[keep]
    cA = np.full(cx_.shape, dx * dy)
    ccomp = np.minimum(ncomp - 1, np.searchsorted(bulkheads, cx_, side="right") - 1)
    cside = (cy_ >= 0).astype(np.int64)
    ncol = cx_.size

    # nodes: compartment k -> (P lower, P upper, S lower, S upper); side 1 = port (y >= 0)
    node_of = {}
    labels, mu, vmax, amin, aov, segStart, segCount, ncent = [], [], [], [], [], [], [], []
    segColL, segAL, segEL = [], [], []
    nseg = 0
    for k in range(ncomp):
        for side in (1, 0):
            cols = np.flatnonzero((ccomp == k) & (cside == side))
            for upper in (False, True):
                n = len(labels)
                node_of[(k, side, upper)] = n
                labels.append(f"C{k}/{'P' if side else 'S'}/{'upper' if upper else 'lower'}")
                perm = float(rng.uniform(0.85, 0.98))
                if upper:
                    a, e = np.maximum(czlo[cols], zdeck), czhi[cols]
                else:
                    a, e = czlo[cols], np.minimum(czhi[cols], zdeck)
                ok = (e - a) > 1e-3
                c, a, e = cols[ok], a[ok], e[ok]
                if c.size == 0:
                    raise RuntimeError("synthetic hull produced an empty node; change the seed")
                segStart.append(nseg); segCount.append(int(c.size)); nseg += int(c.size)
                segColL.append(c); segAL.append(a); segEL.append(e)
                v = cA[c] * (e - a)
                vol, footprint = float(v.sum()), float(cA[c].sum())
                mu.append(perm); vmax.append(perm * vol); amin.append(max(1.0, 0.01 * footprint))
                aov.append(max(1.0, 0.03 * footprint))
                ncent.append(np.array([(v * cx_[c]).sum(), (v * cy_[c]).sum(), (v * 0.5 * (a + e)).sum()]) / vol)
    nn = len(labels)
    segCol, segA, segE = np.concatenate(segColL), np.concatenate(segAL), np.concatenate(segEL)

    # groups
    gS, gC, gN, gM, gpx, gpy, gpz = [], [], [], [], [], [], []

    def group(members, mode, p=(0.0, 0.0, 0.0)):
        gS.append(len(gN)); gC.append(len(members)); gN.extend(members); gM.append(mode)
        gpx.append(p[0]); gpy.append(p[1]); gpz.append(p[2])
        return len(gC) - 1

    lower_group, upper_group = {}, {}
    for k in range(ncomp):
        xc = 0.5 * (bulkheads[k] + bulkheads[k + 1])
        lp, up, ls, us = node_of[(k, 1, False)], node_of[(k, 1, True)], node_of[(k, 0, False)], node_of[(k, 0, True)]
        r = rng.uniform()
        if r < 0.1:
            lower_group[k] = upper_group[k] = group([lp, ls, up, us], 1)
            continue
        if r < 0.55:
            lower_group[k] = group([lp, ls], 1)
        elif r < 0.8:
            lower_group[k] = group([lp, ls], 2, (xc, 0.0, 0.3 * zdeck))
        elif r < 0.9:
            lower_group[k] = group([lp, ls], 3)
        else:
            group([lp], 0); group([ls], 0)
        upper_group[k] = group([up, us], 1)

    # connections: (a, b, law, x, y, z, area/width, coef, kind, idx, en, tOn, skipGroup, monitor)
    conns = []
    K = {name: i for i, name in enumerate(KIND_NAMES)}

    def pair_skip(gr):  # a cross connection inside a group is skipped while that group is merged
        return gr if gr is not None and gr >= 0 else -1

    for k in range(ncomp):
        xc = 0.5 * (bulkheads[k] + bulkheads[k + 1])
        hb = float(half_breadth(xc))
        lp, up, ls, us = node_of[(k, 1, False)], node_of[(k, 1, True)], node_of[(k, 0, False)], node_of[(k, 0, True)]
        skip = pair_skip(lower_group.get(k))
        conns.append((lp, ls, 0, xc, 0.0, 0.6, float(rng.uniform(0.5, 3.0)), 0.6, K["xlow"], k, 1, 0.0, skip, -1))
        for side, lo_, up_ in ((1, lp, up), (0, ls, us)):
            ysgn = 1.0 if side else -1.0
            conns.append((up_, lo_, 0, xc, ysgn * 0.25 * hb, zdeck, float(rng.uniform(0.2, 2.0)), 0.6, K["down"], k, 1, 0.0, -1, -1))
            conns.append((-1, up_, 1, xc, ysgn * 0.6 * hb, D + 0.02 * L * (2 * xc / L) ** 2, float(rng.uniform(1.0, 5.0)), 0.6, K["top"], k,
                          1, 0.0, -1, -1))
    for k in range(ncomp - 1):
        xb = float(bulkheads[k + 1])
        hb = float(half_breadth(xb))
        top = zdeck + (D - zdeck) * float(rng.uniform(0.3, 0.8))
        for side in (1, 0):
            ysgn = 1.0 if side else -1.0
            a_lo, b_lo = node_of[(k, side, False)], node_of[(k + 1, side, False)]
            a_up, b_up = node_of[(k, side, True)], node_of[(k + 1, side, True)]
            conns.append((a_lo, b_lo, 0, xb, ysgn * 0.25 * hb, 1.0, float(rng.uniform(0.5, 4.0)), 0.6, K["door"], k + 2,
                          int(rng.uniform() < 0.3), 0.0, -1, -1))
            conns.append((a_up, b_up, 1, xb, ysgn * 0.5 * hb, top, float(rng.uniform(2.0, 10.0)), 0.6, K["over"], k, 1, 0.0, -1, k))
    nbreach = int(rng.integers(1, 7))
    # extra connections up to the target count (the breaches come last)
    extra = max(0, nconn_target - len(conns) - nbreach)
    for m in range(extra):
        kind = m % 3
        if kind == 0 and ncomp > 1:  # a further door between neighbouring lower spaces
            k = int(rng.integers(ncomp - 1))
            side = int(rng.integers(2))
            xb = float(bulkheads[k + 1])
            hb = float(half_breadth(xb))
            conns.append((node_of[(k, side, False)], node_of[(k + 1, side, False)], 0, xb, (1 if side else -1) * float(rng.uniform(0.1, 0.8)) * hb,
                          float(rng.uniform(0.5, max(0.6, zdeck - 0.5))), float(rng.uniform(0.5, 4.0)), 0.6, K["door"], k + 2,
                          int(rng.uniform() < 0.3), 0.0, -1, -1))
        elif kind == 1:  # a further deck opening
            k = int(rng.integers(ncomp))
            side = int(rng.integers(2))
            xc = float(rng.uniform(bulkheads[k], bulkheads[k + 1]))
            conns.append((node_of[(k, side, True)], node_of[(k, side, False)], 0, xc, (1 if side else -1) * 0.3 * float(half_breadth(xc)), zdeck,
                          float(rng.uniform(0.2, 2.0)), 0.6, K["down"], k, 1, 0.0, -1, -1))
        else:  # an upper cross-flow duct
            k = int(rng.integers(ncomp))
            xc = float(rng.uniform(bulkheads[k], bulkheads[k + 1]))
            conns.append((node_of[(k, 1, True)], node_of[(k, 0, True)], 0, xc, 0.0, zdeck + 0.3 * (D - zdeck), float(rng.uniform(0.5, 3.0)), 0.6,
                          K["xup"], k, 1, 0.0, pair_skip(upper_group.get(k)), -1))
    for m in range(nbreach):
        k = int(rng.integers(max(1, ncomp // 3))) if m else 0
        side = int(rng.integers(2))
        xc = float(rng.uniform(bulkheads[k], bulkheads[k + 1]))
        z = float(max(0.3, T - rng.uniform(0.5, 0.8 * T)))
        tOn = 0.0 if m == 0 or rng.uniform() < 0.6 else float(rng.uniform(0, 1800))
        conns.append((-1, node_of[(k, side, False)], 0, xc, (1 if side else -1) * 0.9 * float(half_breadth(xc)), z, float(rng.uniform(0.01, 1.0)),
                      0.6, K["breach"], m, 1, tOn, -1, -1))
    nc = len(conns)
    col = lambda q, dtp: np.array([c[q] for c in conns], dtype=dtp)

    # hydrostatics at the design draft, and the body floating level there
    wet = np.clip(T - czlo, 0.0, czhi - czlo)
    V = float((cA * wet).sum())
    xB = float((cA * wet * cx_).sum() / V)
    KB = float((cA * wet * (czlo + 0.5 * wet)).sum() / V)
    wp = czlo < T
    IT = float((cA[wp] * cy_[wp] ** 2).sum())
    IL = float((cA[wp] * (cx_[wp] - xB) ** 2).sum())
    Awp = float(cA[wp].sum())
    BMT, BML = IT / V, IL / V
    KG = min(0.6 * D, KB + BMT - 0.04 * B)
    GM = KB + BMT - KG
    ms = rho * V
    mh0, Ith0, Iph0 = 2 * ms, 2 * ms * (0.25 * L) ** 2, 1.2 * ms * (0.38 * B) ** 2
    Kz, Kth, Kph = rho * g * Awp, rho * g * V * BML, rho * g * V * GM
    body = np.array([ms, KG, xB, 0.0, KG, xB, 0.0, mh0, Ith0, Iph0, 2 * 0.6 * math.sqrt(Kz * mh0), 2 * 0.6 * math.sqrt(Kth * Ith0),
                     2 * 0.12 * math.sqrt(Kph * Iph0), 0.0, 0.0, 0.0])
    zO = KG - T
    floor = np.minimum.reduceat(segA, np.asarray(segStart)) - T
    cen = np.array(ncent)
    f64 = lambda v: np.asarray(v, dtype=np.float64)
    i32 = lambda v: np.asarray(v, dtype=np.int32)
    u8 = lambda v: np.asarray(v, dtype=np.uint8)
    sid = f"synth{seed}"
    ship = CompiledShip(sid, f"synthetic hull {seed}", "", cx_, cy_, cA, czlo, czhi, i32(ccomp), i32(cside), labels, f64(mu), f64(vmax),
                        f64(amin), f64(aov), i32(segStart), i32(segCount), i32(segCol), f64(segA), f64(segE))
    mk = [(0.45 * L, 0.0, D + 0.02 * L * 0.81, 0), (0.0, 0.0, D, 0), (-0.48 * L, 0.0, 0.5, 1)]
    sim = CompiledSim(
        "flood", f"synthetic flooding {seed}", sid, "", "", rho, g, dt, i32(gS), i32(gC), i32(gN), u8(gM), f64(gpx), f64(gpy), f64(gpz), 0.05,
        col(0, np.int32), col(1, np.int32), col(2, np.uint8), col(3, np.float64), col(4, np.float64), col(5, np.float64),
        col(6, np.float64), col(7, np.float64), col(8, np.uint8), col(9, np.int32), col(10, np.uint8), col(11, np.float64),
        col(12, np.int32), col(13, np.int32), list(KIND_NAMES), 0.05, [f"over{k}" for k in range(ncomp - 1)],
        [f"Water flows over bulkhead {k}" for k in range(ncomp - 1)], ["bow", "midship", "stern"],
        ["Bow deck goes under", "Midship deck goes under", "Stern keel clears the water"], f64([m[0] for m in mk]),
        f64([m[1] for m in mk]), f64([m[2] for m in mk]), u8([m[3] for m in mk]), f64([-0.5, 1.2, 1.0, 10.0]), body,
        f64([L / 2, -L / 2]), f64([0.0, zO, 0.0, 0.0, 0.0, 0.0, 0.0]), np.zeros(nn), floor, cen[:, 0].copy(), cen[:, 1].copy(),
        cen[:, 2].copy())
    if hashed:
        _hash_in_place(ship, sim)
    return ship, sim


def _ship_arrays(ship):
    return [("columns.x", ship.x), ("columns.y", ship.y), ("columns.dxdy", ship.dxdy), ("columns.zlo", ship.zlo),
            ("columns.zhi", ship.zhi), ("columns.zone", ship.zone), ("columns.side", ship.side), ("nodes.mu", ship.mu),
            ("nodes.vmax", ship.vmax), ("nodes.amin", ship.amin), ("nodes.aov", ship.aov), ("nodes.segStart", ship.segStart),
            ("nodes.segCount", ship.segCount), ("segments.col", ship.segCol), ("segments.a", ship.segA), ("segments.e", ship.segE)]


def _sim_arrays(sim):
    return [("groups.nodeStart", sim.gNodeStart), ("groups.nodeCount", sim.gNodeCount), ("groups.nodes", sim.gNodes),
            ("groups.mode", sim.gMode), ("groups.px", sim.gpx), ("groups.py", sim.gpy), ("groups.pz", sim.gpz),
            ("connections.a", sim.ca), ("connections.b", sim.cb), ("connections.type", sim.cLaw), ("connections.x", sim.cx),
            ("connections.y", sim.cy), ("connections.z", sim.cz), ("connections.area", sim.cArea),
            ("connections.coef", sim.cCoef), ("connections.kind", sim.cKind), ("connections.idx", sim.cIdx),
            ("connections.en", sim.cEn), ("connections.tOn", sim.cTOn), ("connections.skipGroup", sim.cSkipGroup),
            ("connections.monitor", sim.cMonitor), ("marks.x", sim.mkx), ("marks.y", sim.mky), ("marks.z", sim.mkz),
            ("marks.when", sim.mkWhen), ("state0.vol", sim.vol0), ("state0.level", sim.level0), ("state0.cx", sim.cx0),
            ("state0.cy", sim.cy0), ("state0.cz", sim.cz0)]


def _hashes(arrays):
    from .fnv import Fnv1a64
    out, chain = {}, Fnv1a64()
    for name, a in arrays:
        one = Fnv1a64()
        one.update_array(a)
        chain.update_array(a)
        out[name] = one.hex()
    out["all"] = chain.hex()
    return out


def _hash_in_place(ship, sim):
    ship.hashAll = _hashes(_ship_arrays(ship))["all"]
    sim.shipHash = ship.hashAll
    sim.hashAll = _hashes(_sim_arrays(sim))["all"]


def _nums(a):
    return [float(v) if isinstance(v, (float, np.floating)) else int(v) for v in np.asarray(a).tolist()]


def write_compiled(ship, sim, ship_path: str, sim_path: str) -> None:
    """The pair as sinksim.compiled-ship / sinksim.compiled-sim v1 files with every array hashed (docs/FORMATS.md),
    readable by the reference's loaders and by this package's."""
    import json
    sh = _hashes(_ship_arrays(ship))
    ship_doc = {
        "format": "sinksim.compiled-ship", "formatVersion": 1, "id": ship.id,
        "meta": {"title": ship.title, "generator": "sinksim_amd.synth", "engine": "python"},
        "columns": {"n": ship.columns, "x": _nums(ship.x), "y": _nums(ship.y), "dxdy": _nums(ship.dxdy), "zlo": _nums(ship.zlo),
                    "zhi": _nums(ship.zhi), "zone": _nums(ship.zone), "side": _nums(ship.side)},
        "nodes": {"n": ship.nodes, "label": list(ship.nodeLabel), "mu": _nums(ship.mu), "vmax": _nums(ship.vmax),
                  "amin": _nums(ship.amin), "aov": _nums(ship.aov), "segStart": _nums(ship.segStart), "segCount": _nums(ship.segCount)},
        "segments": {"n": ship.segments, "col": _nums(ship.segCol), "a": _nums(ship.segA), "e": _nums(ship.segE)},
        "hash": sh,
    }
    b = sim.body
    mh = _hashes(_sim_arrays(sim))
    sim_doc = {
        "format": "sinksim.compiled-sim", "formatVersion": 1, "id": sim.id, "ship": ship.id, "shipHash": sh["all"],
        "meta": {"title": sim.title, "generator": "sinksim_amd.synth", "engine": "python"},
        "constants": {"rho": sim.rho, "g": sim.g}, "scheme": {"dt": sim.dt}, "params": {},
        "groups": {"n": sim.groups, "nodeStart": _nums(sim.gNodeStart), "nodeCount": _nums(sim.gNodeCount), "nodes": _nums(sim.gNodes),
                   "mode": _nums(sim.gMode), "px": _nums(sim.gpx), "py": _nums(sim.gpy), "pz": _nums(sim.gpz), "margin": sim.gMargin},
        "connections": {"n": sim.connections, "a": _nums(sim.ca), "b": _nums(sim.cb), "type": _nums(sim.cLaw), "x": _nums(sim.cx),
                        "y": _nums(sim.cy), "z": _nums(sim.cz), "area": _nums(sim.cArea), "coef": _nums(sim.cCoef),
                        "kind": _nums(sim.cKind), "idx": _nums(sim.cIdx), "en": _nums(sim.cEn), "tOn": _nums(sim.cTOn),
                        "skipGroup": _nums(sim.cSkipGroup), "monitor": _nums(sim.cMonitor), "kindNames": list(sim.kindNames)},
        "monitors": {"n": sim.monitors, "threshold": sim.monitorThreshold, "id": list(sim.monitorId), "label": list(sim.monitorLabel)},
        "marks": {"n": sim.marks, "id": list(sim.markId), "label": list(sim.markLabel), "x": _nums(sim.mkx), "y": _nums(sim.mky),
                  "z": _nums(sim.mkz), "when": _nums(sim.mkWhen)},
        "founder": dict(zip(("hullTopBelow", "pitchAbsAbove", "buoyancyBelow", "minT"), _nums(sim.founder))),
        "body": {"ms": float(b[0]), "KG": float(b[1]), "g0": _nums(b[2:5]), "XO": float(b[5]), "YO": float(b[6]),
                 **dict(zip(("mh0", "Ith0", "Iph0", "cDz", "cDth", "cDph", "qz", "qth", "qph"), _nums(b[7:16])))},
        "readout": {"xFP": float(sim.readout[0]), "xAP": float(sim.readout[1])},
        "state0": {**dict(zip(("t", "zO", "pitch", "roll", "vz", "wth", "wph"), _nums(sim.state0))), "vol": _nums(sim.vol0),
                   "level": _nums(sim.level0), "cx": _nums(sim.cx0), "cy": _nums(sim.cy0), "cz": _nums(sim.cz0)},
        "hash": mh,
    }
    with open(ship_path, "w") as f:
        json.dump(ship_doc, f)
    with open(sim_path, "w") as f:
        json.dump(sim_doc, f)
