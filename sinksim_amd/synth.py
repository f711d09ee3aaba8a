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
