"""Shared crease-normal cases (argent-mesh normals.rs tests :795-931) for the
oracle tests and the GPU parity tests."""
import numpy as np

import oracle as O
from argent_amd import synth


def mesh_hash32(x):
    """argent-mesh util.rs:244-251."""
    x = np.uint32(x)
    x ^= x >> np.uint32(16)
    x = np.uint32((int(x) * 0x7FEB352D) & 0xFFFFFFFF)
    x ^= x >> np.uint32(15)
    x = np.uint32((int(x) * 0x846CA68B) & 0xFFFFFFFF)
    x ^= x >> np.uint32(16)
    return int(x)


def welded_sphere():
    p, t = synth.unweld(*synth.icosphere(5.0, 4))
    q, s, _ = O.weld(p, t, 1e-6)
    return q, s


def inside_out_sphere():
    p, t = synth.icosphere(3.0, 3)
    return p, t[:, [0, 2, 1]].copy()


def hollow():
    po, to = synth.icosphere(5.0, 3)
    pi, ti = synth.icosphere(3.0, 3)
    ti = ti[:, [0, 2, 1]] + len(po)
    return np.concatenate([po, pi]), np.concatenate([to, ti]).astype(np.uint32)


def inconsistent_sphere():
    p, t = synth.icosphere(2.0, 3)
    t = t.copy()
    for i in range(len(t)):
        if mesh_hash32(i) % 7 == 0:
            t[i] = t[i][[0, 2, 1]]
    return p, t


def broken_fan():
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0], [0, -1, 0]], np.float32)
    t = np.array([[0, 1, 2], [0, 3, 4], [0, 4, 5], [0, 6, 1]], np.uint32)
    return p, t


def cases():
    """(name, positions, triangles, crease angle) for the parity tests."""
    out = [("sphere", *welded_sphere(), 50.0), ("cube", *synth.subdivided_cuboid([2.0, 2.0, 2.0], 3), 50.0),
           ("faceted", *synth.icosphere(1.0, 1), 0.0), ("inside-out", *inside_out_sphere(), 50.0),
           ("hollow", *hollow(), 50.0), ("inconsistent", *inconsistent_sphere(), 50.0),
           ("broken-fan", *broken_fan(), 50.0)]
    p, t, _ = synth.torus_ring()
    out.append(("torus-ring", p, t, 50.0))
    out.append(("torus-ring-20", p, t, 20.0))
    return out
