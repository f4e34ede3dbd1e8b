"""Crease-aware normals on the GPU (ag_mesh_normals) against the oracle's
restatement of normals.rs: positions bit-exact, triangles and report counts
exact, normals within 1e-6 (the corner angles come from the device's f64
atan2 and the host's; everything else is the same f64 arithmetic in the same
order).  Cases: the reference's own normals tests (sphere, creased cube,
faceted, inside-out shell, hollow cavity, inconsistent winding, broken fan),
the torus ring at two crease angles, and the whole ingest chain (binary STL,
weld, normals) of the 1.2 M-triangle signet."""
import numpy as np
import pytest

import argent_amd as ag
import oracle as O
from argent_amd import synth
from normalcases import cases

pytestmark = pytest.mark.gpu


def _same(ctx, m, want, tag):
    wp, wn, wt, wr = want
    gp, gt = m.read()
    gn = m.read_normals()
    assert gp.shape == wp.shape, (tag, gp.shape, wp.shape)
    assert np.array_equal(gp.view(np.uint32), wp.view(np.uint32)), tag
    assert np.array_equal(gt, wt), tag
    err = float(np.abs(gn - wn).max()) if len(gn) else 0.0
    assert err <= 1e-6, (tag, err)


@pytest.mark.parametrize("case", cases(), ids=lambda c: c[0])
def test_normals_match_oracle(ctx, case):
    tag, p, t, crease = case
    m = ag.Mesh.from_arrays(ctx, p, t)
    r = m.normals(crease)
    want = O.normals(p, t, crease)
    for k, v in want[3].items():
        assert r[k] == v, (tag, k, r[k], v)
    _same(ctx, m, want, tag)
    m.close()


def test_signet_ingest_chain(ctx):
    pos, tris, _ = synth.signet()
    data = synth.to_binary_stl(pos, tris)
    m, rep = ag.load_stl(ctx, data)
    r = m.normals(50.0)
    wp, wt, _ = O.load_stl(data)
    want = O.normals(wp, wt, 50.0)
    for k, v in want[3].items():
        assert r[k] == v, (k, r[k], v)
    _same(ctx, m, want, "signet")
    print("signet normals report", r)
    m.close()


def test_weld_drops_normals(ctx):
    p, t = synth.icosphere(1.0, 2)
    m = ag.Mesh.from_arrays(ctx, p, t)
    m.normals(30.0)
    m.weld(1e-6)
    with pytest.raises(ag.ArgentError):
        m.read_normals()
    m.close()
