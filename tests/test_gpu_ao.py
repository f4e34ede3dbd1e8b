"""Cavity ambient occlusion on the GPU (ag_ambient_occlusion) against the
oracle (argent-mesh cavity.rs:95-128 restated) and against the closed form of
a floor under a ceiling.  A vertex's count is an integer; it may differ from
the oracle's only where a ray grazes an edge within the last bits of sin/cos
(the device's and the host's libm agree to an ulp or two), so a per-vertex
difference is allowed on at most 0.5 % of vertices and never by more than 2
rays."""
import numpy as np
import pytest

import argent_amd as ag
import oracle as O
from argent_amd import synth

from aocases import ceiling_counts, desc, floor_and_ceiling, groove_ridge, nearest

pytestmark = pytest.mark.gpu


def _ao(ctx, d, samples, distance):
    sc = ag.Scene(ctx, d)
    try:
        return sc.ambient_occlusion(samples, distance)
    finally:
        sc.close()


def test_floor_under_ceiling_closed_form(ctx):
    pos, tris, nrm = floor_and_ceiling(0.5)
    ao = _ao(ctx, desc(pos, tris, nrm), 48, 1.2)
    counts, clear = ceiling_counts(pos, 41 * 41, 48, 1.2, 0.5)
    got = np.rint(ao[:41 * 41] * 48).astype(int)
    assert np.array_equal(got[clear], counts[clear])


def _compare(gpu, ref, samples):
    a, b = np.rint(gpu * samples).astype(int), np.rint(ref * samples).astype(int)
    diff = np.abs(a - b)
    assert diff.max() <= 2, diff.max()
    assert (diff > 0).mean() <= 0.005, (diff > 0).mean()


@pytest.mark.parametrize("samples,distance", [(48, 1.2), (16, 2.5)])
def test_signet_matches_oracle(ctx, samples, distance):
    pos, tris, nrm = synth.signet(seed=2, band_nu=256, band_nv=64, plate_nt=96, plate_np=192)
    d = desc(pos, tris, nrm)
    gpu = _ao(ctx, d, samples, distance)
    ref = O.OracleScene(d).ambient_occlusion(len(pos), samples, distance)
    assert gpu.mean() > 0.01
    _compare(gpu, ref, samples)


def test_groove_and_edge_cases(ctx):
    pos, tris, nrm = groove_ridge(200)
    ao = _ao(ctx, desc(pos, tris, nrm), 48, 1.2)
    assert ao[nearest(pos, 0.0, 0.0)] == 0.0 and ao[nearest(pos, 3.0, 0.0)] == 0.0
    assert ao[nearest(pos, -3.0, 0.0)] > 0.3
    ref = O.OracleScene(desc(pos, tris, nrm)).ambient_occlusion(len(pos), 48, 1.2)
    _compare(ao, ref, 48)
    # a zero normal gives 0; without normals, the smooth vertex normals (cavity.rs:82-87)
    p2, t2, n2 = floor_and_ceiling(0.5)
    n2 = n2.copy()
    n2[0] = 0.0
    ao2 = _ao(ctx, desc(p2, t2, n2), 8, 1.0)
    assert ao2[0] == 0.0 and ao2[1] > 0.0
    gpu = _ao(ctx, desc(pos, tris, None), 48, 1.2)
    ref = O.OracleScene(desc(pos, tris, None)).ambient_occlusion(len(pos), 48, 1.2)
    _compare(gpu, ref, 48)
    assert gpu[nearest(pos, -3.0, 0.0)] > 0.3
