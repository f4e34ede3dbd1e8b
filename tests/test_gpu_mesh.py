"""Mesh ingest on the GPU (binary STL parse, invalid-triangle filter, weld,
compact) against the oracle's restatement of io/stl.rs, io.rs and weld.rs.
Index and byte work: positions must match bit for bit, triangles and report
counts exactly."""
import struct

import numpy as np
import pytest

import argent_amd as ag
import oracle as O
from argent_amd import synth, workloads
from meshcases import clustered, stl_bytes, stl_with_bad

pytestmark = pytest.mark.gpu

COUNTS = ("input_vertices", "output_vertices", "exact_merges", "tolerance_merges", "dropped_triangles")


def _same(got, want, tag=""):
    gp, gt = got
    wp, wt = want
    assert gp.shape == wp.shape, (tag, gp.shape, wp.shape)
    assert np.array_equal(gp.view(np.uint32), wp.view(np.uint32)), tag
    assert np.array_equal(gt, wt), tag


def _weld_both(ctx, p, t, rel):
    m = ag.Mesh.from_arrays(ctx, p, t)
    r = m.weld(rel)
    got = m.read()
    m.close()
    wp, wt, wr = O.weld(p, t, rel)
    _same(got, (wp, wt), rel)
    for k in COUNTS:
        assert r[k] == wr[k], (k, r[k], wr[k])
    assert np.float32(r["tolerance_mm"]) == np.float32(wr["tolerance_mm"])
    return r


def test_cube_soup(ctx):
    p, t = synth.unweld(*synth.cuboid([2.0, 2.0, 2.0]))
    r = _weld_both(ctx, p, t, 1e-6)
    assert r["output_vertices"] == 8 and r["exact_merges"] == 28


def test_near_duplicates(ctx):
    p, t = synth.unweld(*synth.cuboid([2.0, 2.0, 2.0]))
    rng = np.random.default_rng(5)
    p[:, 0] += ((rng.random(len(p)) - 0.5) * 2e-7).astype(np.float32)
    r = _weld_both(ctx, p, t, 1e-6)
    assert r["output_vertices"] == 8 and r["tolerance_merges"] > 0
    _weld_both(ctx, p, t, 0.0)


def test_collapsed_and_unreferenced(ctx):
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1e-9, 0], [5, 5, 5]], np.float32)
    t = np.array([[0, 1, 2], [0, 1, 3], [1, 3, 2]], np.uint32)
    r = _weld_both(ctx, p, t, 1e-6)
    assert r["dropped_triangles"] == 2


@pytest.mark.parametrize("seed", range(6))
def test_clustered(ctx, seed):
    p, t = clustered(seed)
    for rel in (1e-3, 0.0, 3e-3):
        _weld_both(ctx, p, t, rel)


def test_large_clusters_cross_cells(ctx):
    # 40k clustered points, thousands of tolerance merges and cell crossings
    p, t = clustered(11, n_centres=8000, per=(1, 8))
    r = _weld_both(ctx, p, t, 1e-3)
    assert r["tolerance_merges"] > 1000


def test_special_values(ctx):
    snan = struct.unpack("<f", struct.pack("<I", 0x7F800001))[0]
    qnan = struct.unpack("<f", struct.pack("<I", 0x7FC00001))[0]
    den = np.float32(1e-41)
    p = np.array([[-0.0, 1, 2], [0.0, 1, 2], [den, 3, 4], [den, 3, 4], [snan, 0, 0], [qnan, 0, 0],
                  [np.nan, 1, 1], [1.5, 2.5, 3.5], [1.5, 2.5, 3.5000002], [7, 8, 9]], np.float32)
    p[4, 0] = np.frombuffer(struct.pack("<I", 0x7F800001), np.float32)[0]
    t = np.array([[0, 2, 4], [1, 3, 5], [6, 7, 8], [7, 8, 9], [0, 7, 9]], np.uint32)
    for rel in (0.0, 1e-6, 0.05):
        _weld_both(ctx, p, t, rel)


def test_torus_soup_full(ctx):
    p, t = synth.unweld(*synth.torus(10.0, 2.0, 400, 200))
    r = _weld_both(ctx, p, t, 1e-6)
    assert r["output_vertices"] == 80000


def test_deterministic(ctx):
    p, t = clustered(3, n_centres=4000)
    outs = []
    for _ in range(3):
        m = ag.Mesh.from_arrays(ctx, p, t)
        m.weld(2e-3)
        outs.append(m.read())
        m.close()
    for o in outs[1:]:
        _same(o, outs[0])


def test_compact(ctx):
    p = np.arange(18, dtype=np.float32).reshape(6, 3)
    t = np.array([[0, 2, 4], [1, 1, 3], [4, 2, 5]], np.uint32)
    m = ag.Mesh.from_arrays(ctx, p, t)
    assert m.compact() == 1
    wp, wt, _ = O.compact(p, t)
    _same(m.read(), (wp, wt))


def test_index_out_of_range_is_rejected(ctx):
    m = ag.Mesh.from_arrays(ctx, np.zeros((3, 3), np.float32), np.array([[0, 1, 3]], np.uint32))
    with pytest.raises(ag.ArgentError):
        m.weld(1e-6)


def _load_both(ctx, data, **kw):
    m, r = ag.load_stl(ctx, data, **kw)
    got = m.read()
    m.close()
    wp, wt, wr = O.load_stl(data, **kw)
    _same(got, (wp, wt))
    for k in ("raw_triangles", "raw_vertices", "invalid_triangles", "trailing_bytes", "declared_triangles"):
        assert r[k] == wr[k], k
    if wr["weld"] is not None:
        for k in COUNTS:
            assert r["weld"][k] == wr["weld"][k], k
    return r


def test_stl_cube_padded_scaled(ctx):
    r = _load_both(ctx, stl_bytes(*synth.cuboid([2.0, 3.0, 4.0]), b"solid cube"))
    assert r["weld"]["output_vertices"] == 8
    data = bytearray(stl_bytes(*synth.cuboid([1.0, 1.0, 1.0])))
    data += bytes(7)
    data[:5] = b"BINRY"
    r = _load_both(ctx, bytes(data), unit_scale=25.4)
    assert r["trailing_bytes"] == 7


def test_stl_truncated_and_invalid(ctx):
    data = stl_bytes(*synth.torus(10.0, 2.0, 8, 6))
    _load_both(ctx, data[:84 + 50 * 41 + 17], weld=False)
    _load_both(ctx, data[:84 + 50 * 41 + 17])
    r = _load_both(ctx, stl_with_bad(3))
    assert r["invalid_triangles"] == 17
    for bad in (data[:83], data[:84 + 49]):
        with pytest.raises(ag.ArgentError):
            ag.load_stl(ctx, bad)


def test_stl_signet_full_size(ctx):
    # the 1.2 M-triangle signet as a binary STL soup (3.6 M vertices)
    pos, tris, _ = synth.signet()
    data = synth.to_binary_stl(pos, tris)
    r = _load_both(ctx, data)
    assert r["weld"]["exact_merges"] > 2 * len(pos)
    # the same bytes already in device memory
    buf = ag.DeviceBytes(ctx, data)
    m, r2 = ag.load_stl(ctx, len(data), device_ptr=buf.ptr)
    wp, wt, _ = O.load_stl(data)
    _same(m.read(), (wp, wt))
    m.close()
    buf.close()
