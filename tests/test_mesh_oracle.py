"""The oracle's mesh ingest (binary STL, drop_invalid, weld, compact) pinned
by the reference's own weld and io tests (weld.rs:332-411,
io/tests.rs:13-45), a brute-force weld on clustered points, and the STL
record layout."""
import struct

import numpy as np
import pytest

import oracle
from argent_amd import synth
from meshcases import brute_weld, clustered, signed_volume, stl_bytes, stl_with_bad


def test_cube_soup_welds_to_8():
    # weld.rs:338-347
    p, t = synth.unweld(*synth.cuboid([2.0, 2.0, 2.0]))
    assert len(p) == 36
    q, s, r = oracle.weld(p, t, 1e-6)
    assert len(q) == 8 and len(s) == 12
    assert r["exact_merges"] == 28
    assert abs(signed_volume(q, s) - 8.0) < 1e-5


def test_near_duplicates_merge_within_tolerance_only():
    # weld.rs:349-368: jitter up to 1e-7 mm, tolerance 1e-6 x diag ~ 3.5e-6
    p, t = synth.unweld(*synth.cuboid([2.0, 2.0, 2.0]))
    rng = np.random.default_rng(5)
    p[:, 0] += ((rng.random(len(p)) - 0.5) * 2e-7).astype(np.float32)
    q, _, r = oracle.weld(p, t, 1e-6)
    assert len(q) == 8, r
    assert r["tolerance_merges"] > 0
    p2, t2 = synth.unweld(*synth.cuboid([2.0, 2.0, 2.0]))
    p2[0, 0] += np.float32(1e-4)
    q2, _, _ = oracle.weld(p2, t2, 0.0)
    assert len(q2) == 9


def test_collapsed_triangles_dropped():
    # weld.rs:370-390
    p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1e-9, 0], [5, 5, 5]], np.float32)
    t = np.array([[0, 1, 2], [0, 1, 3], [1, 3, 2]], np.uint32)
    q, s, r = oracle.weld(p, t, 1e-6)
    assert r["dropped_triangles"] == 2
    assert s.tolist() == [[0, 1, 2]]
    assert len(q) == 3


def test_torus_welds_and_is_deterministic():
    # weld.rs:392-404
    p, t = synth.unweld(*synth.torus(10.0, 2.0, 400, 200))
    q1, s1, _ = oracle.weld(p, t, 1e-6)
    q2, s2, _ = oracle.weld(p, t, 1e-6)
    assert len(q1) == 400 * 200
    assert np.array_equal(s1, s2) and np.array_equal(q1, q2)
    # first-occurrence order: triangle 0's corners become vertices 0, 1, 2
    assert s1[0].tolist() == [0, 1, 2]


@pytest.mark.parametrize("seed", range(6))
def test_weld_matches_brute_force(seed):
    p, t = clustered(seed)
    for rel in (1e-3, 0.0, 3e-3):
        q, s, r = oracle.weld(p, t, rel)
        bq, bs, br = brute_weld(p, t, rel)
        assert np.array_equal(q.view(np.uint32), bq.view(np.uint32)), (seed, rel)
        assert np.array_equal(s, bs), (seed, rel)
        for k in br:
            assert r[k] == br[k], (seed, rel, k)
        assert r["input_vertices"] == len(p) and r["output_vertices"] == len(q)
    # the cases exercise both kinds of merge
    assert r["exact_merges"] > 0 and r["tolerance_merges"] > 0


def test_signed_zeros_collide_first_occurrence_kept():
    p = np.array([[-0.0, 1, 2], [0.0, 1, 2], [3, 4, 5], [3, 4, 6]], np.float32)
    t = np.array([[0, 2, 3], [1, 3, 2]], np.uint32)
    q, s, r = oracle.weld(p, t, 0.0)
    assert len(q) == 3 and r["exact_merges"] == 1
    assert np.signbit(q[0, 0])  # the first occurrence was -0.0
    assert s.tolist() == [[0, 1, 2], [0, 2, 1]]


def test_compact_drops_collapsed_and_unreferenced():
    p = np.arange(18, dtype=np.float32).reshape(6, 3)
    t = np.array([[0, 2, 4], [1, 1, 3], [4, 2, 5]], np.uint32)
    q, s, dropped = oracle.compact(p, t)
    assert dropped == 1
    assert np.array_equal(q, p[[0, 2, 4, 5]])
    assert s.tolist() == [[0, 1, 2], [2, 1, 3]]


def test_stl_record_layout_and_cube_load():
    # io/tests.rs:13-31 (weld part): a "solid" header, 12 triangles, 8 welded corners
    pos, tris = synth.cuboid([2.0, 3.0, 4.0])
    data = stl_bytes(pos, tris, b"solid cube")
    assert data.startswith(b"solid") and len(data) == 84 + 50 * 12
    # the record layout: normal, three vertices, attribute word
    v = struct.unpack_from("<9f", data, 84 + 50 * 5 + 12)
    assert np.array_equal(np.array(v, np.float32).reshape(3, 3), pos[tris[5]])
    q, s, r = oracle.load_stl(data)
    assert r["raw_triangles"] == 12 and r["raw_vertices"] == 36
    assert r["weld"]["output_vertices"] == 8 and len(s) == 12
    assert abs(signed_volume(q, s) - 24.0) < 1e-4


def test_stl_padded_and_scaled():
    # io/tests.rs:32-45
    data = bytearray(stl_bytes(*synth.cuboid([1.0, 1.0, 1.0])))
    data += bytes(7)
    data[:5] = b"BINRY"
    q, s, r = oracle.load_stl(bytes(data), unit_scale=25.4)
    assert r["trailing_bytes"] == 7
    assert abs(signed_volume(q, s) - 25.4 ** 3) / 25.4 ** 3 < 1e-5


def test_stl_truncated_uses_available_records():
    data = stl_bytes(*synth.torus(10.0, 2.0, 8, 6))
    cut = data[:84 + 50 * 40 + 17]
    q, s, r = oracle.load_stl(cut, weld=False)
    assert r["declared_triangles"] == 96 and r["raw_triangles"] == 40
    assert len(s) == 40 and len(q) == 120
    with pytest.raises(ValueError):
        oracle.load_stl(data[:83])
    with pytest.raises(ValueError):
        oracle.load_stl(data[:84 + 49])  # a header and no whole record


def test_stl_non_finite_records_dropped():
    data = stl_with_bad(3)
    q, s, r = oracle.load_stl(data)
    assert r["invalid_triangles"] == 17
    assert np.isfinite(q).all()
    # the same as welding the parsed soup with those triangles removed
    p, t = synth.parse_binary_stl(data)
    ok = np.isfinite(p[t]).all(axis=(1, 2))
    q2, s2, _ = oracle.weld(p, t[ok], 1e-6)
    assert np.array_equal(q, q2) and np.array_equal(s, s2)


def test_load_without_weld_is_compact_soup():
    data = stl_bytes(*synth.cuboid([1.0, 2.0, 3.0]))
    q, s, r = oracle.load_stl(data, weld=False)
    assert r["weld"] is None and len(q) == 36 and s.tolist() == np.arange(36).reshape(12, 3).tolist()


def test_bounds_skip_every_nan():
    # weld.rs:126-148 folds with f32::min/max, which return the other operand
    # for any NaN, signalling ones included (glibc's fmin does not)
    p = np.array([[0, 0, 0], [1, 2, 2], [3, 0, 1], [2, 2, 2]], np.float32)
    t = np.array([[0, 1, 2], [1, 2, 3]], np.uint32)
    _, _, r0 = oracle.weld(p, t, 1e-3)
    for bits in (0x7F800001, 0x7FC00000, 0xFFC00001):
        q = np.concatenate([p, p[:1]])
        q[4, 0] = np.frombuffer(np.uint32(bits).tobytes(), np.float32)[0]
        _, _, r = oracle.weld(q, np.concatenate([t, [[4, 1, 2]]]).astype(np.uint32), 1e-3)
        assert np.float32(r["tolerance_mm"]) == np.float32(r0["tolerance_mm"]), hex(bits)
    assert np.float32(r0["tolerance_mm"]) == np.float32(np.float64(np.float32(1e-3)) * np.float64(np.sqrt(np.float32(17))))
