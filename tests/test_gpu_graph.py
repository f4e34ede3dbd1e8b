"""Render passes replayed from a captured HIP graph (tuning launch_graph)
give the films of the same passes launched kernel by kernel, bit for bit:
progressive passes on one film (the graph replays; seeds and sample indices
change through device memory), with eager passes on other films in between,
shape changes (another spp count, another film size, another depth) that
capture again, and a batch of views; enqueued passes stay eager."""
import numpy as np
import pytest

import argent_amd as ag
from argent_amd import synth, workloads

pytestmark = pytest.mark.gpu


def _graph(ctx, on):
    ctx.set_tuning(launch_graph=1 if on else 0)


@pytest.fixture(scope="module")
def torus(ctx):
    pos, tris, nrm = synth.torus_ring(seed=1, nu=256, nv=64)
    piece = workloads.Piece(pos, tris, nrm)
    sc = ag.Scene(ctx, piece.scene("gold14k", 0.07, 0.8, "noir"))
    cam, turn = piece.shot("hero", 100.0, 16.0, 1.0)
    sc.set_turntable(turn)
    yield piece, sc, cam
    sc.close()
    _graph(ctx, False)


def test_progressive_passes_replay_bit_identical(ctx, torus):
    _, sc, cam = torus
    ps = workloads.path_settings(8)
    eager, graph, other = ag.Film(ctx, 96, 64), ag.Film(ctx, 96, 64), ag.Film(ctx, 80, 80)
    c0, r0 = ctx.graph_stats()
    for k in range(5):
        _graph(ctx, False)
        ag.render_pass(ctx, sc, cam, eager, ps, 3, 100 + k)
        ag.render_pass(ctx, sc, cam, other, ps, 2, 7 + k)  # eager work between replays
        _graph(ctx, True)
        ag.render_pass(ctx, sc, cam, graph, ps, 3, 100 + k)
        assert np.array_equal(eager.sums(), graph.sums()), k
        assert ctx.stats()["samples"] == 96 * 64 * 3
    c1, r1 = ctx.graph_stats()
    assert c1 - c0 == 1 and r1 - r0 == 4
    for f in (eager, graph, other):
        f.close()


@pytest.mark.parametrize("shape", ["spp", "film", "depth"])
def test_shape_changes_capture_again(ctx, torus, shape):
    _, sc, cam = torus
    films = {}
    for on in (False, True):
        _graph(ctx, on)
        f1 = ag.Film(ctx, 64, 64)
        ag.render_pass(ctx, sc, cam, f1, workloads.path_settings(8), 2, 5)
        if shape == "spp":
            ag.render_pass(ctx, sc, cam, f1, workloads.path_settings(8), 5, 6)
            f2 = f1
        elif shape == "film":
            f2 = ag.Film(ctx, 48, 72)
            ag.render_pass(ctx, sc, cam, f2, workloads.path_settings(8), 2, 6)
        else:
            ag.render_pass(ctx, sc, cam, f1, workloads.path_settings(3), 2, 6)
            f2 = f1
        films[on] = (f1.sums(), f2.sums())
    assert np.array_equal(films[False][0], films[True][0])
    assert np.array_equal(films[False][1], films[True][1])


def test_views_batch_with_graphs(ctx, torus):
    _, sc, cam = torus
    ps = workloads.path_settings(8)
    out = {}
    for on in (False, True):
        _graph(ctx, on)
        views = [dict(camera=cam, turntable_deg=10.0 * i, count=1 + i % 3, seed=50 + i, film=ag.Film(ctx, 40, 40))
                 for i in range(12)]
        for _ in range(3):
            ag.render_views(ctx, sc, views, ps)
        out[on] = [v["film"].sums() for v in views]
        for v in views:
            v["film"].close()
    for a, b in zip(out[False], out[True]):
        assert np.array_equal(a, b)


def test_async_passes_stay_eager(ctx, torus):
    """An enqueued pass (ag_render_pass_async) launches kernel by kernel
    whatever the tuning; its film equals the synchronous pass's."""
    _, sc, cam = torus
    ps = workloads.path_settings(8)
    _graph(ctx, True)
    c0, r0 = ctx.graph_stats()
    fa, fb = ag.Film(ctx, 64, 64), ag.Film(ctx, 64, 64)
    ag.render_pass_async(ctx, sc, cam, fa, ps, 2, 9).wait()
    assert ctx.graph_stats() == (c0, r0)
    ag.render_pass(ctx, sc, cam, fb, ps, 2, 9)
    assert np.array_equal(fa.sums(), fb.sums())
    fa.close()
    fb.close()
