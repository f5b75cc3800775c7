"""Batched form: many meshes of uneven sizes, grid sizes and modes in one svo_batch_run,
each job's node and data records byte-compared with the plain path (svo_run on that job
alone) and, for the small ones, with the CPU oracle."""
import numpy as np
import pytest

import oracle
from ooc_svo_builder_amd import SvoContext, generators

from test_gpu_parity import _corner_mesh, adversarial_soup

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _mix(seed):
    """A batch exercising every job shape the batch code distinguishes: binary and
    payload, every colour mode, one and several partitions, an empty mesh, a mesh outside
    the grid, a 1^3 and a 2^3 grid, non-unit bboxes."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k, G in enumerate((64, 128, 256)):
        jobs.append(dict(records=generators.icosphere(subdivisions=2 + k,
                                                      radius=0.2 + 0.1 * k, seed=k + 1),
                         payload=False, gridsize=G))
    for mode in range(4):
        jobs.append(dict(records=generators.terrain(n=33 + 16 * mode, seed=10 + mode,
                                                    payload=True),
                         payload=True, gridsize=(64, 128, 64, 128)[mode], color_mode=mode))
    jobs.append(dict(records=generators.terrain(n=41, seed=20, payload=False), payload=False,
                     gridsize=256, memory_limit_mb=2))  # 8 partitions
    jobs.append(dict(records=generators.terrain(n=41, seed=21, payload=True), payload=True,
                     gridsize=256, memory_limit_mb=2))  # 8 partitions, payload
    jobs.append(dict(records=np.zeros((0, 9), np.float32), payload=False, gridsize=64))
    jobs.append(dict(records=np.full((5, 9), 7.5, np.float32), payload=False, gridsize=32))
    jobs.append(dict(records=adversarial_soup(rng, 300, 64), payload=False, gridsize=64))
    jobs.append(dict(records=generators.icosphere(subdivisions=1), payload=False, gridsize=1))
    jobs.append(dict(records=generators.icosphere(subdivisions=1), payload=False, gridsize=2))
    jobs.append(dict(records=generators.icosphere(subdivisions=2) * np.float32(2.0),
                     payload=False, gridsize=128, bbox=(0, 0, 0, 2, 2, 2)))
    jobs.append(dict(records=_corner_mesh(32), payload=False, gridsize=32))  # hazard
    return jobs


def _plain(ctx, j):
    info = ctx.run(j["records"], j["payload"], j["gridsize"], j.get("memory_limit_mb", 2048),
                   bbox=j.get("bbox", (0, 0, 0, 1, 1, 1)), color_mode=j.get("color_mode", 0))
    return info, ctx.nodes(info.n_nodes), ctx.data(info.n_data)


def _check_batch(bctx, pctx, jobs):
    res, tot = bctx.batch_run(jobs)
    assert len(res) == len(jobs)
    assert tot.n_nodes == sum(r.n_nodes for r in res)
    assert tot.n_data == sum(r.n_data for r in res)
    for i, (j, r) in enumerate(zip(jobs, res)):
        info, nodes, data = _plain(pctx, j)
        assert (r.n_partitions, r.nfilled, r.n_nodes, r.n_data, r.hazard) == (
            info.n_partitions, info.nfilled, info.n_nodes, info.n_data, info.hazard), i
        assert bctx.nodes_range(r.node_offset, r.n_nodes).tobytes() == nodes.tobytes(), i
        assert bctx.data_range(r.data_offset, r.n_data).tobytes() == data.tobytes(), i
    return res


def test_batch_matches_plain():
    _say("batch of %d jobs vs plain path" % len(_mix(1)))
    jobs = _mix(1)
    with SvoContext(0) as b, SvoContext(0) as p:
        _check_batch(b, p, jobs)
        # the same context again: a smaller batch, reordered (buffers reused, not cleared)
        sub = [jobs[i] for i in (9, 3, 0, 12, 7)]
        _say("second batch of %d jobs" % len(sub))
        _check_batch(b, p, sub)
        # grids of 16^3 and up only: the build takes the brick stream (with and without a
        # payload job in the pass), the hazard job included
        for idx in ((0, 1, 7, 10, 11, 15), (15, 3, 1)):
            deep = [jobs[i] for i in idx]
            _say("brick-stream batch %s" % (idx,))
            _check_batch(b, p, deep)
        # knobs picked by the autotuner, and extreme ones: the same bytes
        _say("autotune")
        knobs = b.autotune(sub)
        assert b.get_tuning() == knobs
        _check_batch(b, p, jobs)
        b.set_tuning(small_cols=4)
        _check_batch(b, p, sub)


def test_batch_matches_oracle():
    """Anchor: a few batch jobs straight against the CPU oracle's pipeline."""
    _say("batch vs oracle")
    jobs = [dict(records=generators.icosphere(subdivisions=3), payload=False, gridsize=128),
            dict(records=generators.terrain(n=33, seed=5, payload=True), payload=True,
                 gridsize=64, color_mode=0),
            dict(records=generators.terrain(n=33, seed=6, payload=False), payload=False,
                 gridsize=256, memory_limit_mb=2)]
    with SvoContext(0) as b:
        res, _ = b.batch_run(jobs)
        for j, r in zip(jobs, res):
            nodes_w, data_w, hz_w, np_w, nf_w = oracle.pipeline(
                j["records"], j["payload"], j["gridsize"], j.get("memory_limit_mb", 2048),
                color_mode=j.get("color_mode", 0))
            assert (r.n_partitions, r.nfilled, r.hazard) == (np_w, nf_w, hz_w)
            assert b.nodes_range(r.node_offset, r.n_nodes).tobytes() == nodes_w.tobytes()
            assert b.data_range(r.data_offset, r.n_data).tobytes() == data_w.tobytes()


def test_batch_device_meshes_and_plain_after():
    """Meshes already in HBM are used in place; a plain run on the same context after a
    batch starts from a clean state."""
    import torch
    _say("batch with device meshes")
    jobs = _mix(2)[:6]
    djobs = [dict(j, records=torch.from_numpy(np.ascontiguousarray(j["records"])).cuda())
             for j in jobs]
    with SvoContext(0) as b, SvoContext(0) as p:
        _check_batch(b, p, djobs)
        info, nodes, data = _plain(b, jobs[4])
        info2, nodes2, data2 = _plain(p, jobs[4])
        assert nodes.tobytes() == nodes2.tobytes() and data.tobytes() == data2.tobytes()


def test_batch_with_levels_jobs():
    """-levels jobs in a batch (run through the plain path, their records appended after
    the shared jobs'): every job byte-equal to svo_run alone; a batch of -levels jobs only."""
    _say("batch with -levels jobs")
    jobs = _mix(6)[:8]
    jobs[1] = dict(jobs[1], levels=True)
    jobs[5] = dict(jobs[5], levels=True)
    jobs.append(dict(records=generators.terrain(n=41, seed=23, payload=True), payload=True,
                     gridsize=256, memory_limit_mb=2, levels=True))  # 8 partitions
    with SvoContext(0) as b, SvoContext(0) as p:
        res, tot = b.batch_run(jobs)
        assert tot.n_nodes == sum(r.n_nodes for r in res)
        for i, (j, r) in enumerate(zip(jobs, res)):
            info = p.run(j["records"], j["payload"], j["gridsize"], j.get("memory_limit_mb", 2048),
                         bbox=j.get("bbox", (0, 0, 0, 1, 1, 1)),
                         color_mode=j.get("color_mode", 0), levels=j.get("levels", False))
            assert (r.n_nodes, r.n_data, r.nfilled) == (info.n_nodes, info.n_data, info.nfilled), i
            assert b.nodes_range(r.node_offset, r.n_nodes).tobytes() == \
                p.nodes(info.n_nodes).tobytes(), i
            assert b.data_range(r.data_offset, r.n_data).tobytes() == \
                p.data(info.n_data).tobytes(), i
        only = [jobs[1], jobs[-1]]
        res, _ = b.batch_run(only)
        for j, r in zip(only, res):
            nodes_w, data_w, _, _, _ = oracle.pipeline(
                j["records"], j["payload"], j["gridsize"], j.get("memory_limit_mb", 2048),
                color_mode=j.get("color_mode", 0), levels=True)
            assert b.nodes_range(r.node_offset, r.n_nodes).tobytes() == nodes_w.tobytes()
            assert b.data_range(r.data_offset, r.n_data).tobytes() == data_w.tobytes()


def _batch_outputs(ctx, jobs):
    res, tot = ctx.batch_run(jobs)
    out = [((r.n_partitions, r.nfilled, r.n_nodes, r.n_data, r.hazard),
            ctx.nodes_range(r.node_offset, r.n_nodes).tobytes(),
            ctx.data_range(r.data_offset, r.n_data).tobytes()) for r in res]
    return out, tot


def test_batch_graph_replays_match_eager():
    """svo_batch_run with graphs: the first call captures each kernel sequence between
    host synchronizations into a graph, later calls with the same launches replay the kept
    graphs. Every replay (with eager batches, plain runs and synchronizations between
    them) gives the eager batch's bytes; a different batch captures anew."""
    jobs = _mix(3)
    other = [jobs[i] for i in (12, 4, 0, 9)]
    with SvoContext(0) as e, SvoContext(0) as g, SvoContext(0) as p:
        e.set_graphs(False)
        _say("graph test: eager batch of %d jobs" % len(jobs))
        want, tot_e = _batch_outputs(e, jobs)
        assert tot_e.graph_launches == 0
        want_other, _ = _batch_outputs(e, other)
        for rep in range(4):
            _say("graph test: batch call %d" % rep)
            got, tot = _batch_outputs(g, jobs)
            assert tot.graph_launches > 0 and tot.graph_kernels > 0
            if rep == 0:
                assert tot.graph_replays == 0
            elif rep == 1:  # a buffer first allocated after a segment's capture (the
                # payload winners, after the marking) changes that segment's arguments
                assert tot.graph_replays > 0
            else:  # from then on every segment replays
                assert tot.graph_replays == tot.graph_launches, (tot.graph_replays,
                                                                 tot.graph_launches)
            for i, (a, b) in enumerate(zip(got, want)):
                assert a == b, (rep, i)
            # eager work between the replays: a plain run and an eager batch, synchronized
            _plain(p, jobs[rep])
            assert _batch_outputs(e, other)[0] == want_other
        _say("graph test: a different batch, then the first again")
        got, tot = _batch_outputs(g, other)
        assert got == want_other
        got, tot = _batch_outputs(g, jobs)
        assert got == want
