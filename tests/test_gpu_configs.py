"""The workload configurations at their full sizes on the GPU.

* 1025^2 fBm payload terrain (2,097,152 triangles) at -s 1024 -l 2048: nodes and data
  byte-equal to the oracle.
* The same terrain's geometry at -s 2048 -l 8192 (the 2048^3 binary headline): nodes and
  data byte-equal to the oracle.
* The 512-mesh job mix (job_mix(512, seed=4): icosphere / terrain 50/50, 1k-2M
  triangles, gridsize 64-1024, 30% payload) in one svo_batch_run: every job byte-equal to
  svo_run on that job alone, and one job per (kind, payload, gridsize) stratum — the
  smallest — byte-equal to the oracle.
* Served: Poisson arrivals drawn from the same mix plus 2049^2 payload terrains
  (8,388,608 triangles) at gridsize 2048, submitted as they arrive: every result
  byte-equal to the plain path on that job alone, the large job byte-equal to the oracle.
"""
import os
import time

import numpy as np
import pytest

import oracle
from ooc_svo_builder_amd import SvoContext, SvoServer, generators

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _vs_oracle(ctx, tris, payload, G, L, **kw):
    info = ctx.run(tris, payload, G, L, **kw)
    nodes_w, data_w, hz_w, np_w, nf_w = oracle.pipeline(tris, payload, G, L, **kw)
    assert (info.n_partitions, info.nfilled, info.hazard) == (np_w, nf_w, hz_w)
    assert (info.n_nodes, info.n_data) == (len(nodes_w), len(data_w))
    assert ctx.nodes(info.n_nodes).tobytes() == nodes_w.tobytes()
    assert ctx.data(info.n_data).tobytes() == data_w.tobytes()
    return info


def test_payload_terrain_1024_full_size(gpu_ctx):
    _say("payload terrain 1025^2 at 1024 vs oracle")
    tris = generators.terrain(n=1025, seed=2, payload=True)
    info = _vs_oracle(gpu_ctx, tris, True, 1024, 2048)
    assert info.n_partitions == 1 and info.nfilled == 1_980_808


def test_binary_terrain_2048_headline_full_size(gpu_ctx):
    _say("binary terrain 1025^2 at 2048 vs oracle")
    tris = generators.terrain(n=1025, seed=2, payload=False)
    info = _vs_oracle(gpu_ctx, tris, False, 2048, 8192)
    assert info.n_partitions == 1 and info.nfilled == 7_923_772


@pytest.fixture(scope="module")
def mix():
    import torch
    workers = max(1, min(16, os.cpu_count() or 1))
    t = time.time()
    jobs = generators.job_mix(512, seed=4, workers=workers)
    spec = generators.mix_spec(512, seed=4)
    _say("job mix generated in %.1f s" % (time.time() - t))
    dev = [dict(j, records=torch.from_numpy(j["records"]).cuda()) for j in jobs]
    torch.cuda.synchronize()
    return jobs, dev, spec


def _device_copy(ctx_copy, n, rec):
    import torch
    out = torch.empty(n * rec, dtype=torch.uint8, device="cuda:0")
    ctx_copy(out.data_ptr(), n)
    return out


def test_job_mix_512_batched(mix):
    import torch
    jobs, dev, spec = mix
    _say("512-mesh mix: batch")
    with SvoContext(0) as b, SvoContext(0) as p:
        res, tot = b.batch_run(dev)
        assert len(res) == 512
        same = 0
        for i, (j, r) in enumerate(zip(dev, res)):
            info = p.run(j["records"], j["payload"], j["gridsize"])
            assert (r.n_partitions, r.nfilled, r.n_nodes, r.n_data, r.hazard) == (
                info.n_partitions, info.nfilled, info.n_nodes, info.n_data, info.hazard), i
            pn = _device_copy(lambda d, n: p.copy_nodes_to(d, 0, n), info.n_nodes, 24)
            pd = _device_copy(lambda d, n: p.copy_data_to(d, 0, n), info.n_data, 32)
            bn = _device_copy(lambda d, n: b.copy_nodes_to(d, r.node_offset, n), r.n_nodes, 24)
            bd = _device_copy(lambda d, n: b.copy_data_to(d, r.data_offset, n), r.n_data, 32)
            assert torch.equal(pn, bn) and torch.equal(pd, bd), i
            same += 1
        assert same == 512
        _say("512-mesh mix: stratified sample vs oracle")
        strata = {}
        for i, (kind, size, payload, G, _) in enumerate(spec):
            key = (kind, payload, G)
            n = jobs[i]["records"].shape[0]
            if key not in strata or n < jobs[strata[key]]["records"].shape[0]:
                strata[key] = i
        assert len(strata) == 20
        for key, i in sorted(strata.items()):
            j, r = jobs[i], res[i]
            nodes_w, data_w, hz_w, np_w, nf_w = oracle.pipeline(j["records"], j["payload"],
                                                                j["gridsize"], 2048)
            assert (r.n_partitions, r.nfilled, r.hazard) == (np_w, nf_w, hz_w), key
            assert b.nodes_range(r.node_offset, r.n_nodes).tobytes() == nodes_w.tobytes(), key
            assert b.data_range(r.data_offset, r.n_data).tobytes() == data_w.tobytes(), key


def test_served_poisson_mix_with_large_jobs(mix):
    import torch
    jobs, dev, spec = mix
    large_host = generators.terrain(n=2049, seed=50, payload=True)
    large = dict(records=torch.from_numpy(large_host).cuda(), payload=True, gridsize=2048)
    rng = np.random.default_rng(5)
    lam, n = 500.0, 200
    arrive = np.cumsum(rng.exponential(1.0 / lam, size=n))
    pick = [(-1 if rng.uniform() < 0.02 else int(rng.integers(0, 512))) for _ in range(n)]
    pick[n // 2] = -1  # at least one large job, in the middle of the stream
    # the plain path's bytes for every job the stream uses, kept on the device
    want = {}
    with SvoContext(0) as p:
        for k in sorted(set(pick)):
            j = large if k < 0 else dev[k]
            info = p.run(j["records"], j["payload"], j["gridsize"])
            want[k] = (info.n_nodes, info.n_data,
                       _device_copy(lambda d, m: p.copy_nodes_to(d, 0, m), info.n_nodes, 24),
                       _device_copy(lambda d, m: p.copy_data_to(d, 0, m), info.n_data, 32))
    torch.cuda.synchronize()
    _say("served: %d Poisson arrivals at %.0f/s, %d large" % (n, lam, sum(k < 0 for k in pick)))
    tickets = []
    with SvoServer(0, budget_mb=200 * 1024, slots=2) as srv:
        t0 = time.perf_counter()
        for i in range(n):
            while time.perf_counter() - t0 < arrive[i]:
                time.sleep(0.0002)
            tickets.append(srv.submit(large if pick[i] < 0 else dev[pick[i]]))
        large_bytes = None
        for t, k in zip(tickets, pick):
            r, lat = srv.wait(t)
            nn, nd, wn, wd = want[k]
            assert (r.n_nodes, r.n_data) == (nn, nd), k
            gn = _device_copy(lambda d, m: srv.copy_nodes_to(t, d, m), nn, 24)
            gd = _device_copy(lambda d, m: srv.copy_data_to(t, d, m), nd, 32)
            assert torch.equal(gn, wn) and torch.equal(gd, wd), k
            if k < 0 and large_bytes is None:
                large_bytes = (gn.cpu().numpy().tobytes(), gd.cpu().numpy().tobytes())
            srv.release(t)
        st = srv.stats()
        assert st.completed == n and st.failed == 0
        assert st.batches < n  # arrivals while others ran were batched
    _say("served: large job vs oracle")
    nodes_w, data_w, _, np_w, _ = oracle.pipeline(large_host, True, 2048, 2048)
    assert np_w == 8
    assert large_bytes[0] == nodes_w.tobytes()
    assert large_bytes[1] == data_w.tobytes()
