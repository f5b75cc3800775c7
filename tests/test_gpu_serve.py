"""Served form: jobs submitted as they arrive (a burst, and from several threads), each
result byte-compared with the plain path on that job alone; admission by HBM budget."""
import threading

import numpy as np
import pytest

from ooc_svo_builder_amd import SvoContext, SvoServer

from test_gpu_batch import _mix, _plain

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _expected(jobs):
    with SvoContext(0) as p:
        out = []
        for j in jobs:
            info, nodes, data = _plain(p, j)
            out.append((info, nodes.tobytes(), data.tobytes()))
        return out


def _check(srv, ticket, want):
    info, nodes, data = want
    r, lat = srv.wait(ticket)
    assert lat >= 0.0
    assert (r.n_partitions, r.nfilled, r.n_nodes, r.n_data, r.hazard) == (
        info.n_partitions, info.nfilled, info.n_nodes, info.n_data, info.hazard)
    assert srv.nodes(ticket, r.n_nodes).tobytes() == nodes
    assert srv.data(ticket, r.n_data).tobytes() == data
    srv.release(ticket)


def test_served_burst_matches_plain():
    jobs = _mix(3)
    want = _expected(jobs)
    _say("served burst of %d jobs" % len(jobs))
    with SvoServer(0, slots=2) as srv:
        tickets = [srv.submit(j) for j in jobs]
        for t, w in zip(tickets, want):
            _check(srv, t, w)
        st = srv.stats()
        assert st.completed == len(jobs) and st.failed == 0
        assert st.batches < len(jobs)  # the burst was batched


def test_served_threads_and_alone():
    """Two client threads submitting interleaved jobs, then every job again alone: the
    same bytes in the mix and alone."""
    jobs = _mix(4)
    want = _expected(jobs)
    _say("served from two threads")
    errors = []
    with SvoServer(0, slots=2) as srv:
        def client(idx):
            try:
                ts = [(srv.submit(jobs[i]), i) for i in idx]
                for t, i in ts:
                    _check(srv, t, want[i])
            except Exception as e:  # reported by the main thread
                errors.append(e)
        th = [threading.Thread(target=client, args=(list(range(k, len(jobs), 2)),))
              for k in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errors, errors
        _say("served jobs alone")
        for i, j in enumerate(jobs):
            _check(srv, srv.submit(j), want[i])


def test_served_admission_budget():
    """A 1 MB budget admits one job at a time (each alone exceeds it): batches of one,
    results unchanged; a held result keeps later jobs waiting until it is released."""
    jobs = _mix(5)[:5]
    want = _expected(jobs)
    _say("served admission budget")
    with SvoServer(0, budget_mb=1, slots=2) as srv:
        tickets = [srv.submit(j) for j in jobs]
        _check(srv, tickets[0], want[0])
        for t, w in zip(tickets[1:], want[1:]):
            _check(srv, t, w)
        st = srv.stats()
        assert st.max_batch_jobs_seen == 1 and st.completed == len(jobs)


def test_served_levels_and_lone_jobs_match_oracle():
    """-levels jobs and lone jobs (run through the plain path by the server): the served
    bytes equal the oracle's."""
    import oracle
    from ooc_svo_builder_amd import generators
    _say("served -levels and lone jobs vs oracle")
    jobs = [dict(records=generators.terrain(n=49, seed=40, payload=True), payload=True,
                 gridsize=128, levels=True, color_mode=2),
            dict(records=generators.icosphere(subdivisions=3), payload=False, gridsize=64,
                 levels=True),
            dict(records=generators.terrain(n=33, seed=41, payload=True), payload=True,
                 gridsize=64)]
    with SvoServer(0, slots=2) as srv:
        for j in jobs:  # one at a time: each finds the queue empty
            t = srv.submit(j)
            r, _ = srv.wait(t)
            nodes_w, data_w, hz_w, np_w, nf_w = oracle.pipeline(
                j["records"], j["payload"], j["gridsize"], 2048,
                color_mode=j.get("color_mode", 0), levels=j.get("levels", False))
            assert (r.n_nodes, r.n_data, r.nfilled) == (len(nodes_w), len(data_w), nf_w)
            assert srv.nodes(t, r.n_nodes).tobytes() == nodes_w.tobytes()
            assert srv.data(t, r.n_data).tobytes() == data_w.tobytes()
            srv.release(t)
        # the same jobs in one burst (levels jobs alone, the others batched)
        tickets = [srv.submit(j) for j in jobs + jobs]
        for k, t in enumerate(tickets):
            j = (jobs + jobs)[k]
            r, _ = srv.wait(t)
            nodes_w, data_w, _, _, _ = oracle.pipeline(
                j["records"], j["payload"], j["gridsize"], 2048,
                color_mode=j.get("color_mode", 0), levels=j.get("levels", False))
            assert srv.nodes(t, r.n_nodes).tobytes() == nodes_w.tobytes()
            assert srv.data(t, r.n_data).tobytes() == data_w.tobytes()
            srv.release(t)
