#!/usr/bin/env python3
"""Throughput of the partition -> voxelize -> SVO-build pipeline on one MI355X.

A step is one full pass of the pipeline (partition, voxelize every partition, build the
octree) over a synthetic mesh already resident in HBM, with the octree left in HBM.
The default workload is the 2048^3 binary headline: a 1025 x 1025-vertex fBm terrain
(2,097,152 triangles, geometry only) voxelized at gridsize 2048 with -l 8192 (one
partition). Prints one JSON line.

The CPU baseline is the single-threaded CPU restatement of the reference algorithm
(oracle/), timed on the same host on the same workload.

Two more sections report the batched and served forms on a 512-mesh mix (icosphere /
terrain 50/50, 1k-2M triangles log-uniform, gridsize 64-1024, 30% payload; meshes
resident in HBM):
  * batched: the whole mix in one svo_batch_run against a plain per-mesh loop of
    svo_run, every job's node and data bytes compared between the two;
  * served: open-loop Poisson arrivals (50-6000 jobs/s, plus 2% large jobs: a 2049^2
    terrain, 8,388,608 triangles, payload, gridsize 2048) into svo_server with a 200 GB
    admission budget; p50/p99/p999 latency and jobs/s, next to the plain loop's latency
    under the same arrivals (a FIFO queue served at the measured per-job plain times).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (generator, payload, gridsize, memory limit MB, description)
    "terrain2048": ("terrain", False, 2048, 8192,
                    "fBm terrain 1025x1025 vertices (2,097,152 tris), binary, -s 2048 -l 8192"),
    "icosphere2048": ("icosphere", False, 2048, 8192,
                      "icosphere subdiv 5 (20,480 tris), binary, -s 2048 -l 8192"),
    "icosphere256": ("icosphere", False, 256, 2048,
                     "icosphere subdiv 5 (20,480 tris), binary, -s 256 -l 2048"),
    "terrain1024p": ("terrain", True, 1024, 2048,
                     "fBm terrain 1025x1025 vertices (2,097,152 tris), payload, -s 1024 -l 2048"),
}


TRAFFIC_FILES = {"terrain2048": "hbm_traffic_terrain2048_binary.md"}


def measured_traffic(workload, timer):
    """(HBM bytes per launch, SIMD VALU issue share) of the kernel behind 'timer' from the
    committed rocprofv3 counter summary of this workload (None where there is none)."""
    name = TRAFFIC_FILES.get(workload)
    path = os.path.join(ROOT, "profiles", name) if name else None
    if not path or not os.path.exists(path):
        return None, None
    traffic = share = None
    kernel = None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        try:
            if len(cells) == 8 and cells[1] == timer:
                kernel = cells[0]
                traffic = float(cells[7]) * 1e6
            elif len(cells) == 9 and kernel and cells[0] == kernel and cells[8] != "nan":
                share = float(cells[8])
        except ValueError:
            pass
    return traffic, share


def make_mesh(kind, payload):
    from ooc_svo_builder_amd import generators
    if kind == "terrain":
        return generators.terrain(n=1025, seed=2, payload=payload)
    return generators.icosphere()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="terrain2048", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched section")
    ap.add_argument("--no-served", action="store_true", help="skip the served section")
    ap.add_argument("--only", choices=["headline", "batched", "served"],
                    help="run one section only (profiling)")
    ap.add_argument("--knobs", default=None,
                    help="voxelizer knobs small_cols,direct_cols,tile_blocks,large_blocks "
                         "for the headline (default: the library's)")
    ap.add_argument("--mix-jobs", type=int, default=512, help="meshes in the job mix")
    ap.add_argument("--served-seconds", type=float, default=2.0,
                    help="arrival window per rate of the served sweep")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (the reference kernel "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy()
    if args.only in ("batched", "served"):
        args.no_cpu = True
        args.no_batched = args.only != "batched"
        args.no_served = args.only != "served"
        args.steps, args.warmup = 1, 0
    elif args.only == "headline":
        args.no_batched = args.no_served = True
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(the multi-GPU sharding is a design, not a measured path)" % args.gpus,
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from ooc_svo_builder_amd import SvoContext

    kind, payload, G, L, desc = WORKLOADS[args.workload]
    tris = make_mesh(kind, payload)
    N = tris.shape[0]
    R = 84 if payload else 36
    dev = torch.device("cuda:0")
    d_tris = torch.from_numpy(tris).to(dev)
    torch.cuda.synchronize()

    ctx = SvoContext(0)
    if args.knobs:
        ctx.set_knobs(*[int(x) for x in args.knobs.split(",")])
    ctx.load_mesh(d_tris, payload)
    stream = torch.cuda.ExternalStream(ctx.stream)
    for _ in range(args.warmup):
        info = ctx.rerun(payload, G, L)
    torch.cuda.synchronize()

    # timed steps: nothing but the pipeline between the two events
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    st = {"partition": 0.0, "voxelize": 0.0, "build": 0.0}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        info = ctx.rerun(payload, G, L)
        st["partition"] += info.ms_partition
        st["voxelize"] += info.ms_voxelize
        st["build"] += info.ms_build
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_total = e0.elapsed_time(e1)
    ms_step = ms_total / args.steps
    V = int(info.nfilled)
    n_nodes = int(info.n_nodes)
    n_data = int(info.n_data)

    # the same steps with the mesh handed over in host memory (pageable numpy array):
    # the upload over PCIe is inside each step; never the headline value
    h_steps = max(1, min(args.steps, 5))
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(h_steps):
        ctx.run(tris, payload, G, L)
    torch.cuda.synchronize()
    pcie_s = (time.perf_counter() - t) / h_steps
    ctx.load_mesh(d_tris, payload)

    # per-kernel device time: the same steps again with HIP events around every launch
    kt_steps = max(1, min(args.steps, 5))
    ctx.set_kernel_timing(True)
    ctx.reset_kernel_timing()
    for _ in range(kt_steps):
        ctx.rerun(payload, G, L)
    torch.cuda.synchronize()

    # dominant kernel: the instrumented kernel with the most device time
    names = ["part_hitbox", "vox_small_touch", "vox_large_touch", "vox_small_mark",
             "vox_large_mark", "brick_count", "brick_emit",
             "vox_small_winner", "vox_large_winner", "payload_records", "brick_clear",
             "brick_level", "emit_children", "emit_bricks", "emit_leaves"]
    ktimes = {}
    for nm in names:
        n, ms = ctx.kernel_timing(nm)
        if n:
            ktimes[nm] = (n, ms)
    ctx.set_kernel_timing(False)
    dom = max(ktimes, key=lambda k: ktimes[k][1])
    launches, dom_ms = ktimes[dom]
    per_step_launches = launches / kt_steps
    # algorithmic bytes (per stage, per unit): voxelize N*R + 8V (payload N*R + 32V);
    # build 8V + 24 n_nodes + 32 n_data (payload + 24V)
    vox_bytes = N * R + (32 if payload else 8) * V
    build_bytes = 8 * V + 24 * n_nodes + 32 * n_data + (24 * V if payload else 0)
    if dom.startswith("vox_") or dom in ("brick_count", "brick_emit", "brick_clear",
                                         "payload_records", "part_hitbox"):
        algo_bytes_per_launch = vox_bytes / per_step_launches
    else:
        algo_bytes_per_launch = build_bytes / per_step_launches
    avg_launch_ms = dom_ms / launches
    achieved = algo_bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    traffic, valu_share = measured_traffic(args.workload, dom)

    value = N * args.steps / (ms_total * 1e-3)
    stage_ms = {k: v / args.steps for k, v in st.items()}
    out = {
        "metric": "triangles/s voxelized->SVO @2048^3 (binary terrain), one MI355X",
        "value": value,
        "unit": "triangles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": desc, "triangles": N, "gridsize": G, "memory_limit_mb": L,
                   "partitions": int(info.n_partitions), "voxels": V, "nodes": n_nodes,
                   "payload": payload},
        "host_input_triangles_per_s": N / pcie_s,
        "voxels_per_s": V / (ms_step * 1e-3),
        "nodes_per_s": n_nodes / (ms_step * 1e-3),
        "stage_ms": stage_ms,
        "stage_hbm_frac": {
            "voxelize": vox_bytes / (stage_ms["voxelize"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "build": build_bytes / (stage_ms["build"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "kernels_ms_per_step": {k: v[1] / kt_steps for k, v in ktimes.items()},
        "knobs": dict(zip(("small_cols", "direct_cols", "tile_blocks", "large_blocks"),
                          ctx.get_tuning())),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "simd_valu_share": valu_share,
                     "avg_launch_ms": avg_launch_ms,
                     "algo_bytes_per_launch": algo_bytes_per_launch},
        "wall_s_timed": wall,
    }

    if not args.no_cpu:
        import oracle
        t = time.perf_counter()
        nodes_w, data_w, hz_w, np_w, nf_w = oracle.pipeline(tris, payload, G, L)
        cpu_s = time.perf_counter() - t
        # same workload, so the CPU result doubles as a parity check of this run
        same = (nf_w == V and len(nodes_w) == n_nodes and
                np.array_equal(ctx.nodes(n_nodes), nodes_w))
        out["cpu_baseline"] = {"value": N / cpu_s, "unit": "triangles/s", "cores": 1,
                               "kind": "port",
                               "sample": "the full workload once (%d triangles), %.1f s" % (N, cpu_s),
                               "gpu_speedup": value / (N / cpu_s),
                               "bit_exact_vs_gpu": bool(same)}
    ctx.close()

    mix = None
    if not (args.no_batched and args.no_served):
        try:
            mix = MixBench(args.mix_jobs)
        except Exception as e:  # the headline line is printed regardless
            out["batched"] = out["served"] = {"error": repr(e)}
    if mix is not None and not args.no_batched:
        try:
            out["batched"] = mix.batched()
        except Exception as e:
            out["batched"] = {"error": repr(e)}
    if mix is not None and not args.no_served:
        mix.drop_outputs()  # the served sweep needs the HBM
        try:
            out["served"] = mix.served(args.served_seconds)
        except Exception as e:
            out["served"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)
    return 0


def calibrate_copy():
    """A device-to-device copy of exactly 1 GiB, 5 times: its counters read 1 GiB fetched
    and 1 GiB written per launch on an ideal memory system."""
    import torch
    n = 1 << 30
    a = torch.ones(n // 4, dtype=torch.int32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_per_launch": n, "launches": 5}), flush=True)
    return 0


def _log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def _pct(a, q):
    import numpy as np
    return float(np.percentile(np.asarray(a, dtype=np.float64), q)) if len(a) else None


class MixBench:
    """The 512-mesh mix on the device, its plain per-job results kept for comparison."""

    def __init__(self, n_jobs):
        import numpy as np
        import torch

        from ooc_svo_builder_amd import SvoContext, generators
        self.np, self.torch = np, torch
        workers = max(1, min(16, os.cpu_count() or 1))
        t = time.perf_counter()
        jobs = generators.job_mix(n_jobs, seed=4, workers=workers)
        _log("mix of %d meshes generated in %.1f s" % (n_jobs, time.perf_counter() - t))
        dev = torch.device("cuda:0")
        self.jobs = [dict(j, records=torch.from_numpy(j["records"]).to(dev)) for j in jobs]
        self.n_tris = int(sum(j["records"].shape[0] for j in jobs))
        del jobs
        torch.cuda.synchronize()
        self.ctx = SvoContext(0)
        # plain per-job loop: warm-up, timed pass, then a pass that keeps every output
        for j in self.jobs[:8]:
            self._plain(j)
        torch.cuda.synchronize()
        self.plain_ms = []
        t0 = time.perf_counter()
        for j in self.jobs:
            t = time.perf_counter()
            self._plain(j)
            self.plain_ms.append((time.perf_counter() - t) * 1e3)
        self.plain_s = time.perf_counter() - t0
        _log("plain loop: %d jobs in %.3f s" % (len(self.jobs), self.plain_s))
        self.want = []
        for j in self.jobs:
            info = self._plain(j)
            nodes = torch.empty(info.n_nodes * 24, dtype=torch.uint8, device=dev)
            data = torch.empty(info.n_data * 32, dtype=torch.uint8, device=dev)
            self.ctx.copy_nodes_to(nodes.data_ptr(), 0, info.n_nodes)
            self.ctx.copy_data_to(data.data_ptr(), 0, info.n_data)
            self.want.append((info.n_nodes, info.n_data, info.nfilled, nodes, data))

    def _plain(self, j):
        return self.ctx.run(j["records"], j["payload"], j["gridsize"])

    def drop_outputs(self):
        self.want = []
        self.torch.cuda.empty_cache()

    def batched(self):
        torch = self.torch
        from ooc_svo_builder_amd import SvoContext
        b = SvoContext(0)
        res, info = b.batch_run(self.jobs)  # warm-up (allocations)
        t = time.perf_counter()
        knobs = b.autotune(self.jobs)
        _log("autotuned small_cols/tile_blocks/large_blocks = %s in %.1f s" % (
            knobs, time.perf_counter() - t))
        reps = 3
        times = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            res, info = b.batch_run(self.jobs)
            times.append(time.perf_counter() - t)
        best = min(times)
        same = 0
        dev = torch.device("cuda:0")
        for r, w in zip(res, self.want):
            nn, nd, nf, nodes, data = w
            if (r.n_nodes, r.n_data, r.nfilled) != (nn, nd, nf):
                continue
            got_n = torch.empty_like(nodes)
            got_d = torch.empty_like(data)
            b.copy_nodes_to(got_n.data_ptr(), r.node_offset, nn)
            b.copy_data_to(got_d.data_ptr(), r.data_offset, nd)
            same += int(torch.equal(got_n, nodes) and torch.equal(got_d, data))
        b.close()
        n = len(self.jobs)
        _log("batched: %d jobs in %.3f s, %d/%d identical" % (n, best, same, n))
        return {"jobs": n, "triangles": self.n_tris, "seconds": best,
                "jobs_per_s": n / best, "triangles_per_s": self.n_tris / best,
                "stage_ms": {"partition": info.ms_partition, "voxelize": info.ms_voxelize,
                             "build": info.ms_build},
                "plain_loop_seconds": self.plain_s, "plain_jobs_per_s": n / self.plain_s,
                "speedup_vs_plain": self.plain_s / best,
                "identical_jobs": same, "identical": same == n,
                "autotuned": {"small_cols": knobs[0], "direct_cols": knobs[1],
                              "tile_blocks": knobs[2], "large_blocks": knobs[3]},
                "voxels": int(info.nfilled), "nodes": int(info.n_nodes)}

    def served(self, seconds):
        np, torch = self.np, self.torch
        from ooc_svo_builder_amd import SvoServer, generators
        large = dict(records=torch.from_numpy(generators.terrain(n=2049, seed=50, payload=True))
                     .to("cuda:0"), payload=True, gridsize=2048)
        torch.cuda.synchronize()
        # plain service time of the large job (the plain loop's per-job times otherwise)
        self._plain(large)
        torch.cuda.synchronize()
        t = time.perf_counter()
        self._plain(large)
        large_ms = (time.perf_counter() - t) * 1e3
        rng = np.random.default_rng(5)
        rates = (50, 200, 1000, 2000, 4000, 6000)
        rows = []
        # the collector thread must get the interpreter promptly while the arrival loop spins
        sys.setswitchinterval(0.0002)
        srv = SvoServer(0, budget_mb=200 * 1024)
        # warm the slots (allocations) with one pass over a few jobs
        for tk in [srv.submit(j) for j in self.jobs[:16] + [large]]:
            srv.wait(tk)
            srv.release(tk)
        for lam in rates:
            n = max(1, int(lam * seconds))
            gaps = rng.exponential(1.0 / lam, size=n)
            arrive = np.cumsum(gaps)
            pick = [(-1 if rng.uniform() < 0.02 else int(rng.integers(0, len(self.jobs))))
                    for _ in range(n)]
            import threading
            lat = []
            done_t = []
            tickets = []
            cv = threading.Condition()

            failed = []

            def collector():
                k = 0
                while k < n:
                    with cv:
                        while len(tickets) <= k:
                            cv.wait()
                        tk = tickets[k]
                    try:
                        r, ms = srv.wait(tk)
                        lat.append(ms)
                    except Exception as e:  # counted and reported, the sweep goes on
                        failed.append(repr(e))
                    done_t.append(time.perf_counter())
                    srv.release(tk)
                    k += 1
            st0 = srv.stats()
            th = threading.Thread(target=collector)
            th.start()
            t0 = time.perf_counter()
            for i in range(n):
                while True:  # sleep through long gaps, spin through short ones
                    left = arrive[i] - (time.perf_counter() - t0)
                    if left <= 0:
                        break
                    if left > 0.002:
                        time.sleep(left - 0.001)
                tk = srv.submit(large if pick[i] < 0 else self.jobs[pick[i]])
                with cv:
                    tickets.append(tk)
                    cv.notify()
            th.join()
            span = max(done_t) - t0
            st1 = srv.stats()
            # the plain loop under the same arrivals: FIFO, measured per-job service times
            svc = np.array([large_ms if p < 0 else self.plain_ms[p] for p in pick]) / 1e3
            fin = np.empty(n)
            tcur = 0.0
            for i in range(n):
                tcur = max(tcur, arrive[i]) + svc[i]
                fin[i] = tcur
            plat = (fin - arrive) * 1e3
            row = {"rate_jobs_per_s": lam, "jobs": n, "large_jobs": int(sum(p < 0 for p in pick)),
                   "failed_jobs": len(failed), "first_failure": failed[0] if failed else None,
                   "jobs_per_s": n / span, "p50_ms": _pct(lat, 50), "p99_ms": _pct(lat, 99),
                   "p999_ms": _pct(lat, 99.9),
                   "mean_batch_jobs": (st1.batched_jobs - st0.batched_jobs) /
                   max(1, st1.batches - st0.batches),
                   "mean_batch_ms": (st1.batch_ms - st0.batch_ms) /
                   max(1, st1.batches - st0.batches),
                   "peak_admitted_gb": st1.peak_admitted_bytes / 1e9,
                   "plain_jobs_per_s": n / fin[-1], "plain_p50_ms": _pct(plat, 50),
                   "plain_p99_ms": _pct(plat, 99), "plain_p999_ms": _pct(plat, 99.9)}
            rows.append(row)
            _log("served %d/s: %.1f jobs/s p50 %.2f p99 %.2f ms (plain p99 %.1f ms)" % (
                lam, row["jobs_per_s"], row["p50_ms"], row["p99_ms"], row["plain_p99_ms"]))
        srv.close()
        cap_plain = len(self.jobs) / self.plain_s
        return {"budget_gb": 200, "slots": 2, "window_s": seconds, "sweep": rows,
                "plain_capacity_jobs_per_s": cap_plain,
                "large_job_plain_ms": large_ms,
                "plain_latency_model": "FIFO queue served at the measured per-job svo_run times"}


if __name__ == "__main__":
    sys.exit(main())
