#!/usr/bin/env python3
"""Throughput of the partition -> voxelize -> SVO-build pipeline on one MI355X.

A step is one full pass of the pipeline (partition, voxelize every partition, build the
octree) over a synthetic mesh already resident in HBM, with the octree left in HBM.
The default workload is the 2048^3 binary headline: a 1025 x 1025-vertex fBm terrain
(2,097,152 triangles, geometry only) voxelized at gridsize 2048 with -l 8192 (one
partition). Prints one JSON line.

The CPU baseline is the single-threaded CPU restatement of the reference algorithm
(oracle/), timed on the same host on the same workload.

More sections:
  * cli: the drop-in CLI end to end (.tri/.tridata file in -> .octree* files out) on the
    headline mesh, next to the oracle CLI on the same file (one core), files compared;
  * out_of_core: the 5x10^7-triangle sphere soup at -s 8192 -l 65536 (8 partitions of
    4096^3), streamed build with the records collected in HBM;
  * batched and served, on a 512-mesh mix (icosphere / terrain 50/50, 1k-2M triangles
    log-uniform, gridsize 64-1024, 30% payload; meshes resident in HBM):
      - batched: the whole mix in one svo_batch_run against a plain per-mesh loop of
        svo_run, every job's node and data bytes compared between the two;
      - served: open-loop Poisson arrivals (50-6000 jobs/s, plus 2% large jobs: a 2049^2
        terrain, 8,388,608 triangles, payload, gridsize 2048) into svo_server with a
        200 GB admission budget; p50/p99/p999 latency and jobs/s, next to a plain loop
        (svo_run, one job after another in arrival order) measured under the same
        arrivals, and a sample of the served results byte-compared with the plain path.
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
    "soup8192": ("soup", False, 8192, 65536,
                 "sphere soup 5e7 tris (200 spheres, edges log-normal median 4 voxels), "
                 "binary, -s 8192 -l 65536 (8 partitions)"),
}

# MI355X: 256 CUs at 2.4 GHz; a wave64 VALU instruction issues at one per CU per cycle
# (4 SIMDs, 4 cycles each), so the non-packed VALU issue peak is 256 * 2.4e9 per second
VALU_ISSUE_PEAK = 256 * 2.4e9

COUNTER_FILES = {"terrain2048": "hbm_traffic_terrain2048_binary.md"}


def _git_blob_sha(path):
    import hashlib
    data = open(path, "rb").read()
    return hashlib.sha1(b"blob %d\0" % len(data) + data).hexdigest()


def measured_counters(workload, timer):
    """From the committed rocprofv3 counter summary of this workload, for the kernel behind
    'timer': HBM bytes per launch, VALU instructions per launch, the SIMD VALU issue
    share, the summary's file name and whether the kernel sources it names are unchanged
    since (their git blob hashes)."""
    out = {"traffic": None, "valu_insts": None, "valu_share": None, "from": None,
           "current": None}
    name = COUNTER_FILES.get(workload)
    path = os.path.join(ROOT, "profiles", name) if name else None
    if not path or not os.path.exists(path):
        return out
    out["from"] = "profiles/" + name
    kernel = None
    current = True
    for line in open(path):
        if line.startswith("sources:"):
            for item in line.split(":", 1)[1].split(","):
                src, _, sha = item.strip().partition(" ")
                fp = os.path.join(ROOT, src)
                current = current and os.path.exists(fp) and _git_blob_sha(fp) == sha.strip()
            out["current"] = current
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        try:
            if len(cells) == 8 and cells[1] == timer:
                kernel = cells[0]
                out["traffic"] = float(cells[7]) * 1e6
            elif len(cells) == 9 and kernel and cells[0] == kernel and cells[8] != "nan":
                out["valu_insts"] = float(cells[2])
                out["valu_share"] = float(cells[8])
        except ValueError:
            pass
    return out


def host_cpu():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count(), len(os.sched_getaffinity(0))


def make_mesh(kind, payload):
    from ooc_svo_builder_amd import generators
    if kind == "terrain":
        return generators.terrain(n=1025, seed=2, payload=payload)
    if kind == "soup":
        return generators.sphere_soup(50_000_000, workers=max(1, min(16, os.cpu_count() or 1)))
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
    ap.add_argument("--no-cli", action="store_true", help="skip the CLI end-to-end section")
    ap.add_argument("--no-ooc", action="store_true", help="skip the out-of-core section")
    ap.add_argument("--only", choices=["headline", "batched", "served", "cli", "ooc"],
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
    if args.only in ("batched", "served", "cli", "ooc"):
        args.no_cpu = True
        args.no_batched = args.only != "batched"
        args.no_served = args.only != "served"
        args.no_cli = args.only != "cli"
        args.no_ooc = args.only != "ooc"
        args.steps, args.warmup = 1, 0
    elif args.only == "headline":
        args.no_batched = args.no_served = args.no_cli = args.no_ooc = True
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
    cnt = measured_counters(args.workload, dom)
    valu_issue = (cnt["valu_insts"] / (avg_launch_ms * 1e-3 * VALU_ISSUE_PEAK)
                  if cnt["valu_insts"] else None)
    # the binding roof: VALU issue when the kernel's SIMDs spend most cycles issuing
    # vector instructions (SQ counters), else HBM
    bound = "valu" if (cnt["valu_share"] or 0) > 0.8 else "hbm"

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
        "roofline": {"bound": bound, "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": cnt["traffic"], "hbm_frac": achieved / HBM_PEAK_GBS,
                     "valu_issue_frac": valu_issue, "simd_valu_share": cnt["valu_share"],
                     "valu_insts_per_launch": cnt["valu_insts"],
                     "counters_from": cnt["from"], "counters_current": cnt["current"],
                     "avg_launch_ms": avg_launch_ms,
                     "algo_bytes_per_launch": algo_bytes_per_launch},
        "launches_per_step": int(info.launches), "host_syncs_per_step": int(info.host_syncs),
        "wall_s_timed": wall,
    }

    if not args.no_cpu:
        import oracle
        model, ncpu, naff = host_cpu()
        # one core, like the reference's single-threaded build: this thread pinned to
        # one CPU of the ones it may use
        keep = os.sched_getaffinity(0)
        core = min(keep)
        os.sched_setaffinity(0, {core})
        try:
            t = time.perf_counter()
            nodes_w, data_w, hz_w, np_w, nf_w = oracle.pipeline(tris, payload, G, L)
            cpu_s = time.perf_counter() - t
        finally:
            os.sched_setaffinity(0, keep)
        # same workload, so the CPU result doubles as a parity check of this run
        same = (nf_w == V and len(nodes_w) == n_nodes and len(data_w) == n_data and
                np.array_equal(ctx.nodes(n_nodes), nodes_w) and
                np.array_equal(ctx.data(n_data), data_w))
        out["cpu_baseline"] = {"value": N / cpu_s, "unit": "triangles/s", "cores": 1,
                               "kind": "port",
                               "sample": "the full workload once (%d triangles), %.1f s, "
                                         "pinned to CPU %d" % (N, cpu_s, core),
                               "host_cpu": model, "host_nproc": ncpu,
                               "host_cpus_allowed": naff,
                               "gpu_speedup": value / (N / cpu_s),
                               "bit_exact_vs_gpu": bool(same)}
    ctx.close()
    del d_tris
    torch.cuda.empty_cache()

    if not args.no_cli and args.workload == "terrain2048":
        try:
            out["cli"] = cli_end_to_end(tris, G, L)
        except Exception as e:  # the headline line is printed regardless
            out["cli"] = {"error": repr(e)}
    if not args.no_ooc and args.workload != "soup8192":
        try:
            out["out_of_core"] = out_of_core(max(2, min(args.steps, 3)))
        except Exception as e:
            out["out_of_core"] = {"error": repr(e)}

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


def cli_end_to_end(tris, G, L):
    """The drop-in CLI on the headline mesh: .tri/.tridata in, .octree* out (wall clock,
    process start and device set-up included), twice; the oracle CLI on the same file,
    pinned to one core; the output files compared."""
    import shutil
    import subprocess
    import tempfile

    import oracle
    from ooc_svo_builder_amd import CLI_BINARY, write_tri
    from ooc_svo_builder_amd.formats import files_identical
    d = tempfile.mkdtemp(prefix="svo_cli_")
    try:
        res = {"mesh_bytes": int(tris.nbytes)}
        for who in ("gpu", "ref"):
            os.makedirs(os.path.join(d, who))
            tri = write_tri(os.path.join(d, who, "terrain"), tris, False)
            cmd = [CLI_BINARY, "-f", tri, "-s", str(G), "-l", str(L)]
            runs = 2
            if who == "ref":
                cmd = ["taskset", "-c", str(min(os.sched_getaffinity(0))),
                       oracle.CLI_BINARY] + cmd[1:]
                runs = 1
            times = []
            for _ in range(runs):
                t = time.perf_counter()
                r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
                times.append(time.perf_counter() - t)
                if r.returncode != 0:
                    raise RuntimeError("%s CLI failed: %s" % (who, r.stderr[-500:]))
            res[who + "_s"] = times
        from ooc_svo_builder_amd import estimate_partitions
        base = "terrain%d_%d" % (G, estimate_partitions(G, L))
        res["identical"] = files_identical(os.path.join(d, "gpu", base),
                                           os.path.join(d, "ref", base))
        res["nodes_bytes"] = os.path.getsize(os.path.join(d, "gpu", base + ".octreenodes"))
        res["speedup_vs_oracle_cli"] = res["ref_s"][0] / min(res["gpu_s"])
        _log("cli: gpu %s s, oracle %.1f s, identical %s" % (
            ["%.2f" % x for x in res["gpu_s"]], res["ref_s"][0], res["identical"]))
        return res
    finally:
        shutil.rmtree(d, ignore_errors=True)


def out_of_core(steps):
    """The 5x10^7-triangle sphere soup at -s 8192 -l 65536: 8 partitions of 4096^3, each
    voxelized in a 64 GiB dense grid and built alone, the records collected in HBM."""
    import torch

    from ooc_svo_builder_amd import SvoContext
    t = time.perf_counter()
    tris = make_mesh("soup", False)
    gen_s = time.perf_counter() - t
    N = tris.shape[0]
    d = torch.from_numpy(tris).cuda()
    del tris
    torch.cuda.synchronize()
    ctx = SvoContext(0)
    try:
        ctx.load_mesh(d, False)
        info = ctx.rerun(False, 8192, 65536)  # warm-up: allocations, the grid's first clear
        torch.cuda.synchronize()
        t = time.perf_counter()
        st = {"partition": 0.0, "voxelize": 0.0, "build": 0.0}
        for _ in range(steps):
            info = ctx.rerun(False, 8192, 65536)
            st["partition"] += info.ms_partition
            st["voxelize"] += info.ms_voxelize
            st["build"] += info.ms_build
        torch.cuda.synchronize()
        s_step = (time.perf_counter() - t) / steps
        st = {k: v / steps for k, v in st.items()}
        V, nn = int(info.nfilled), int(info.n_nodes)
        R = 36
        vox_bytes = N * R + 8 * V
        build_bytes = 8 * V + 24 * nn + 32 * int(info.n_data)
        part_bytes = 2 * N * R + 4 * N  # read twice, lists written
        res = {"workload": WORKLOADS["soup8192"][4], "triangles": N, "steps": steps,
               "generate_s": gen_s, "s_per_step": s_step, "triangles_per_s": N / s_step,
               "partitions": int(info.n_partitions), "voxels": V, "nodes": nn,
               "voxels_per_s": V / s_step, "nodes_per_s": nn / s_step, "stage_ms": st,
               "stage_hbm_frac": {
                   "partition": part_bytes / (st["partition"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                   "voxelize": vox_bytes / (st["voxelize"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                   "build": build_bytes / (st["build"] * 1e-3) / 1e9 / HBM_PEAK_GBS},
               "launches_per_step": int(info.launches),
               "host_syncs_per_step": int(info.host_syncs)}
        _log("out of core: %.3f s per step, %d voxels, %d nodes" % (s_step, V, nn))
        return res
    finally:
        ctx.close()
        del d
        torch.cuda.empty_cache()


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
            self.want.append(self._keep(self._plain(j)))

    def _keep(self, info):
        """The plain path's records of the job just run, kept on the device."""
        torch = self.torch
        nodes = torch.empty(info.n_nodes * 24, dtype=torch.uint8, device="cuda:0")
        data = torch.empty(info.n_data * 32, dtype=torch.uint8, device="cuda:0")
        self.ctx.copy_nodes_to(nodes.data_ptr(), 0, info.n_nodes)
        self.ctx.copy_data_to(data.data_ptr(), 0, info.n_data)
        return (info.n_nodes, info.n_data, info.nfilled, nodes, data)

    def _plain(self, j):
        return self.ctx.run(j["records"], j["payload"], j["gridsize"])

    def drop_outputs(self, keep_every=16):
        """Free the plain outputs but every keep_every-th job's (the served sample)."""
        self.want = [w if i % keep_every == 0 else None for i, w in enumerate(self.want)]
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

        def timed():
            times = []
            for _ in range(reps):
                torch.cuda.synchronize()
                t = time.perf_counter()
                r, i = b.batch_run(self.jobs)
                times.append(time.perf_counter() - t)
            return min(times), r, i
        # the same batch without graphs (every kernel launched by the host), then with
        # them (the default: kernel sequences replayed from the graphs the warm-up kept)
        b.set_graphs(False)
        best_eager, _, info_eager = timed()
        b.set_graphs(True)
        b.batch_run(self.jobs)  # captures the graphs
        best, res, info = timed()
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
                "launches_per_batch": int(info.launches),
                "host_syncs_per_batch": int(info.host_syncs),
                "graph_launches_per_batch": int(info.graph_launches),
                "graph_replays_per_batch": int(info.graph_replays),
                "kernels_in_graphs_per_batch": int(info.graph_kernels),
                "eager": {"seconds": best_eager, "jobs_per_s": n / best_eager,
                          "launches_per_batch": int(info_eager.launches),
                          "host_syncs_per_batch": int(info_eager.host_syncs)},
                "identical_jobs": same, "identical": same == n,
                "autotuned": {"small_cols": knobs[0], "direct_cols": knobs[1],
                              "tile_blocks": knobs[2], "large_blocks": knobs[3]},
                "voxels": int(info.nfilled), "nodes": int(info.n_nodes)}

    def served(self, seconds):
        np, torch = self.np, self.torch
        import threading

        from ooc_svo_builder_amd import SvoServer, generators
        large = dict(records=torch.from_numpy(generators.terrain(n=2049, seed=50, payload=True))
                     .to("cuda:0"), payload=True, gridsize=2048)
        torch.cuda.synchronize()
        self._plain(large)
        large_want = self._keep(self._plain(large))
        t = time.perf_counter()
        self._plain(large)
        large_ms = (time.perf_counter() - t) * 1e3
        rng = np.random.default_rng(5)
        rates = (50, 200, 1000, 2000, 4000, 6000, 9000, 12000)
        rows = []
        srv = SvoServer(0, budget_mb=200 * 1024)
        # warm the slots: one burst of the whole mix and a few large jobs, so that their
        # scratch reaches the sizes of the largest batches before anything is timed (a
        # batch larger than any before grows the slot's buffers with the device drained;
        # in a cold server such batches left the GPU idle for 1-2 s each, measured)
        for tk in [srv.submit(j) for j in self.jobs + [large] * 8]:
            srv.wait(tk)
            srv.release(tk)
        for _ in range(4):  # and large jobs alone (the plain path's buffers of each slot)
            tk = srv.submit(large)
            srv.wait(tk)
            srv.release(tk)

        def wait_until(t0, at):
            # sleep, never spin: a spinning loop holds the interpreter lock that the
            # collector needs to release finished jobs (whose held results count against
            # admission). A late wake-up submits the overdue arrivals at once, so the
            # arrival times stay those of the schedule.
            left = at - (time.perf_counter() - t0)
            if left > 0:
                time.sleep(left)

        for lam in rates:
            n = max(1, int(lam * seconds))
            arrive = np.cumsum(rng.exponential(1.0 / lam, size=n))
            pick = [(-1 if rng.uniform() < 0.02 else int(rng.integers(0, len(self.jobs))))
                    for _ in range(n)]
            lat, done_t, tickets, failed, late = [], [], [], [], []
            stride = max(1, n // 48)
            checked = [0, 0]  # sampled results compared, identical
            cv = threading.Condition()

            def collector():
                for k in range(n):
                    with cv:
                        while len(tickets) <= k:
                            cv.wait()
                        tk = tickets[k]
                    try:
                        r, ms = srv.wait(tk)
                        lat.append(ms + late[k])  # from the scheduled arrival
                        w = large_want if pick[k] < 0 else self.want[pick[k]]
                        # a bounded sample (the copies and compares slow the releases,
                        # and with them admission, at high rates)
                        if w is not None and k % stride == 0 and checked[0] < 24:
                            nn, nd, _, wn, wd = w
                            gn = torch.empty_like(wn)
                            gd = torch.empty_like(wd)
                            srv.copy_nodes_to(tk, gn.data_ptr(), nn)
                            srv.copy_data_to(tk, gd.data_ptr(), nd)
                            checked[0] += 1
                            checked[1] += int((r.n_nodes, r.n_data) == (nn, nd) and
                                              torch.equal(gn, wn) and torch.equal(gd, wd))
                    except Exception as e:  # counted and reported, the sweep goes on
                        failed.append(repr(e))
                    done_t.append(time.perf_counter())
                    srv.release(tk)
            st0 = srv.stats()
            th = threading.Thread(target=collector)
            th.start()
            t0 = time.perf_counter()
            for i in range(n):
                wait_until(t0, arrive[i])
                lt = (time.perf_counter() - t0 - arrive[i]) * 1e3
                tk = srv.submit(large if pick[i] < 0 else self.jobs[pick[i]])
                with cv:
                    late.append(lt)
                    tickets.append(tk)
                    cv.notify()
            th.join()
            span = max(done_t) - t0
            st1 = srv.stats()
            # the plain loop under the same arrivals: svo_run one job after another in
            # arrival order (a job waits while the loop is busy), measured
            plat = []
            t0 = time.perf_counter()
            for i in range(n):
                wait_until(t0, arrive[i])
                self._plain(large if pick[i] < 0 else self.jobs[pick[i]])
                plat.append((time.perf_counter() - t0 - arrive[i]) * 1e3)
            plain_span = time.perf_counter() - t0
            row = {"rate_jobs_per_s": lam, "jobs": n, "large_jobs": int(sum(p < 0 for p in pick)),
                   "failed_jobs": len(failed), "first_failure": failed[0] if failed else None,
                   "jobs_per_s": n / span, "p50_ms": _pct(lat, 50), "p99_ms": _pct(lat, 99),
                   "p999_ms": _pct(lat, 99.9),
                   "mean_batch_jobs": (st1.batched_jobs - st0.batched_jobs) /
                   max(1, st1.batches - st0.batches),
                   "mean_batch_ms": (st1.batch_ms - st0.batch_ms) /
                   max(1, st1.batches - st0.batches),
                   "peak_admitted_gb": st1.peak_admitted_bytes / 1e9,
                   "sampled_results": checked[0], "sampled_identical": checked[1],
                   "plain_jobs_per_s": n / plain_span, "plain_p50_ms": _pct(plat, 50),
                   "plain_p99_ms": _pct(plat, 99), "plain_p999_ms": _pct(plat, 99.9)}
            rows.append(row)
            _log("served %d/s: %.1f jobs/s p50 %.2f p99 %.2f ms; plain %.1f jobs/s p99 %.1f ms;"
                 " %d/%d sampled identical" % (
                     lam, row["jobs_per_s"], row["p50_ms"], row["p99_ms"],
                     row["plain_jobs_per_s"], row["plain_p99_ms"], checked[1], checked[0]))
        budget = srv.stats().budget_bytes
        srv.close()
        cap_plain = len(self.jobs) / self.plain_s
        peak = max(r["jobs_per_s"] for r in rows)
        peak_plain = max(r["plain_jobs_per_s"] for r in rows)
        return {"budget_gb": 200, "budget_gb_in_force": budget / 1e9, "slots": 2,
                "window_s": seconds, "sweep": rows,
                "plain_capacity_jobs_per_s": cap_plain,
                "peak_served_jobs_per_s": peak, "peak_vs_plain_capacity": peak / cap_plain,
                # the same arrivals (mix and large jobs) through the plain loop, at its peak
                "peak_plain_same_arrivals_jobs_per_s": peak_plain,
                "peak_vs_plain_same_arrivals": peak / peak_plain,
                "large_job_plain_ms": large_ms,
                "plain_latency_model": "measured: svo_run one job after another in arrival order",
                "sampled_check": "per rate up to 24 results (jobs at even spacing in the "
                                 "stream that are large or among every 16th mix job) "
                                 "byte-compared with the plain path"}


if __name__ == "__main__":
    sys.exit(main())
