#!/usr/bin/env python3
"""Benchmark of the MI355X path-tracing backend: Msamples/s of
RenderSampleBatch on one GPU.

A step is one full-frame batch (every pixel, all samples of the workload)
rendered into device-resident AOV buffers (inputs and outputs stay in HBM;
no host copies inside the timed region).  Default workload: c3 -- 3840x2160,
16 spp, max depth 16 with roulette from depth 3, 5,000,000 triangles
(displaced spheres and boxes), a 2048x1024 importance-sampled dome plus a
directional light, a 40/30/20/10 diffuse/metal/dielectric/coated mix (see
shiro_amd/workloads.py): the scene on which the wavefront form is compared
with the plain megakernel.  Prints one JSON line with the throughput, the
roofline of the dominant render kernel, the megakernel's time on the same
frame and a CPU baseline (the CPU oracle on this host's cores, bounded
sample).

    python bench.py [--steps K] [--warmup W] [--workload c3|c1|c2|c4|c5]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# The headline metric: render throughput of the wavefront form against the
# plain megakernel on frames that agree, and adjoint gradient evaluations per
# second (the c5 workload reports the latter under the same name).
METRIC = "Msamples/s (wavefront vs megakernel) at matched relMSE; adjoint grads/s"
HBM_PEAK_GBS = 8000.0
NODE_BYTES, TRI_BYTES, TEXEL_BYTES, CDF_BYTES, HIT_BYTES = 64, 48, 16, 4, 96


def workload_desc(wl):
    s = wl.settings
    return (f"{wl.name}: {s.width}x{s.height}, {s.samplesPerPixel} spp, max depth {s.maxDepth}, "
            f"{wl.scene.triangle_count} triangles, {len(wl.scene.domeLights)} dome / "
            f"{len(wl.scene.distantLights)} directional lights")


def _tail_name(depth):
    """tuning.tailDepth or tuning.fetchMode as reported: the value, or "auto"."""
    from shiro_amd import abi
    return "auto" if depth == abi.TAIL_AUTO else depth


def host_cpu():
    """CPU model and logical CPU counts of this host (the CPU baseline's
    machine), from /proc/cpuinfo and the process's affinity mask."""
    import os
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}; {os.cpu_count()} logical CPUs, {len(os.sched_getaffinity(0))} available to this process"


def cpu_baseline(wl, target_seconds, threads):
    """Times the CPU oracle (binary SAH BVH, this host's cores) on a band of
    full-width rows of the same frame at the full spp."""
    from oracle.bindings import OracleScene
    t0 = time.perf_counter()
    orc = OracleScene(wl.scene, bvh=True)
    build_s = time.perf_counter() - t0
    st = wl.settings
    spp = st.samplesPerPixel
    orc.render_region(wl.camera, st, 0, st.height // 2, st.width, 1, 0, 1, threads)  # warm caches and threads

    def timed(rows):
        y0 = (st.height - rows) // 2
        t0 = time.perf_counter()
        orc.render_region(wl.camera, st, 0, y0, st.width, rows, 0, spp, threads)
        return y0, max(time.perf_counter() - t0, 1e-3)

    # Grow the band of rows until it takes about target_seconds; the last
    # band timed is the one reported.
    rows2 = 4
    y0, dt = timed(rows2)
    for _ in range(3):
        if dt >= 0.6 * target_seconds or rows2 >= st.height:
            break
        rows2 = int(min(st.height, max(rows2 + 1, rows2 * target_seconds / dt)))
        y0, dt = timed(rows2)
    samples = st.width * rows2 * spp
    orc.close()
    return {
        "value": samples / dt / 1e6,
        "unit": "Msamples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"rows {y0}-{y0 + rows2 - 1} of the same frame ({st.width}x{rows2} px x {spp} spp = "
                  f"{samples} samples) in {dt:.1f} s; CPU oracle with a binary SAH BVH "
                  f"(built in {build_s:.1f} s, not timed), row-parallel threads",
        "per_core": samples / dt / 1e6 / threads,
        "host": host_cpu(),
    }


def bench_c4(args):
    """Aggregate Msamples/s of the c4 batch (32 scenes, 256-1024 px square,
    4/16/64 spp, 10k-500k triangles, own 1024x512 dome each) in the fastest
    batch mode; every mode (persistent work-list megakernel with autotuned
    grid, replayed hipGraph of per-scene megakernel launches, sequential
    launches, per-scene wavefront, and the wavefront of all scenes in one
    captured graph of autotuned concurrent branches, replayed) is timed and
    reported beside it."""
    import torch
    from shiro_amd import abi, workloads
    from shiro_amd.hip import Context
    t0 = time.perf_counter()
    wls = workloads.c4()
    gen_s = time.perf_counter() - t0
    ctxs, frames = [], []
    t0 = time.perf_counter()
    for wl in wls:
        c = Context(0)
        c.upload(wl.scene)
        ctxs.append(c)
        n = wl.settings.width * wl.settings.height
        frames.append([torch.empty(n * k, dtype=torch.float32, device="cuda") for k in (4, 3, 3, 1)])
    upload_s = time.perf_counter() - t0
    launcher = Context(0)
    items = Context.batch_items([(c, wl.camera, wl.settings, [b.data_ptr() for b in f], 0, None)
                                 for wl, c, f in zip(wls, ctxs, frames)])
    samples = sum(wl.samples for wl in wls)
    tris = sum(wl.scene.triangle_count for wl in wls)
    print(f"# c4: {len(wls)} scenes, {tris} triangles, {samples} samples per batch; generated in {gen_s:.1f} s, "
          f"built and uploaded in {upload_s:.1f} s", file=sys.stderr, flush=True)
    best, report = launcher.autotune_many(items)
    print("# batch autotune (persistent grid):\n" + report, file=sys.stderr, flush=True)
    modes = {}
    names = {"persistent": abi.MANY_PERSISTENT, "graph": abi.MANY_GRAPH, "sequential": abi.MANY_SEQUENTIAL,
             "wavefront": abi.MANY_WAVEFRONT, "wavefront_graph": abi.MANY_WAVEFRONT_GRAPH}
    for name, mode in names.items():
        launcher.render_many(items, mode)
        modes[name] = round(launcher.render_many(items, mode), 3)
        print(f"# {name}: {modes[name]} ms", file=sys.stderr, flush=True)
    fastest = min(modes, key=modes.get)
    for i in range(args.warmup):
        launcher.render_many(items, names[fastest])
    torch.cuda.synchronize()
    gpu_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        gpu_ms.append(launcher.render_many(items, names[fastest]))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cpu = None
    if not args.no_cpu:
        wl = sorted(wls, key=lambda w: w.samples)[len(wls) // 2]
        cpu = cpu_baseline(wl, args.cpu_seconds, args.cpu_threads)
        cpu["sample"] = f"{wl.name}: " + cpu["sample"]
    out = {
        "metric": METRIC, "value": round(samples * args.steps / wall / 1e6, 3), "unit": "Msamples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (32 seeded scenes and domes)",
        "config": {"workload": f"c4: {len(wls)} scenes, {tris} triangles, {samples} samples per step",
                   "batch_mode": fastest, "batch_ms": modes,
                   "speedup_vs_sequential": round(modes["sequential"] / (sum(gpu_ms) / len(gpu_ms)), 3),
                   "tuning": {"batchBlocksPerCU": best.batchBlocksPerCU, "manyBranches": best.manyBranches,
                              "tailDepth": _tail_name(best.tailDepth)},
                   "graph_captures": launcher.many_graph_stats()[0]},
        "roofline": None, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_c5(args):
    """Gradient evaluations per second of the texture extension's
    differentiable render (c5: 256x256, 32 spp forward + 32 spp path-replay
    adjoint, depth 4, 50,002 triangles, 128x128 base-colour and roughness
    textures and the 256x128 dome map as parameters): one evaluation =
    textured forward render (the dome's importance tables rebuilt from the
    current map on the device), MSE to a target image, adjoint into both
    textures and the map, one Adam update, as a PyTorch loop."""
    import numpy as np
    import torch
    from shiro_amd import workloads
    from shiro_amd.diff import TextureRenderer
    wl = workloads.c5()
    st = wl.settings
    alb_t, rough_t = wl.textures_truth
    alb0, rough0 = wl.textures_init
    from shiro_amd.scene import Texture
    import copy
    scene = copy.copy(wl.scene)
    scene.textures = [Texture(alb_t), Texture(rough_t)]
    print(f"# workload {workload_desc(wl)}", file=sys.stderr, flush=True)
    r = TextureRenderer(scene, wl.camera, st, env_domes=[0])
    with torch.no_grad():
        target = r.render(torch.from_numpy(alb_t).cuda(), torch.from_numpy(rough_t).cuda(),
                          env=[torch.from_numpy(wl.env_truth).cuda()]).clone()
    albedo = torch.from_numpy(alb0.copy()).cuda().requires_grad_(True)
    rough = torch.from_numpy(rough0.copy()).cuda().requires_grad_(True)
    sky = torch.from_numpy(wl.env_init.copy()).cuda().requires_grad_(True)
    opt = torch.optim.Adam([{"params": [albedo, rough], "lr": 0.02}, {"params": [sky], "lr": 0.05}])
    spp = st.samplesPerPixel
    losses, fwd_ms, adj_ms = [], [], []

    def step(i):
        opt.zero_grad()
        img = r.render(albedo, rough, env=[sky])
        fwd_ms.append(r.ctx.last_render_ms())
        loss = ((img - target) ** 2).mean()
        loss.backward()
        adj_ms.append(r.ctx.last_render_ms())
        opt.step()
        with torch.no_grad():
            albedo.clamp_(0.0, 1.0)
            rough.clamp_(0.05, 1.0)
            sky.clamp_(min=0.0)
        losses.append(loss.item())

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    fwd_ms.clear()
    adj_ms.clear()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(args.warmup + i)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = args.steps / wall
    cpu = None
    if not args.no_cpu:
        from oracle.bindings import OracleScene
        orc = OracleScene(scene, bvh=True)

        def cpu_eval(rows, reps):
            # `reps` evaluations of a band of `rows` central rows: forward
            # render of the band, adjoint of a loss weighting only the band.
            y0 = (st.height - rows) // 2
            dimg = np.zeros((st.height, st.width, 4), np.float32)
            dimg[y0:y0 + rows, :, :3] = 1.0 / (st.width * st.height * 3)
            t0 = time.perf_counter()
            for _ in range(reps):
                orc.render_region(wl.camera, st, 0, y0, st.width, rows, 0, spp, args.cpu_threads)
                orc.render_adjoint(wl.camera, st, dimg, wanted=[0, 1], env_wanted=[0], threads=args.cpu_threads)
            return y0, time.perf_counter() - t0

        # Calibrate on 32 rows, then size the sample to about cpu_seconds:
        # a taller band, or whole frames repeated.
        _, dt = cpu_eval(32, 1)
        frame_s = dt * st.height / 32
        rows = max(32, min(st.height, int(st.height * args.cpu_seconds / frame_s)))
        reps = max(1, int(round(args.cpu_seconds / frame_s))) if rows == st.height else 1
        y0, dt = cpu_eval(rows, reps)
        cpu = {"value": reps * rows / st.height / dt, "unit": "grads/s", "cores": args.cpu_threads, "kind": "port",
               "sample": f"{reps} x rows {y0}-{y0 + rows - 1} ({st.width}x{rows} px, {spp} spp forward + {spp} spp "
                         f"adjoint into both textures and the map) in {dt:.1f} s, scaled to the full frame; "
                         f"CPU oracle, row-parallel threads",
               "per_core": reps * rows / st.height / dt / args.cpu_threads, "host": host_cpu()}
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "grads/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded sphere, target textures and dome)",
        "config": {"workload": workload_desc(wl),
                   "parameters": "128x128x3 base colour + 128x128x1 roughness + 256x128x3 dome map texels",
                   "forward_ms": round(sum(fwd_ms) / len(fwd_ms), 3), "adjoint_ms": round(sum(adj_ms) / len(adj_ms), 3),
                   "loss_first": losses[0], "loss_last": losses[-1],
                   "adam_steps": len(losses), "loss_curve": [float(f"{x:.5g}") for x in losses]},
        "roofline": None, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def _csv_rows(directory, suffix):
    import csv
    import glob
    rows = []
    for path in sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True)):
        with open(path, newline="") as f:
            rows.extend(csv.DictReader(f))
    return rows


def _kernel_symbol(name):
    """rocprofv3's kernel name without the namespace and the argument list."""
    name = name.strip()
    if name.startswith("void "):
        name = name[5:]
    depth, cut = 0, len(name)
    for i, ch in enumerate(name):  # the argument list is the first '(' outside template brackets
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    return name[:cut].replace("shiro_hip::", "").strip()


def summarize_profile(args):
    """Per-kernel HBM bytes per dispatch from rocprofv3 output directories
    (csv): kernel stats of a --kernel-trace --stats run, FETCH_SIZE and
    WRITE_SIZE runs of the workload, and the same two counters of a device
    copy of known size (bench.py --copy-probe), which give the factors that
    turn the counters into bytes.  Prints a markdown table and the JSON
    record kept in profiles/hbm_traffic.json."""
    stats_dir, fetch_dir, write_dir, copyf_dir, copyw_dir = args.summarize_profile

    def counter_avgs(directory, counter):
        acc = {}
        for r in _csv_rows(directory, "counter_collection.csv"):
            if r.get("Counter_Name") != counter:
                continue
            k = _kernel_symbol(r["Kernel_Name"])
            acc.setdefault(k, []).append(float(r["Counter_Value"]))
        return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}

    fetch, write = counter_avgs(fetch_dir, "FETCH_SIZE"), counter_avgs(write_dir, "WRITE_SIZE")
    cf, cw = counter_avgs(copyf_dir, "FETCH_SIZE"), counter_avgs(copyw_dir, "WRITE_SIZE")
    # The copy kernel reads and writes the buffer (the fill that made it only
    # writes): the kernel of both runs with the largest read.
    copy_kernel = max((k for k in cf if k in cw), key=lambda k: cf[k][0])
    copy_bytes = float(args.copy_bytes)
    fscale = copy_bytes / (cf[copy_kernel][0] * 1024.0)
    wscale = copy_bytes / (cw[copy_kernel][0] * 1024.0)
    stats = {}
    for r in _csv_rows(stats_dir, "kernel_stats.csv"):
        stats[_kernel_symbol(r["Name"])] = (int(r["Calls"]), float(r["TotalDurationNs"]) / 1e6,
                                            float(r["AverageNs"]) / 1e6)
    kernels = {}
    lines = ["| kernel | calls (trace) | avg ms (trace) | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM MB per dispatch | GB/s |",
             "|---|---|---|---|---|---|---|"]
    for k in sorted(stats, key=lambda k: -stats[k][1]):
        calls, total, avg = stats[k]
        if k.startswith("at::"):
            continue  # torch's own kernels (the frame comparisons), not the render
        if k not in fetch or k not in write:
            lines.append(f"| `{k}` | {calls} | {avg:.4f} | - | - | - | - |")
            continue
        b = fetch[k][0] * 1024.0 * fscale + write[k][0] * 1024.0 * wscale
        kernels[k] = {"hbm_bytes_per_dispatch": int(b), "fetch_kib": round(fetch[k][0], 1),
                      "write_kib": round(write[k][0], 1), "avg_ms": round(avg, 4), "calls": calls}
        lines.append(f"| `{k}` | {calls} | {avg:.4f} | {fetch[k][0]:.0f} | {write[k][0]:.0f} | {b / 1e6:.1f} | "
                     f"{b / (avg * 1e-3) / 1e9:.0f} |")
    print(f"Copy calibration ({copy_kernel}, {copy_bytes / 2**30:.2f} GiB): FETCH_SIZE x {fscale:.3f}, "
          f"WRITE_SIZE x {wscale:.3f}\n")
    print("\n".join(lines))
    print()
    print(json.dumps({"fetch_scale": round(fscale, 4), "write_scale": round(wscale, 4), "kernels": kernels},
                     indent=1))


def counter_table(directory):
    """Per kernel symbol: dispatches and the mean of every counter of a
    rocprofv3 --pmc csv directory."""
    acc = {}
    for r in _csv_rows(directory, "counter_collection.csv"):
        k = _kernel_symbol(r["Kernel_Name"])
        d = acc.setdefault(k, {})
        d.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in d.items()} | {"_n": max(len(v) for v in d.values())}
            for k, d in acc.items()}


def summarize_valu(args):
    """VALU issue rate, lane utilisation, memory waits and measured occupancy
    per kernel, against the issue rate the FMA probe (bench.py --valu-probe)
    sustains on the same counters: SQ_INSTS_VALU per second of the probe
    kernel is the ceiling."""
    sq_dir, occ_dir, stats_dir, probe_sq_dir, probe_stats_dir = args.summarize_valu
    sq, occ = counter_table(sq_dir), counter_table(occ_dir)
    stats = {}
    for d in (stats_dir, probe_stats_dir):
        for r in _csv_rows(d, "kernel_stats.csv"):
            stats[_kernel_symbol(r["Name"])] = (int(r["Calls"]), float(r["AverageNs"]) / 1e6)
    # The probe runs a short warm-up dispatch and the measured one: take the
    # largest count and the longest duration (the measured dispatch).
    pk, n_probe = None, 0.0
    for r in _csv_rows(probe_sq_dir, "counter_collection.csv"):
        if r["Counter_Name"] == "SQ_INSTS_VALU" and _kernel_symbol(r["Kernel_Name"]).startswith("k_valu_probe"):
            if float(r["Counter_Value"]) > n_probe:
                pk, n_probe = _kernel_symbol(r["Kernel_Name"]), float(r["Counter_Value"])
    probe_ms = max(float(r["MaxNs"]) for r in _csv_rows(probe_stats_dir, "kernel_stats.csv")
                   if _kernel_symbol(r["Name"]) == pk) / 1e6
    ceiling = n_probe / (probe_ms * 1e-3)
    print(f"VALU issue ceiling: {pk} issues {n_probe:.4g} VALU wave-instructions (SQ_INSTS_VALU) in "
          f"{probe_ms:.3f} ms = {ceiling:.4g} per second\n")
    print("| kernel | avg ms | VALU wave-instr per dispatch | VALU issue rate / ceiling | lane utilisation | "
          "wait share | issue-wait share | MeanOccupancyPerActiveCU |")
    print("|---|---|---|---|---|---|---|---|")
    rows = []
    for k, c in sq.items():
        if k.startswith("at::") or k not in stats or "SQ_INSTS_VALU" not in c:
            continue
        ms = stats[k][1]
        rate = c["SQ_INSTS_VALU"] / (ms * 1e-3) / ceiling
        util = c.get("SQ_THREAD_CYCLES_VALU", 0.0) / max(64.0 * c.get("SQ_ACTIVE_INST_VALU", 1.0), 1.0)
        # Shares of resident wave time: waiting on anything (memory,
        # dependencies, barriers: SQ_WAIT_ANY) and ready but waiting to be
        # issued (SQ_WAIT_INST_ANY).
        cyc = max(c.get("SQ_WAVE_CYCLES", 1.0), 1.0)
        wait = c["SQ_WAIT_ANY"] / cyc if "SQ_WAIT_ANY" in c else None
        iwait = c.get("SQ_WAIT_INST_ANY", 0.0) / cyc
        o = occ.get(k, {}).get("MeanOccupancyPerActiveCU")
        rows.append((ms * stats[k][0], k, ms, c["SQ_INSTS_VALU"], rate, util, wait, iwait, o))
    out = {}
    for _, k, ms, n, rate, util, wait, iwait, o in sorted(rows, reverse=True):
        print(f"| `{k}` | {ms:.4f} | {n:.4g} | {rate:.3f} | {util:.2f} | {'-' if wait is None else f'{wait:.2f}'} | "
              f"{iwait:.2f} | {'-' if o is None else f'{o:.1f}'} |")
        out[k] = {"valu_issue_frac": round(rate, 4), "lane_utilisation": round(util, 3),
                  "wait_share": None if wait is None else round(wait, 3), "issue_wait_share": round(iwait, 3),
                  "mean_occupancy": None if o is None else round(o, 2)}
    print()
    print(json.dumps({"valu_ceiling_per_s": ceiling, "kernels": out}, indent=1))


def _profile_child(out_dir, counters, child_args, seconds=300):
    """Runs this script under `rocprofv3 --pmc <counters>` as a child process
    (csv output in out_dir) and returns the child's last JSON stdout line."""
    import subprocess
    cmd = ["timeout", "-s", "KILL", str(seconds), "rocprofv3", "--pmc", *counters, "--output-format", "csv",
           "-d", out_dir, "-o", "run", "--", sys.executable, os.path.join(ROOT, "bench.py"), *child_args]
    r = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, env={**os.environ, "TMPDIR": "/tmp"})
    if r.returncode != 0:
        raise RuntimeError(f"rocprofv3 {' '.join(counters)}: exit {r.returncode}: {r.stderr[-400:]}")
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    return json.loads(lines[-1]) if lines else None


def measure_counters(args, tuning, kernel, probe_ms):
    """Counter measurements of this run for `kernel` (a kernel symbol of
    the workload): HBM bytes per dispatch (FETCH_SIZE and WRITE_SIZE child
    runs of one frame with the same tuning, scaled by what the same counters
    read for a 1 GiB device copy) and VALU wave-instructions per dispatch
    (SQ counters) against those of the FMA issue-rate probe.  probe_ms(n)
    times the probe of n iterations outside the profiler (a profiled
    dispatch's events include the counter set-up)."""
    import shutil
    import tempfile
    work = tempfile.mkdtemp(prefix="shiro_prof_", dir="/tmp")
    frame = ["--workload", args.workload, "--transport", args.transport, "--no-cpu", "--no-count", "--no-parity",
             "--no-traffic", "--no-e2e", "--steps", "1", "--warmup", "0", "--tuning-json", json.dumps(tuning)]
    copy_mib = 1024
    out = {}
    try:
        for name, counters, child in (("fetch", ["FETCH_SIZE"], frame), ("write", ["WRITE_SIZE"], frame),
                                      ("copyf", ["FETCH_SIZE"], ["--copy-probe", str(copy_mib)]),
                                      ("copyw", ["WRITE_SIZE"], ["--copy-probe", str(copy_mib)])):
            d = os.path.join(work, name)
            _profile_child(d, counters, child)
            out[name] = counter_table(d)
            print(f"# counters {name}: {len(out[name])} kernels", file=sys.stderr, flush=True)
        copy_bytes = copy_mib * (1 << 20)
        copy_kernel = max((k for k in out["copyf"] if k in out["copyw"]),
                          key=lambda k: out["copyf"][k]["FETCH_SIZE"])
        fscale = copy_bytes / (out["copyf"][copy_kernel]["FETCH_SIZE"] * 1024.0)
        wscale = copy_bytes / (out["copyw"][copy_kernel]["WRITE_SIZE"] * 1024.0)
        res = {"fetch_scale": round(fscale, 4), "write_scale": round(wscale, 4)}
        f, w = out["fetch"].get(kernel), out["write"].get(kernel)
        if f and w:
            res["hbm_bytes_per_dispatch"] = int(f["FETCH_SIZE"] * 1024.0 * fscale + w["WRITE_SIZE"] * 1024.0 * wscale)
            res["dispatches_profiled"] = int(f["_n"])
        sq = ["SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_WAVE_CYCLES", "SQ_WAIT_ANY",
              "SQ_WAIT_INST_ANY"]
        d = os.path.join(work, "sq")
        _profile_child(d, sq, frame)
        k = counter_table(d).get(kernel)
        print("# counters sq", file=sys.stderr, flush=True)
        d = os.path.join(work, "probe")
        _profile_child(d, ["SQ_INSTS_VALU"], ["--valu-probe", "100000"])
        # The probe's timed dispatch is its larger one (a short warm-up runs
        # first).
        pn = max((float(r["Counter_Value"]) for r in _csv_rows(d, "counter_collection.csv")
                  if r["Counter_Name"] == "SQ_INSTS_VALU"
                  and _kernel_symbol(r["Kernel_Name"]).startswith("k_valu_probe")), default=0.0)
        if k and pn > 0:
            res["valu_wave_instructions_per_dispatch"] = k["SQ_INSTS_VALU"]
            res["valu_ceiling_per_s"] = pn / (min(probe_ms(100000) for _ in range(3)) * 1e-3)
            res["lane_utilisation"] = round(k["SQ_THREAD_CYCLES_VALU"] / max(64.0 * k["SQ_ACTIVE_INST_VALU"], 1.0), 3)
            res["wait_share"] = round(k["SQ_WAIT_ANY"] / max(k["SQ_WAVE_CYCLES"], 1.0), 3)
            res["issue_wait_share"] = round(k["SQ_WAIT_INST_ANY"] / max(k["SQ_WAVE_CYCLES"], 1.0), 3)
        return res
    finally:
        shutil.rmtree(work, ignore_errors=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 5; c5: 100 Adam steps, whose loss curve is reported)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--transport", default="wavefront", choices=["megakernel", "wavefront", "megakernel_plus"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-count", action="store_true")
    ap.add_argument("--no-parity", action="store_true", help="skip the other transport's comparison frame")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default wavefront tuning")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 child runs that measure the dominant kernel's HBM bytes and VALU issue")
    ap.add_argument("--no-e2e", action="store_true", help="skip the host-buffer render with readback")
    ap.add_argument("--tuning-json", default=None, help="set the wavefront tuning (JSON of its fields), no autotune")
    ap.add_argument("--fetch-mode", type=int, default=None, help="force the traversal fetch mode (0, 1, 2)")
    ap.add_argument("--paths-per-pass", type=int, default=None, help="force the wavefront paths per pass")
    ap.add_argument("--primary-packets", type=int, default=None, choices=[0, 1, 4, 16, 64],
                    help="camera-ray packets off (0) or of 1, 4, 16 or 64 samples per packet")
    ap.add_argument("--bvh-width", type=int, default=None, choices=[4, 8],
                    help="force the tree width the wavefront traversal walks")
    ap.add_argument("--shade-order", type=int, default=None, choices=[0, 1],
                    help="force shading-class ordering of the hit queue off (0) or on (1)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 3 times (calibrates HBM byte counters)")
    ap.add_argument("--summarize-profile", nargs=5, default=None,
                    metavar=("STATS", "FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="summarise rocprofv3 csv output directories into HBM bytes per dispatch (no GPU)")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30, help="bytes the --copy-probe run copied")
    ap.add_argument("--valu-probe", type=int, default=0, metavar="ITERS",
                    help="only run the FMA issue-rate probe (ITERS iterations of 8 FMAs per thread)")
    ap.add_argument("--summarize-valu", nargs=5, default=None,
                    metavar=("SQ", "OCCUPANCY", "STATS", "PROBE_SQ", "PROBE_STATS"),
                    help="summarise rocprofv3 SQ counter directories into VALU issue rates (no GPU)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 100 if args.workload == "c5" else 5
    if args.summarize_profile:
        summarize_profile(args)
        return
    if args.summarize_valu:
        summarize_valu(args)
        return
    if args.valu_probe:
        from shiro_amd.hip import Context
        ctx = Context(0)
        ms, n = ctx.valu_probe(args.valu_probe)
        print(json.dumps({"valu_probe_ms": ms, "wave_instructions": n, "per_second": n / (ms * 1e-3)}), flush=True)
        return
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4, "repeats": 3}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this backend has no multi-GPU path "
              "(a frame shards by sample range or by tile; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "c5":
        bench_c5(args)
        return
    if args.workload == "c4":
        bench_c4(args)
        return
    import torch
    from shiro_amd import abi, workloads
    from shiro_amd.hip import Context
    from shiro_amd.scene import RenderSettings
    wl = workloads.WORKLOADS[args.workload]()
    st = wl.settings
    st.transport = {"wavefront": abi.TRANSPORT_WAVEFRONT, "megakernel": abi.TRANSPORT_MEGAKERNEL,
                    "megakernel_plus": abi.TRANSPORT_MEGAKERNEL_PLUS}[args.transport]
    print(f"# workload {workload_desc(wl)}", file=sys.stderr, flush=True)
    ctx = Context(0)
    t0 = time.perf_counter()
    ctx.upload(wl.scene)
    upload_s = time.perf_counter() - t0
    info = ctx.scene_info()
    print(f"# BVH4 nodes {info.bvhNodeCount} leaves {info.bvhLeafCount} depth {info.bvhDepth} stack {info.bvhStack}; "
          f"8-wide nodes {info.bvh8NodeCount} depth {info.bvh8Depth} stack {info.bvh8Stack}; "
          f"build {info.buildSeconds:.2f}s upload {info.uploadSeconds:.2f}s", file=sys.stderr, flush=True)
    npx = st.width * st.height
    tuning = None
    if args.transport == "wavefront":
        if args.tuning_json:
            ctx.set_tuning(**json.loads(args.tuning_json))
        elif args.fetch_mode is not None or args.paths_per_pass is not None:
            kw = {}
            if args.fetch_mode is not None:
                kw["fetchMode"] = args.fetch_mode
            if args.paths_per_pass is not None:
                kw["pathsPerPass"] = args.paths_per_pass
            ctx.set_tuning(**kw)
        elif not args.no_autotune:
            # Trials render the whole frame: the number of passes a frame
            # takes (paths per pass against spp x pixels) is part of what is
            # tuned, and a shorter trial hides it (c3 at 8 spp renders one
            # pass with 64M and with 128M paths per pass; at 16 spp the
            # first takes two passes and 7% longer).
            trial = st.samplesPerPixel
            t0 = time.perf_counter()
            _, report = ctx.autotune(wl.camera, st, trial)
            print(f"# autotune ({trial} spp trials, {time.perf_counter() - t0:.1f} s):\n" + report, file=sys.stderr,
                  flush=True)
        if args.shade_order is not None:
            ctx.set_tuning(shadeOrder=args.shade_order)
        if args.primary_packets is not None:
            ctx.set_tuning(primaryPackets=args.primary_packets)
        if args.bvh_width is not None:
            ctx.set_tuning(bvhWidth=args.bvh_width)
        t = ctx.get_tuning()
        tuning = {"pathsPerPass": t.pathsPerPass, "fetchMode": _tail_name(t.fetchMode), "traceBlocks": t.traceBlocks,
                  "stageBlocks": t.stageBlocks, "shadeOrder": t.shadeOrder, "primaryPackets": t.primaryPackets,
                  "bvhWidth": t.bvhWidth if info.bvh8NodeCount else 4, "tailDepth": _tail_name(t.tailDepth),
                  "shadowOverlap": t.shadowOverlap,
                  "autotuned": (args.fetch_mode is None and args.paths_per_pass is None and not args.no_autotune
                                and not args.tuning_json)}
    dev = torch.device("cuda:0")
    bufs = [torch.empty(npx * k, dtype=torch.float32, device=dev) for k in (4, 3, 3, 1)]
    ptrs = [b.data_ptr() for b in bufs]
    torch.cuda.synchronize()
    for i in range(args.warmup):
        ctx.render_device(wl.camera, st, ptrs, 0, st.samplesPerPixel, synchronize=True)
        print(f"# warmup {i} {ctx.last_render_ms():.1f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    ctx.synchronize()
    gpu_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctx.render_device(wl.camera, st, ptrs, 0, st.samplesPerPixel, synchronize=True)
        gpu_ms.append(ctx.last_render_ms())
        print(f"# step {i} {gpu_ms[-1]:.1f} ms", file=sys.stderr, flush=True)
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    samples = npx * st.samplesPerPixel
    value = samples * args.steps / wall / 1e6
    ms_per_step = wall * 1e3 / args.steps
    frame = bufs[0].view(-1, 4)
    ok = bool(torch.isfinite(frame).all().item())

    # Roofline of the dominant render kernel.  Algorithmic bytes come from
    # fetch counters: BVH4 nodes 64 B (8-wide 128 B), triangles 48 B, hit
    # data 96 B, environment texels 16 B, CDF entries 4 B, AOV write 44 B per
    # pixel per batch.  The wavefront's traversal kernels count their own
    # fetches (counting instantiations of the same kernels over the whole
    # frame); hit, texel and CDF reads come from the megakernel's counters
    # (the shading is the same in both transports).  Each wavefront kernel
    # symbol is credited with its own bytes: each traversal kernel with the
    # nodes and triangles its load instructions request, each record once
    # per wave instruction that reads it (the camera packets read each node
    # once for 64 rays; in the per-lane kernels the lanes of one instruction
    # that read the same node, rays of one path group walking in step, share
    # one request), with every ray's own fetches summed beside it
    # (frac_per_ray_sum: can exceed 1, since the rays of a path group share
    # their nodes through L2 and the last-level cache as well); the shade
    # kernels with the hit, texel and CDF bytes; queue and path-state
    # traffic earns nothing.  Kernel durations come from events recorded
    # after every launch on the render stream, in one extra (untimed) render
    # call.  The dominant kernel is the symbol with the largest time per
    # frame.  `traffic` is what HBM moved per launch by the rocprofv3
    # counters.
    roofline = None
    stages = None
    if not args.no_count:
        cspp = min(4, st.samplesPerPixel)
        c = ctx.count_fetches(wl.camera, st, 0, cspp)
        ns = max(c.samples, 1)
        camera_b = (c.primaryNodeFetches * NODE_BYTES + c.primaryTriFetches * TRI_BYTES) / ns
        bounce_b = ((c.nodeFetches - c.primaryNodeFetches) * NODE_BYTES +
                    (c.triFetches - c.primaryTriFetches) * TRI_BYTES) / ns
        shadow = (c.shadowNodeFetches * NODE_BYTES + c.shadowTriFetches * TRI_BYTES) / ns
        other = (c.hitFetches * HIT_BYTES + c.envTaps * TEXEL_BYTES + c.cdfSteps * CDF_BYTES) / ns
        wf = None
        if args.transport == "wavefront":
            # The traversal kernels' own counts over the whole frame (the
            # counting instantiations: same traversal, same work
            # distribution).  The
            # camera packets fetch each node and triangle once per packet of
            # 64 rays, so k_wf_primary is credited with its per-packet
            # fetches; the per-lane kernels with every ray's own fetches.
            w = ctx.count_wavefront(wl.camera, st, 0, st.samplesPerPixel)
            nsw = max(w.samples, 1)
            nb = NODE_BYTES * (2 if w.bvhWidth == 8 else 1)

            def per_ray(k):
                return (w.nodeFetches[k] * nb + w.triFetches[k] * TRI_BYTES) / nsw

            def per_wave(k):
                return (w.waveNodeFetches[k] * nb + w.waveTriFetches[k] * TRI_BYTES) / nsw

            camera_b, bounce_b, shadow = per_ray(0), per_ray(1), per_ray(2)
            wf = {"camera_rays": w.rays[0] / nsw, "bounce_rays": w.rays[1] / nsw, "shadow_rays": w.rays[2] / nsw,
                  "camera_packet_bytes": round(per_wave(0), 1),
                  "bytes_per_ray_sum": {"camera": round(camera_b, 1), "bounce": round(bounce_b, 1),
                                        "shadow": round(shadow, 1)},
                  "bytes_per_wave_instruction": {"camera": round(per_wave(0), 1), "bounce": round(per_wave(1), 1),
                                                 "shadow": round(per_wave(2), 1)},
                  "node_fetches_per_ray": [round(w.nodeFetches[k] / max(w.rays[k], 1), 2) for k in range(3)],
                  "tri_tests_per_ray": [round(w.triFetches[k] / max(w.rays[k], 1), 2) for k in range(3)],
                  "shadow_occluded": round(w.shadowOccluded / max(w.rays[2], 1), 4)}
        per_sample = camera_b + bounce_b + shadow + other + 44.0 / st.samplesPerPixel
        ctx.set_stage_timing(True)
        ctx.render_device(wl.camera, st, ptrs, 0, st.samplesPerPixel, synchronize=True)
        stages = ctx.stage_times()
        ctx.set_stage_timing(False)
        fm, bw = (tuning["fetchMode"], tuning["bvhWidth"]) if tuning else (0, 4)
        if args.transport == "wavefront":
            tex = "true" if wl.scene.textures else "false"
            plain = "true" if ctx.scene_info().plainBsdf else "false"
            shade_first = f"k_wf_shade<{tex}, true, {plain}>"
            shade_later = f"k_wf_shade<{tex}, false, {plain}>"
            wv = wf["bytes_per_wave_instruction"]
            packets = bool(stages.get("primary"))
            kernels = {
                "primary": (f"k_wf_primary<{bw}, false>", wv["camera"], camera_b),
                "extend": (f"k_wf_trace<false, {fm}, {bw}, false>",
                           wv["bounce"] if packets else wv["bounce"] + wv["camera"],
                           bounce_b if packets else bounce_b + camera_b),
                "shadow": (f"k_wf_trace<true, {fm}, {bw}, false>", wv["shadow"], shadow),
                "shade_first": (shade_first, None, None),
                "shade": (shade_later, None, None),
            }
        else:
            sym = "k_render_plus<false>" if args.transport == "megakernel_plus" else "k_render<false, false>"
            kernels = {"megakernel": (sym, per_sample, per_sample)}
        entries = []
        for stage, (sym, credit, ray_credit) in kernels.items():
            if stage not in stages:
                continue
            ms, n = stages[stage]["ms"], stages[stage]["launches"]
            e = {"kernel": sym, "stage": stage, "launches_per_step": n, "ms_per_step": round(ms, 3),
                 "avg_launch_ms": round(ms / n, 4)}
            if credit is not None:
                bpl = credit * samples / n
                e["bytes_per_sample"] = round(credit, 1)
                e["bytes_per_launch"] = int(bpl)
                e["achieved"] = round(bpl / (ms / n * 1e-3) / 1e9, 2)
                e["frac"] = round(e["achieved"] / HBM_PEAK_GBS, 5)
                # Every ray charged its own fetches (no merging of lanes
                # that read the same record in one instruction).
                e["bytes_per_sample_per_ray_sum"] = round(ray_credit, 1)
                e["frac_per_ray_sum"] = round(ray_credit * samples / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)
            entries.append(e)
        dom = max(entries, key=lambda e: e["ms_per_step"])
        if "achieved" not in dom:
            # A shade kernel: credited with the hit, texel and CDF bytes of
            # every bounce in proportion to its share of the shade time.
            sh = sum(e["ms_per_step"] for e in entries if e["stage"] in ("shade", "shade_first"))
            credit = other * dom["ms_per_step"] / sh
            bpl = credit * samples / dom["launches_per_step"]
            dom.update(bytes_per_sample=round(credit, 1), bytes_per_launch=int(bpl),
                       achieved=round(bpl / (dom["avg_launch_ms"] * 1e-3) / 1e9, 2))
            dom["frac"] = round(dom["achieved"] / HBM_PEAK_GBS, 5)
        # Measured HBM bytes and VALU issue of the dominant symbol: counted in
        # this run by rocprofv3 child runs of one frame with the same tuning
        # (FETCH_SIZE and WRITE_SIZE scaled by a device copy of known size;
        # SQ counters against the FMA probe's issue rate); the committed
        # counter summaries of this workload only if that fails.
        traffic, traffic_src, valu = None, None, None
        measured = None
        if not args.no_traffic and args.transport == "wavefront":
            try:
                measured = measure_counters(args, {k: getattr(ctx.get_tuning(), k) for k, _ in abi.Tuning._fields_},
                                            dom["kernel"], lambda n: ctx.valu_probe(n)[0])
            except Exception as exc:  # noqa: BLE001 - reported in the line, the committed summary is used
                print(f"# counter runs failed: {exc}", file=sys.stderr, flush=True)
                measured = {"error": str(exc)[:300]}
        if measured and "hbm_bytes_per_dispatch" in measured:
            traffic = measured["hbm_bytes_per_dispatch"]
            traffic_src = (f"measured in this run: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of one frame "
                           f"({measured['dispatches_profiled']} dispatches), x{measured['fetch_scale']} / "
                           f"x{measured['write_scale']} from a 1 GiB device copy")
        if measured and "valu_wave_instructions_per_dispatch" in measured:
            rate = measured["valu_wave_instructions_per_dispatch"] / (dom["avg_launch_ms"] * 1e-3)
            valu = {"issue_frac": round(rate / measured["valu_ceiling_per_s"], 4),
                    "lane_utilisation": measured["lane_utilisation"], "wait_share": measured["wait_share"],
                    "issue_wait_share": measured["issue_wait_share"],
                    "source": "measured in this run: rocprofv3 --pmc SQ_INSTS_VALU ... of one frame, against "
                              "the SQ_INSTS_VALU rate of k_valu_probe"}
        tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
        if traffic is None and os.path.exists(tpath):
            rec = json.load(open(tpath)).get(args.workload, {})
            k = rec.get("kernels", {}).get(dom["kernel"])
            if k:
                traffic = k["hbm_bytes_per_dispatch"]
                traffic_src = f"{rec['source']} (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, commit {rec['commit']})"
        # VALU issue rate of the dominant symbol against the FMA probe's
        # measured ceiling, from the committed rocprofv3 SQ-counter summary.
        vpath = os.path.join(ROOT, "profiles", "valu_issue.json")
        if valu is None and os.path.exists(vpath):
            rec = json.load(open(vpath)).get(args.workload, {})
            k = rec.get("kernels", {}).get(dom["kernel"])
            if k:
                valu = {"issue_frac": k["valu_issue_frac"], "lane_utilisation": k["lane_utilisation"],
                        "source": f"{rec['source']} (rocprofv3 --pmc SQ_INSTS_VALU ..., commit {rec['commit']})"}
        roofline = {"bound": "hbm", "achieved": dom["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": dom["frac"], "traffic": traffic, "traffic_source": traffic_src, "valu": valu,
                    "traffic_frac": (round(traffic / (dom["avg_launch_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)
                                     if traffic else None),
                    "kernel": dom["kernel"], "launches_per_step": dom["launches_per_step"],
                    "avg_launch_ms": dom["avg_launch_ms"], "bytes_per_launch": dom["bytes_per_launch"],
                    "bytes_per_sample": {"kernel": dom["bytes_per_sample"], "path": round(per_sample, 1)},
                    "kernels": entries,
                    "stage_ms": {k: round(v["ms"], 2) for k, v in stages.items() if k != "complete"},
                    "per_sample": {"rays": c.rays / ns, "camera_rays": c.primaryRays / ns,
                                   "shadow_rays": c.shadowRays / ns,
                                   "node_fetches": c.nodeFetches / ns, "tri_tests": c.triFetches / ns,
                                   "camera_node_fetches": c.primaryNodeFetches / ns,
                                   "camera_tri_tests": c.primaryTriFetches / ns,
                                   "shadow_node_fetches": c.shadowNodeFetches / ns,
                                   "shadow_tri_tests": c.shadowTriFetches / ns,
                                   "env_taps": c.envTaps / ns, "cdf_steps": c.cdfSteps / ns},
                    "wavefront_counts": wf, "counter_runs": measured}
        if args.transport == "wavefront":
            roofline["note"] = ("achieved: bytes the traversal kernel's load instructions request (each node "
                                "and triangle once per wave instruction that reads it, as the lanes of one "
                                "instruction reading one record share the request), counted by the kernel's "
                                "own counting instantiation; frac_per_ray_sum in `kernels` charges every ray "
                                "its own fetches instead; traffic_frac is what HBM moved (rocprofv3 counters), "
                                "valu.issue_frac the binding limit")
    # Outside the timed region: the other transport's frame on the same
    # samples (the throughput comparison is at matched error only if the
    # frames agree; both transports evaluate the same per-sample operations,
    # so they are expected to be identical) and the end-to-end time of a
    # host-buffer batch, readback included.
    parity = None
    if not args.no_parity:
        other = RenderSettings(**{**st.__dict__, "transport": abi.TRANSPORT_MEGAKERNEL
                                  if st.transport == abi.TRANSPORT_WAVEFRONT else abi.TRANSPORT_WAVEFRONT})
        obufs = [torch.empty_like(b) for b in bufs]
        ctx.render_device(wl.camera, other, [b.data_ptr() for b in obufs], 0, st.samplesPerPixel, synchronize=True)
        other_ms = ctx.last_render_ms()
        torch.cuda.synchronize()
        diff = (bufs[0] != obufs[0]).view(-1, 4).any(dim=1).float().mean().item()
        parity = {"other_transport": "megakernel" if other.transport == abi.TRANSPORT_MEGAKERNEL else "wavefront",
                  "other_ms": round(other_ms, 2), "speedup_vs_other": round(other_ms / (sum(gpu_ms) / len(gpu_ms)), 3),
                  "beauty_identical": all(bool(torch.equal(a, b)) for a, b in zip(bufs, obufs)),
                  "pixels_differing": diff}
        if st.transport == abi.TRANSPORT_WAVEFRONT:
            # The megakernel with path regeneration and LDS-ring stacks, a
            # stronger single-kernel baseline, on the same samples.
            plus = RenderSettings(**{**st.__dict__, "transport": abi.TRANSPORT_MEGAKERNEL_PLUS})
            ctx.render_device(wl.camera, plus, [b.data_ptr() for b in obufs], 0, st.samplesPerPixel, synchronize=True)
            plus_ms = ctx.last_render_ms()
            torch.cuda.synchronize()
            parity.update(megakernel_plus_ms=round(plus_ms, 2),
                          speedup_vs_megakernel_plus=round(plus_ms / (sum(gpu_ms) / len(gpu_ms)), 3),
                          megakernel_plus_identical=all(bool(torch.equal(a, b)) for a, b in zip(bufs, obufs)))
        del obufs
    e2e_ms = None
    if not args.no_e2e:
        t0 = time.perf_counter()
        host = ctx.render(wl.camera, st)
        e2e_ms = round((time.perf_counter() - t0) * 1e3, 2)
        del host
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(wl, args.cpu_seconds, args.cpu_threads)
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2),
        "gpu_ms_per_step": round(sum(gpu_ms) / len(gpu_ms), 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded procedural scene and dome map)",
        "config": {"workload": workload_desc(wl), "transport": args.transport, "tuning": tuning,
                   "samples_per_step": samples,
                   "bvh": f"BVH4 {info.bvhNodeCount} nodes, depth {info.bvhDepth}; 8-wide {info.bvh8NodeCount} nodes, "
                          f"depth {info.bvh8Depth}",
                   "scene_upload_s": round(upload_s, 2), "finite_output": ok,
                   "end_to_end_ms_with_readback": e2e_ms, "parity": parity},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
