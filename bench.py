"""Benchmark of the displacement hot path on one MI355X.

Workload (one step = one pass of the line-integral sum over one sample):
  FCC a0 = 3.615, 100^3 cells = 4,000,000 atoms; 50 glide loops (radius
  U(10, 60) A, 60 % circles of 48 segments, 40 % hexagons; {111} planes,
  a/2<110> in-plane Burgers vectors; seed 20241); periodic images +-2 (125);
  FP32 pair terms with FP64 image sums and FP64 near the cut (DISCO_PREC_FP32C).
Atoms are resident in HBM (FP64 SoA) before the timed region; the loop
geometry is uploaded once (a plan).

Prints one JSON line: evals/s (atom x segment x image evaluations per second),
ms per 4M-atom sample, the error of the timed output against the FP64 sum, the
roofline of the dominant kernel (HIP events around it, live) and the CPU
baseline (the reference's own kroupaCPU, compiled from its sources, and the FP64
oracle, on the host cores, on bounded atom samples).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("atom·segment·image evals/s and ms per 4M-atom sample at "
          "max|Δu|≤1e-5 Å")
# Algorithmic FLOPs per atom x segment x image of the line integral: 93
# add/sub/mul + 4 div + 1 sqrt + 1 log + 1 atan2, each counted as 1.
FLOP_PER_EVAL = 100.0
FP32_VECTOR_PEAK_TFLOPS = 157.3   # MI355X published FP32 vector peak (packed FP32 FMA)
FP32_UNPACKED_PEAK_TFLOPS = 78.6  # the same without packed FP32 (one FMA per lane per clock)
# HBM bytes per atom of the far pass, measured with rocprofv3 PMC counters
# (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, each scaled by a plain copy of
# known size, `--workload copy`): 194.0e6 B per dispatch over 4,000,000 atoms.
FAR_PASS_HBM_BYTES_PER_ATOM = 48.50
# cubature form of the sample (reported beside the plain sum): exact within 3 cut
# radii of a loop image, the two-ring cubature from 3 to 5 radii (loops it
# represents; the others stay exact there), the 9-point cubature beyond; +-1
# images on an L/32 grid per octant of the atoms (those far from the octant's
# atoms; interpolation error bound 3e-7 A each), the far shell on an L/12 grid
# (shells 3 and beyond on L/6), all interpolated by 6-node stencils
K_EXACT, K_BAND, NEAR_TOL = 5.0, 3.0, 3e-7
TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this kernel on this workload: "
                  "profiles/line_kernel_far__fcc4m_50loops_img2__pmc_traffic.md")


def host_cpu():
    """CPU model, physical cores and the CPUs this process may run on (the
    host's /proc/cpuinfo and the scheduler affinity)."""
    model, cores = None, set()
    try:
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"cpu_model": model, "host_physical_cores": len(cores) or None,
            "host_logical_cpus": os.cpu_count(), "usable_cpus": usable}


def _time_runs(fn, runs):
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), times


def cpu_baseline(sample, stride, runs=3, ref_stride=64):
    """The CPU path timed on the host cores, each figure the median of `runs`
    runs on a bounded atom sample of the same workload:
      * the reference's own kroupaCPU (Code/backend_cpu.cpp:28-99, compiled from
        its sources with its OpenMP flags), when that build is present: every
        `ref_stride`-th atom, all loops, the +-1 images it sums (its image range
        is fixed);
      * the FP64 oracle (the same loop structure, generalised to +-img) on every
        `stride`-th atom at the workload's +-img ("port").
    `value` is the reference's rate when it was timed, else the port's."""
    import oracle
    segs = sample.segments()
    idx = np.arange(0, sample.n, stride)
    x, y, z = sample.x[idx], sample.y[idx], sample.z[idx]
    dt, times = _time_runs(lambda: oracle.line_sum(x, y, z, sample.loops, sample.box,
                                                   sample.img), runs)
    port = {"value": idx.size * segs * (2 * sample.img + 1) ** 3 / dt, "unit": "evals/s",
            "cores": oracle.threads(), "kind": "port",
            "sample": f"every {stride}th atom ({idx.size} atoms) of the same workload, all "
                      f"{len(sample.loops)} loops, images +-{sample.img}; median of {runs} "
                      f"runs, {dt:.1f} s each", "seconds": dt, "runs": times}
    if oracle.ref_available():
        ridx = np.arange(0, sample.n, ref_stride)
        rx, ry, rz = sample.x[ridx], sample.y[ridx], sample.z[ridx]
        rdt, rtimes = _time_runs(lambda: oracle.ref_kroupa_cpu_line(rx, ry, rz, sample.loops,
                                                                     sample.box), runs)
        out = {"value": ridx.size * segs * 27 / rdt, "unit": "evals/s",
               "cores": oracle.threads(), "kind": "reference",
               "sample": f"the reference's kroupaCPU on every {ref_stride}th atom "
                         f"({ridx.size} atoms) of the same workload, all {len(sample.loops)} "
                         f"loops, images +-1 (its fixed range); median of {runs} runs, "
                         f"{rdt:.1f} s each", "seconds": rdt, "runs": rtimes, "port": port}
    else:
        out = dict(port)
    out["threads"] = oracle.threads()
    out.update(host_cpu())
    return out


def copy_reference(args):
    """y.copy_(x) of 256 MiB, 3 dispatches: reads and writes 262,144 KiB each, the
    known size that scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    x = torch.randn(32 * 1024 * 1024, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(3):
        y.copy_(x)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / 3
    nbytes = 2 * x.numel() * 8
    print(json.dumps({"metric": "device copy GB/s (counter calibration)", "value": nbytes / ms / 1e6,
                      "unit": "GB/s", "n_gpus": 1, "steps": 3, "warmup": 0, "ms_per_step": ms,
                      "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                      "dtype": "f64", "data": "synthetic",
                      "config": {"workload": "y.copy_(x), 256 MiB read + 256 MiB written"}}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-stride", type=int, default=256)
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-e2e", action="store_true")
    ap.add_argument("--no-cubature", action="store_true")
    ap.add_argument("--c5-all", action="store_true",
                    help="c5: check the error on every atom (about 50 s of FP64) instead of "
                         "every 8th")
    ap.add_argument("--workload", default="c2",
                    choices=["c2", "c3", "c4", "c5", "copy", "lattice", "writers", "prismatic",
                             "apply", "datagen"],
                    help="c2 (default): the 4M-atom line-integral sample; c3: loopons; "
                         "c4: 256 batched samples; c5: images +-4 in cubature form; copy: a plain "
                         "256 MiB device copy (x3), the size reference for HBM byte counters; "
                         "lattice: the 4M-atom lattice generated on the device; writers: the "
                         "4M-atom sample's final positions and .xyz / .lmp files; prismatic: "
                         "slab selection of 20 prismatic loops over the 4M atoms; apply: the "
                         "far-shell MLP over the 4M atoms and 200 loops; datagen: several "
                         "256-sample batches with the host producer overlapped")
    ap.add_argument("--batches", type=int, default=4, help="datagen: batches")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): atoms shard "
              "across GPUs with no exchange, and this build measures one GPU's share.",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "copy":
        copy_reference(args)
        return
    if args.workload != "c2":
        import bench_workloads
        bench_workloads.run(args)
        return

    import torch
    from disco_amd import backend as B
    from disco_amd import geometry as G

    t0 = time.time()
    s = G.config_c2()
    print(f"[bench] workload built: {s.n} atoms, {len(s.loops)} loops, {s.segments()} segments, "
          f"images +-{s.img} ({time.time() - t0:.1f} s)", flush=True)
    dev = torch.device("cuda:0")
    x, y, z = (torch.from_numpy(a).to(dev) for a in (s.x, s.y, s.z))
    ux, uy, uz = (torch.zeros(s.n, dtype=torch.float64, device=dev) for _ in range(3))
    ctx = B.Context(0)
    plan = B.Plan(ctx, s.loops, s.box, s.img, B.PREC_FP32C)
    plan.profile(True)
    evals = s.n * plan.evals_per_atom()
    torch.cuda.synchronize()

    for w in range(args.warmup):
        plan.run_torch(x, y, z, ux, uy, uz)
        torch.cuda.synchronize()
        print(f"[bench] warmup {w + 1}/{args.warmup} done", flush=True)

    # Two plans of the same sample take the steps in turn: step k + 1 is queued
    # before the host waits for step k's kernel times, so the device never idles
    # between steps while the host wakes up.
    plans = (plan, B.Plan(ctx, s.loops, s.box, s.img, B.PREC_FP32C))
    plans[1].profile(True)
    plans[1].run_torch(x, y, z, ux, uy, uz)
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    far_ms, near_ms = [], []
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    ev0.record(stream)
    if args.steps > 0:
        plans[0].run_torch(x, y, z, ux, uy, uz)
    for k in range(args.steps):
        if k + 1 < args.steps:
            plans[(k + 1) % 2].run_torch(x, y, z, ux, uy, uz)
        a, b = plans[k % 2].kernel_ms()          # waits for step k's kernels
        far_ms.append(a)
        near_ms.append(b)
        print(f"[bench] step {k + 1}/{args.steps}: far {a:.1f} ms, near {b:.1f} ms", flush=True)
    ev1.record(stream)
    plans[1].free()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    total_ms = ev0.elapsed_time(ev1)
    ms_step = total_ms / args.steps
    value = evals / (ms_step * 1e-3)

    # accuracy of the timed output against the GPU FP64 sum, over every atom
    err = None
    fp64 = None
    v = None
    if not args.no_check:
        v = [torch.zeros_like(x) for _ in range(3)]
        p64 = B.Plan(ctx, s.loops, s.box, s.img, B.PREC_FP64)
        p64.profile(True)
        p64.run_torch(x, y, z, *v)
        ms64 = p64.kernel_ms()[0]
        err = max(float((a - b).abs().max()) for a, b in zip((ux, uy, uz), v))
        fp64 = {"atoms": s.n, "ms": ms64, "evals_per_s": evals / (ms64 * 1e-3)}
        p64.free()
        print(f"[bench] max|du| vs FP64 over all {s.n} atoms: {err:.3e} A; FP64 mode: "
              f"{ms64:.1f} ms ({fp64['evals_per_s']:.3e} evals/s)", flush=True)

    # the same sample in cubature form (exact within K_BAND loop radii of a loop
    # image, the two-ring cubature to K_EXACT, the 9-point cubature beyond, far
    # images on grids): ms per sample and its error against the FP64 plain sum
    # over every atom
    cub = None
    if not args.no_cubature:
        from disco_amd import cubature as C
        cs = C.CubatureSum(ctx, x, y, z, s.loops, s.box, s.img, K_EXACT, near_tol=NEAR_TOL,
                           k_band=K_BAND)
        w = [torch.zeros_like(x) for _ in range(3)]
        cs.run(x, y, z, *w)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            cs.run(x, y, z, *w)
        e1.record(stream)
        torch.cuda.synchronize()
        cms = e0.elapsed_time(e1) / args.steps
        cerr = max(float((a - b).abs().max()) for a, b in zip(w, v)) if v is not None else None
        cub = {"ms_per_sample": cms, "max_abs_err_A": cerr, "err_atoms": s.n if v is not None else None,
               "speedup_vs_plain": ms_step / cms, "k_exact": K_EXACT, "k_band": K_BAND,
               "near_tol_A": NEAR_TOL, "rule_loops": cs.eligible, "band_loops": cs.band,
               "images_on_grid": cs.skipped,
               "form": "exact line integral within k_band R of each loop image (FP32, FP64 "
                       "near slab; to k_exact R for loops the two-ring cubature does not "
                       "represent), the two-ring cubature (25 points) to k_exact R, the 9-point "
                       "cubature of the cut beyond; +-1 images far from the atoms of an octant "
                       "on that octant's L/32 grid, the far shell on an L/12 grid (6-node "
                       "stencils)"}
        cs.free()
        del w
        print(f"[bench] cubature form: {cms:.1f} ms per sample ({ms_step / cms:.1f}x the plain "
              f"sum), max|du| vs FP64 over all atoms {cerr:.3e} A", flush=True)
    del v

    # kroupaGPU end to end: host arrays in, plan + upload + kernels + download +
    # accumulation into the caller's u out (Code/backend_cuda.cu:568-722 times the
    # same phases)
    e2e = None
    if not args.no_e2e:
        runs = []
        hu = [np.zeros(s.n) for _ in range(3)]
        for k in range(3):
            for a in hu:
                a[:] = 0.0
            t = B.BackendTiming()
            t0 = time.perf_counter()
            B.line_sum(ctx, s.x, s.y, s.z, s.loops, s.box, s.img, B.PREC_FP32C, u=hu, timing=t)
            runs.append({"wall_ms": (time.perf_counter() - t0) * 1e3, "upload_ms": t.upload * 1e3,
                         "kernels_ms": t.kernels * 1e3, "download_ms": t.download * 1e3})
            print(f"[bench] end to end {k + 1}/3: {runs[-1]['wall_ms']:.1f} ms (upload "
                  f"{runs[-1]['upload_ms']:.1f}, kernels {runs[-1]['kernels_ms']:.1f}, download "
                  f"{runs[-1]['download_ms']:.1f})", flush=True)
        med = sorted(runs, key=lambda r: r["wall_ms"])[1]
        e2e = dict(med, runs=runs,
                   max_abs_diff_vs_device_run_A=max(float(np.abs(a - b.cpu().numpy()).max())
                                                    for a, b in zip(hu, (ux, uy, uz))))

    far = float(np.mean(far_ms))
    achieved = evals * FLOP_PER_EVAL / (far * 1e-3) / 1e12
    roofline = {"bound": "valu", "achieved": achieved, "peak": FP32_VECTOR_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP32_VECTOR_PEAK_TFLOPS,
                "traffic": FAR_PASS_HBM_BYTES_PER_ATOM * s.n, "traffic_unit": "bytes",
                "traffic_source": TRAFFIC_SOURCE,
                "peak_unpacked": FP32_UNPACKED_PEAK_TFLOPS,
                "frac_unpacked": achieved / FP32_UNPACKED_PEAK_TFLOPS,
                "kernel": "line_kernel<PASS_FAR, LOOP_GLIDE>", "kernel_ms": far,
                "near_kernel_ms": float(np.mean(near_ms)),
                "flop_per_eval": FLOP_PER_EVAL, "evals_per_launch": evals}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(s, args.cpu_stride, args.cpu_runs)
        extra = (f" (port at +-{s.img}: {cpu['port']['value']:.3e})" if "port" in cpu else "")
        print(f"[bench] cpu baseline ({cpu['kind']}): {cpu['value']:.3e} evals/s on "
              f"{cpu['cores']} threads{extra}", flush=True)
    out = {"metric": METRIC, "value": value, "unit": "evals/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
           "ms_per_sample": ms_step, "max_abs_err_A": err, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 image sums, f64 near cut)",
           "data": "synthetic",
           "config": {"workload": "FCC a0=3.615 100^3 cells (4,000,000 atoms), 50 loops "
                                  "(circles 48 seg / hexagons), images +-2, line integral",
                      "atoms": s.n, "loops": len(s.loops), "segments": s.segments(),
                      "images": (2 * s.img + 1) ** 3, "precision": "FP32C", "seed": 20241,
                      "evals_per_step": evals},
           "e2e_ms": e2e["wall_ms"] if e2e else None, "e2e": e2e, "max_abs_err_atoms": s.n,
           "cubature": cub,
           "roofline": roofline, "cpu_baseline": cpu, "fp64_mode": fp64, "wall_s": wall}
    print(json.dumps(out), flush=True)
    plan.free()
    ctx.close()


if __name__ == "__main__":
    main()
