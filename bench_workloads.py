"""The other workloads of the benchmark (opt-in: ``bench.py --workload c3|c4|c5``).

Each prints one JSON line in bench.py's format.  The default bench line (the
4M-atom line-integral sample) stays in bench.py.

* c3 -- loopons: BCC a0 = 2.8553, 60^3 cells (432,000 atoms), 20 circular loops
        R = 40 A on {110} (2,001 loopons each), images +-1, FP32 pair terms.
* c4 -- batched datagen: 256 BCC samples (20-60 cells, 1-16 loops of 16-128
        segments: circles, rectangles, parallelograms), images +-1, one
        persistent launch per pass replayed as a HIP graph, autotuned shape.
* c5 -- images +-4: FCC 100^3 (4M atoms), 200 loops as the default sample,
        in cubature form (exact near each loop image, a two-ring and then a
        9-point cubature of the cut farther out, far images on grids), against
        the FP64 plain sum.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

FLOP_LINE = 100.0
FLOP_LOOPON = 58.0
FP32_PEAK = 157.3
BF16_PEAK = 2500.0


def _emit(d):
    print(json.dumps(d), flush=True)


def _cpu_line(sample, stride, img=None):
    import oracle
    idx = np.arange(0, sample.n, stride)
    img = sample.img if img is None else img
    t0 = time.perf_counter()
    oracle.line_sum(sample.x[idx], sample.y[idx], sample.z[idx], sample.loops, sample.box, img)
    dt = time.perf_counter() - t0
    ev = idx.size * sample.segments() * (2 * img + 1) ** 3
    return ev / dt, oracle.threads(), f"every {stride}th atom ({idx.size}), images +-{img}, {dt:.1f} s"


def _cpu_reference(kind, sample, stride):
    """The reference's own kroupaCPU (Code/backend_cpu.cpp:28-99, compiled from
    its sources) on every `stride`-th atom at its +-1 images, or None when that
    build is absent: (evals/s, threads, description)."""
    import oracle
    if not oracle.ref_available():
        return None
    idx = np.arange(0, sample.n, stride)
    fn = oracle.ref_kroupa_cpu_line if kind == "line" else oracle.ref_kroupa_cpu_loopon
    t0 = time.perf_counter()
    fn(sample.x[idx], sample.y[idx], sample.z[idx], sample.loops, sample.box)
    dt = time.perf_counter() - t0
    ev = idx.size * sample.segments() * 27
    return ev / dt, oracle.threads(), (f"the reference's kroupaCPU on every {stride}th atom "
                                       f"({idx.size}), images +-1, {dt:.1f} s")


def run_c3(args):
    import torch
    import oracle
    from disco_amd import backend as B
    from disco_amd import geometry as G
    s = G.config_c3()
    dev = torch.device("cuda:0")
    x, y, z = (torch.from_numpy(a).to(dev) for a in (s.x, s.y, s.z))
    u = [torch.zeros_like(x) for _ in range(3)]
    ctx = B.Context(0)
    plan = B.Plan(ctx, s.loops, s.box, s.img, B.PREC_FP32C)
    plan.profile(True)
    evals = s.n * plan.evals_per_atom()
    for _ in range(args.warmup):
        plan.run_torch(x, y, z, *u)
    torch.cuda.synchronize()
    ms = []
    for k in range(args.steps):
        plan.run_torch(x, y, z, *u)
        ms.append(plan.kernel_ms()[0])
        print(f"[c3] step {k + 1}: {ms[-1]:.1f} ms", flush=True)
    # the FP64 mode over every atom: the accuracy check of the timed output and
    # the FP64 rate
    v = [torch.zeros_like(x) for _ in range(3)]
    p64 = B.Plan(ctx, s.loops, s.box, s.img, B.PREC_FP64)
    p64.profile(True)
    p64.run_torch(x, y, z, *v)
    ms64 = p64.kernel_ms()[0]
    err = max(float((a - b).abs().max()) for a, b in zip(u, v))
    print(f"[c3] FP64 mode: {ms64:.1f} ms ({evals / (ms64 * 1e-3):.3e} evals/s); max|du| of the "
          f"FP32 pass vs FP64 (all atoms) {err:.3e} A", flush=True)
    cpu = None
    if not args.no_cpu:
        sub = np.arange(0, s.n, 128)
        t0 = time.perf_counter()
        oracle.loopon_sum(s.x[sub], s.y[sub], s.z[sub], s.loops, s.box, 1)
        dt = time.perf_counter() - t0
        cpu = {"value": sub.size * s.segments() * 27 / dt, "unit": "evals/s",
               "cores": oracle.threads(), "kind": "port",
               "sample": f"every 128th atom ({sub.size}), all 20 loops, +-1; {dt:.1f} s"}
        ref = _cpu_reference("loopon", s, 128)
        if ref is not None:
            cpu = {"value": ref[0], "unit": "evals/s", "cores": ref[1], "kind": "reference",
                   "sample": ref[2] + ", all 20 loops", "port": cpu}
    k = float(np.mean(ms))
    ach = evals * FLOP_LOOPON / (k * 1e-3) / 1e12
    _emit({"metric": "atom·loopon·image evals/s (loopon method)", "value": evals / (k * 1e-3),
           "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": k, "max_abs_err_A": err, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32 (f64 partial-sum flushes)", "data": "synthetic",
           "config": {"workload": "BCC 60^3 (432,000 atoms), 20 loops x 2,001 loopons, images +-1",
                      "evals_per_step": evals, "fp64_mode_ms": ms64,
                      "fp64_mode_evals_per_s": evals / (ms64 * 1e-3)},
           "roofline": {"bound": "valu", "achieved": ach, "peak": FP32_PEAK, "unit": "TFLOP/s",
                        "frac": ach / FP32_PEAK, "traffic": None, "kernel": "loopon_kernel<DISCO_PREC_FP32C>",
                        "kernel_ms": k, "flop_per_eval": FLOP_LOOPON},
           "cpu_baseline": cpu})
    ctx.close()


def run_c4(args):
    import torch
    from disco_amd import backend as B
    from disco_amd import geometry as G
    ctx = B.Context(0)
    G._nlib()
    # sample generation: loop contours by the C++ producer, lattices on the device
    t0 = time.perf_counter()
    samples = G.config_c4_specs()
    t_host = time.perf_counter() - t0
    dev = torch.device("cuda:0")
    ntot = sum(s.n for s in samples)
    x, y, z = (torch.empty(ntot, dtype=torch.float64, device=dev) for _ in range(3))
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    off = 0
    for s in samples:
        N = s.meta["cells"]
        B.lattice_torch(ctx, "BCC", N, N, N, s.meta["a0"],
                        out=[t[off:off + s.n] for t in (x, y, z)])
        off += s.n
    e1.record()
    torch.cuda.synchronize()
    t_lat = e0.elapsed_time(e1)
    print(f"[c4] {len(samples)} samples, {ntot} atoms: contours {t_host:.3f} s on the host "
          f"(C++ producer), lattices {t_lat:.2f} ms on the device", flush=True)
    u = [torch.zeros_like(x) for _ in range(3)]
    bt = B.Batch(ctx, samples, 1, B.PREC_FP32C)
    evals = bt.evals()
    default_shape = bt.tuning
    bt.run_torch(x, y, z, *u)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    bt.run_torch(x, y, z, *u)
    e1.record()
    torch.cuda.synchronize()
    ms_default = e0.elapsed_time(e1)
    t0 = time.perf_counter()
    best, best_ms = bt.autotune(x, y, z, *u, stride=8)
    tune_s = time.perf_counter() - t0
    print(f"[c4] default {default_shape}: {ms_default:.1f} ms; tuned {best} "
          f"(est. {best_ms:.1f} ms, {tune_s:.1f} s)", flush=True)
    for _ in range(args.warmup):
        bt.run_torch(x, y, z, *u)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(args.steps):
        bt.run_torch(x, y, z, *u)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    # accuracy: FP64 sums of every 16th sample
    errs = []
    host = lambda k: [t[bt.offsets[k]:bt.offsets[k] + samples[k].n].cpu().numpy()
                      for t in (x, y, z)]
    for k in range(0, len(samples), 16):
        s = samples[k]
        off = bt.offsets[k]
        ref = B.line_sum(ctx, *host(k), s.loops, s.box, 1, B.PREC_FP64)
        got = [t[off:off + s.n].cpu().numpy() for t in u]
        errs.append(max(float(np.abs(a - b).max()) for a, b in zip(got, ref)))
    cpu = None
    if not args.no_cpu:
        hx, hy, hz = host(0)
        s0 = G.Sample("C4", hx, hy, hz, samples[0].loops, samples[0].box, 1)
        v, c, desc = _cpu_line(s0, 64)
        cpu = {"value": v, "unit": "evals/s", "cores": c, "kind": "port",
               "sample": "sample 0 of the batch, " + desc}
        ref = _cpu_reference("line", s0, 64)
        if ref is not None:
            cpu = {"value": ref[0], "unit": "evals/s", "cores": ref[1], "kind": "reference",
                   "sample": "sample 0 of the batch, " + ref[2], "port": cpu}
    ach = evals * FLOP_LINE / (ms * 1e-3) / 1e12
    _emit({"metric": "atom·segment·image evals/s (batched samples)", "value": evals / (ms * 1e-3),
           "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms, "max_abs_err_A": max(errs), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32 (f64 image sums, f64 near cut)",
           "data": "synthetic",
           "config": {"workload": "256 BCC samples (20-60 cells, 1-16 loops, 16-128 segments), +-1",
                      "atoms": int(sum(s.n for s in samples)), "evals_per_step": evals,
                      "tuning": best, "default_shape_ms": ms_default, "autotune_s": tune_s,
                      "datagen": {"contours_host_s": t_host, "lattices_device_ms": t_lat}},
           "roofline": {"bound": "valu", "achieved": ach, "peak": FP32_PEAK, "unit": "TFLOP/s",
                        "frac": ach / FP32_PEAK, "traffic": None,
                        "kernel": "batch_line_kernel (far + near passes)", "kernel_ms": ms},
           "cpu_baseline": cpu})
    ctx.close()


def run_c5(args):
    """The 200-loop sample at images +-4 (729 per loop) in cubature form: exact
    line integral within 3 cut radii of each loop image (FP32, FP64 near slab;
    5 radii for the hexagons, which the two-ring cubature does not represent),
    the two-ring cubature from 3 to 5 radii of the circles, the 9-point cubature
    of the cut beyond, the +-1 images far from the atoms of an octant on that
    octant's L/32 grid, the far shell's images 2 on an L/12 grid and 3..4 on an
    L/6 grid (6-node interpolation stencils).
    Error against the FP64 plain sum on every 8th atom (every atom with
    --c5-all); the plain GPU sum timed on a contiguous eighth of the atoms and
    scaled; the oracle timed on a few atoms."""
    import torch
    from disco_amd import backend as B
    from disco_amd import cubature as C
    from disco_amd import geometry as G
    img, k_exact, k_band, near_tol = 4, 5.0, 3.0, 3e-7
    t0 = time.time()
    c = G.config_c2(seed=20245, nloops=200)
    print(f"[c5] workload {c.n} atoms, {len(c.loops)} loops, {c.segments()} segments "
          f"({time.time() - t0:.1f} s)", flush=True)
    dev = torch.device("cuda:0")
    x, y, z = (torch.from_numpy(a).to(dev) for a in (c.x, c.y, c.z))
    ctx = B.Context(0)
    t0 = time.perf_counter()
    cs = C.CubatureSum(ctx, x, y, z, c.loops, c.box, img, k_exact, near_tol=near_tol,
                       k_band=k_band)
    t_setup = time.perf_counter() - t0
    u = [torch.zeros_like(x) for _ in range(3)]
    for _ in range(max(1, args.warmup)):
        cs.run(x, y, z, *u)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ms = []
    for k in range(args.steps):
        ev[0].record(stream)
        cs.run(x, y, z, *u)
        ev[1].record(stream)
        torch.cuda.synchronize()
        ms.append(ev[0].elapsed_time(ev[1]))
        print(f"[c5] step {k + 1}: {ms[-1]:.1f} ms", flush=True)
    t_step = float(np.mean(ms))
    cs.near.profile(True)
    cs.run(x, y, z, *u)
    far_ms, near_ms = cs.near.kernel_ms()
    torch.cuda.synchronize()
    # accuracy against the FP64 plain sum
    stride = 1 if getattr(args, "c5_all", False) else 8
    idx = torch.arange(0, c.n, stride, device=dev)
    xs, ys, zs = x[idx].contiguous(), y[idx].contiguous(), z[idx].contiguous()
    v = [torch.zeros_like(xs) for _ in range(3)]
    p64 = B.Plan(ctx, c.loops, c.box, img, B.PREC_FP64)
    t0 = time.perf_counter()
    p64.run_torch(xs, ys, zs, *v)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    p64.free()
    err = max(float((a[idx] - b).abs().max()) for a, b in zip(u, v))
    which = "all atoms" if stride == 1 else f"every {stride}th"
    print(f"[c5] max|du| vs the FP64 plain sum on {idx.numel()} atoms ({which}): "
          f"{err:.3e} A ({t64:.1f} s of FP64)", flush=True)
    del v
    # the plain GPU sum (FP32C) on a contiguous eighth, scaled to the sample
    m = c.n // 8
    w = [torch.zeros(m, dtype=torch.float64, device=dev) for _ in range(3)]
    plain = B.Plan(ctx, c.loops, c.box, img, B.PREC_FP32C)
    plain.profile(True)
    plain.run_torch(x[:m], y[:m], z[:m], *w)
    pf, pn = plain.kernel_ms()
    plain.free()
    t_plain = (pf + pn) * c.n / m
    print(f"[c5] plain +-4 FP32C sum: {t_plain / 1e3:.2f} s per sample (scaled from the first "
          f"eighth); cubature form {t_step:.1f} ms ({t_plain / t_step:.0f}x)", flush=True)
    cpu = None
    if not args.no_cpu:
        import oracle
        sub = np.arange(0, c.n, c.n // 256)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            oracle.line_sum(c.x[sub], c.y[sub], c.z[sub], c.loops, c.box, img)
            times.append(time.perf_counter() - t0)
        dt = float(np.median(times))
        ev_cpu = sub.size * c.segments() * (2 * img + 1) ** 3
        cpu = {"value": ev_cpu / dt, "unit": "evals/s", "cores": oracle.threads(), "kind": "port",
               "sample": f"{sub.size} atoms spread over the sample, all 200 loops, images +-4, "
                         f"plain sum; median of 3 runs, {dt:.1f} s each",
               "s_per_sample_implied": c.n * c.segments() * (2 * img + 1) ** 3 / (ev_cpu / dt)}
    plain_evals = c.n * c.segments() * (2 * img + 1) ** 3
    _emit({"metric": "ms per 4M-atom sample, images +-4 (cubature form)",
           "value": t_step, "unit": "ms", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": t_step, "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "max_abs_err_A": err,
           "max_abs_err_atoms": int(idx.numel()),
           "dtype": "f32 line and cubature, f64 image sums and near slab; f32 loopon "
                    "cubature on grids",
           "data": "synthetic",
           "config": {"workload": "FCC 100^3 (4M atoms), 200 loops as the 4M-atom sample, "
                                  "images +-4", "k_exact": k_exact, "k_band": k_band,
                      "near_tol_A": near_tol, "rule_loops": cs.eligible, "band_loops": cs.band,
                      "images_on_grid": cs.skipped,
                      "setup_s": t_setup, "plain_gpu_ms": t_plain,
                      "speedup_vs_plain": t_plain / t_step, "plain_evals": plain_evals,
                      "plain_evals_per_s_equivalent": plain_evals / (t_step * 1e-3)},
           "roofline": {"bound": "valu", "achieved": None, "peak": FP32_PEAK, "unit": "TFLOP/s",
                        "frac": None, "traffic": None,
                        "kernel": "line_kernel<PASS_FAR, LOOP_GLIDE, true> (cubature far pass)",
                        "kernel_ms": far_ms, "near_kernel_ms": near_ms,
                        "grids_ms": t_step - far_ms - near_ms},
           "cpu_baseline": cpu})
    cs.free()
    ctx.close()


def run_lattice(args):
    """The 4M-atom sample's lattice (FCC a0 = 3.615, 100^3 cells) generated in HBM:
    24 B written per atom, HBM-bound; against the reference's generateLattice on the
    host (the compiled reference when present, else the restatement) plus the copy
    of its positions to the device."""
    import torch
    import oracle
    from disco_amd import backend as B
    from disco_amd import geometry as G
    kind, nc, a0 = "FCC", 100, 3.615
    ctx = B.Context(0)
    n = B.lattice_atoms(kind, nc, nc, nc)
    out = [torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(3)]
    for _ in range(max(1, args.warmup)):
        B.lattice_torch(ctx, kind, nc, nc, nc, a0, out=out)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(args.steps):
        B.lattice_torch(ctx, kind, nc, nc, nc, a0, out=out)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    rx, ry, rz = G.lattice(kind, nc, nc, nc, a0)
    exact = all(np.array_equal(t.cpu().numpy(), r) for t, r in zip(out, (rx, ry, rz)))
    print(f"[lattice] {n} sites in {ms:.3f} ms; bit-identical to the host restatement: {exact}",
          flush=True)
    # host path: generateLattice on the host cores, then the H2D copy of the positions
    cpu = None
    if not args.no_cpu:
        try:
            t0 = time.perf_counter()
            hx, hy, hz = oracle.ref_lattice(kind, nc, nc, nc, a0)
            dt, k = time.perf_counter() - t0, "reference"
        except Exception:
            t0 = time.perf_counter()
            hx, hy, hz = G.lattice(kind, nc, nc, nc, a0)
            dt, k = time.perf_counter() - t0, "port"
        pin = [torch.from_numpy(a).pin_memory() for a in (hx, hy, hz)]
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        dev = [p.to("cuda", non_blocking=True) for p in pin]
        torch.cuda.synchronize()
        h2d = time.perf_counter() - t1
        cpu = {"value": n / dt, "unit": "sites/s", "cores": 1, "kind": k,
               "sample": f"the whole lattice ({n} sites), one host thread: {dt * 1e3:.1f} ms; "
                         f"then the pinned H2D copy of 96 MB: {h2d * 1e3:.1f} ms",
               "h2d_ms": h2d * 1e3}
        del dev
    gbs = 24.0 * n / (ms * 1e-3) / 1e9
    _emit({"metric": "lattice sites/s generated in HBM", "value": n / (ms * 1e-3),
           "unit": "sites/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": "FCC a0=3.615, 100^3 cells (4,000,000 sites), the 4M-atom "
                                  "sample's lattice", "bit_identical": exact},
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s",
                        "frac": gbs / 8000.0, "traffic": 24.0 * n,
                        "kernel": "lattice_kernel<4>", "kernel_ms": ms},
           "cpu_baseline": cpu})
    ctx.close()


def run_writers(args):
    """The 4M-atom sample's output stage (Code/main.cpp:797-852): final positions
    (x + u wrapped into the box, shifted by +L/2) and both files, extended XYZ and
    the LAMMPS data file, from the device displacements of the FP32C sum; against
    the reference's host writers (the oracle's iostream restatement, one host
    thread, as the reference writes) on a bounded atom sample."""
    import shutil
    import tempfile
    import torch
    import oracle
    from disco_amd import backend as B
    from disco_amd import geometry as G
    from disco_amd import writers as W
    s = G.config_c2()
    ctx = B.Context(0)
    dev = torch.device("cuda:0")
    x, y, z = (torch.from_numpy(a).to(dev) for a in (s.x, s.y, s.z))
    u = [torch.zeros_like(x) for _ in range(3)]
    plan = B.Plan(ctx, s.loops, s.box, s.img, B.PREC_FP32C)
    plan.run_torch(x, y, z, *u)
    torch.cuda.synchronize()
    plan.free()
    box = np.asarray(s.box, dtype=np.float64)
    Hp = np.diag(box)
    n = s.n
    tmp = tempfile.mkdtemp(prefix="disco_out_")
    name = os.path.join(tmp, "sample.xyz")
    try:
        # device kernels alone: final positions, then the text of both formats in HBM
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        f = W.final_positions_torch(ctx, x, y, z, *u, box)
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(args.steps):
            W.final_positions_torch(ctx, x, y, z, *u, box, out=f)
        ev[1].record()
        torch.cuda.synchronize()
        fp_ms = ev[0].elapsed_time(ev[1]) / args.steps
        W.text_profile(ctx, True)
        _, frame = W.lammps_box(Hp)
        kern = {}
        for fmt in ("xyz", "lammps"):
            t = W.format_text_torch(ctx, fmt, *f, element="Cu", frame=frame, host=False)
            ms = []
            for _ in range(args.steps):
                W.format_text_torch(ctx, fmt, *f, element="Cu", frame=frame, host=False)
                ms.append(W.text_kernel_ms(ctx))
            kern[fmt] = {"bytes": int(t.numel()),
                         "count_ms": float(np.mean([m[0] for m in ms])),
                         "scan_ms": float(np.mean([m[1] for m in ms])),
                         "write_ms": float(np.mean([m[2] for m in ms]))}
            del t
        W.text_profile(ctx, False)
        print(f"[writers] final positions {fp_ms:.3f} ms; text kernels (count/scan/write, ms): "
              + "; ".join(f"{k} {v['count_ms']:.3f}/{v['scan_ms']:.3f}/{v['write_ms']:.3f} "
                          f"({v['bytes'] / 1e6:.1f} MB)" for k, v in kern.items()), flush=True)

        # whole output step, files included
        for _ in range(max(1, args.warmup)):
            W.write_outputs(ctx, name, "both", x, y, z, *u, box, "Cu", 63.546)
        torch.cuda.synchronize()
        walls = []
        for k in range(args.steps):
            t0 = time.perf_counter()
            W.write_outputs(ctx, name, "both", x, y, z, *u, box, "Cu", 63.546)
            torch.cuda.synchronize()
            walls.append(time.perf_counter() - t0)
            print(f"[writers] step {k + 1}/{args.steps}: {walls[-1] * 1e3:.1f} ms", flush=True)
        wall = float(np.mean(walls))
        nx, nl = os.path.getsize(name), os.path.getsize(name[:-3] + "lmp")

        # parity on this workload: the first 200,000 atoms, against the oracle
        fh = [t.cpu().numpy() for t in f]
        uh = [t.cpu().numpy() for t in u]
        m = 200_000
        want = oracle.final_positions(s.x[:m], s.y[:m], s.z[:m], *(a[:m] for a in uh), box)
        exact_pos = all(np.array_equal(a[:m], b) for a, b in zip(fh, want))
        oracle.write_xyz(os.path.join(tmp, "o.xyz"), *want, Hp, "Cu")
        with open(name, "rb") as g, open(os.path.join(tmp, "o.xyz"), "rb") as o:
            g.readline()
            o.readline()
            ref_bytes = o.read()                      # Lattice line + the first m atom lines
            exact_txt = g.read(len(ref_bytes)) == ref_bytes
        cpu = None
        if not args.no_cpu:
            ms_ = n // 8
            t0 = time.perf_counter()
            fo = oracle.final_positions(s.x[:ms_], s.y[:ms_], s.z[:ms_], *(a[:ms_] for a in uh),
                                        box)
            oracle.write_xyz(os.path.join(tmp, "c.xyz"), *fo, Hp, "Cu")
            oracle.write_lammps(os.path.join(tmp, "c.lmp"), *fo, Hp, "Cu", 63.546)
            dt = time.perf_counter() - t0
            cpu = {"value": ms_ / dt, "unit": "atoms/s", "cores": 1, "kind": "port",
                   "sample": f"the first {ms_} atoms (one eighth), final positions + both files "
                             f"through iostreams on one host thread: {dt:.2f} s",
                   "seconds_per_sample_scaled": dt * n / ms_}
            print(f"[writers] host writers: {dt:.2f} s for {ms_} atoms "
                  f"(x8 = {dt * 8:.1f} s per sample)", flush=True)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    wk = kern["xyz"]
    gbs = (24.0 * n + wk["bytes"]) / (wk["write_ms"] * 1e-3) / 1e9
    _emit({"metric": "atoms/s written (final positions + XYZ + LAMMPS files)",
           "value": n / wall, "unit": "atoms/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": wall * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64 (text: fixed, 8 decimals)",
           "data": "synthetic",
           "config": {"workload": "FCC a0=3.615 100^3 (4,000,000 atoms) + its FP32C "
                                  "displacements: final positions, .xyz and .lmp files",
                      "xyz_bytes": nx, "lmp_bytes": nl, "final_positions_ms": fp_ms,
                      "text_kernels": kern, "bit_exact_positions_first_200k": exact_pos,
                      "byte_exact_xyz_first_200k": exact_txt,
                      "files_per_s_GBps": (nx + nl) / wall / 1e9},
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s",
                        "frac": gbs / 8000.0, "traffic": None,
                        "kernel": "text_write_kernel<DISCO_TEXT_XYZ>",
                        "kernel_ms": wk["write_ms"],
                        "bytes_per_launch": 24.0 * n + wk["bytes"]},
           "cpu_baseline": cpu})
    ctx.close()


def prismatic_loops(a0, cells, count=20, seed=7):
    """The prismatic-slab workload's loops: FCC {110} interstitial / vacancy loops
    (b = a0/2<110> along n: 2 planes), circles, rounded heptagons and rounded
    rectangles of 20-60 A at random snapped centres."""
    from disco_amd import geometry as G
    rng = np.random.default_rng(seed)
    normals = [[1, 1, 0], [1, -1, 0], [0, 1, 1], [0, 1, -1], [1, 0, 1], [1, 0, -1]]
    out = []
    for t in range(count):
        n = normals[int(rng.integers(6))]
        m = n if t % 2 == 0 else [-v for v in n]
        c, sp = G.native_snap_to_mid_plane(list(rng.uniform(-0.4, 0.4, 3) * cells * a0), n, a0,
                                           cells, cells, cells, "FCC")
        R = float(rng.uniform(20, 60))
        if t % 3 == 0:
            lp = G.Loop("circle", c, m, n, a0, radius=R)
        elif t % 3 == 1:
            lp = G.Loop("polygon", c, m, n, a0, vertices=G.native_regular_polygon(7, R, m, n),
                        corner_r=0.2 * R)
        else:
            lp = G.Loop("rect", c, m, n, a0, half_m=R, half_p=0.6 * R, corner_r=0.25 * R)
        out.append((G.native_loop_state(lp), sp))
    return out


def run_prismatic(args):
    """Prismatic slab selection (Code/prismatic.cpp:30-74) for 20 loops over the
    4M-atom FCC sample: per loop one flag pass over the atoms (24 B read + 1 B
    written per atom), a scan and an ordered scatter; against the oracle's
    restatement of the reference's host scan (one thread, as the reference)."""
    import torch
    import oracle
    from disco_amd import backend as B
    from disco_amd import prismatic as P
    a0, cells = 3.615, 100
    ctx = B.Context(0)
    x, y, z = B.lattice_torch(ctx, "FCC", cells, cells, cells, a0)
    n = x.numel()
    box = [cells * a0] * 3
    loops = prismatic_loops(a0, cells)
    for st, sp in loops[:2]:
        P.prismatic_slab_torch(ctx, st, sp, x, y, z, box)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    per_loop = []
    counts = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        for st, sp in loops:
            ev[0].record()
            a, sd, info = P.prismatic_slab_torch(ctx, st, sp, x, y, z, box)
            ev[1].record()
            torch.cuda.synchronize()
            per_loop.append(ev[0].elapsed_time(ev[1]))
            if k == 0:
                counts.append(info["count"])
    wall = (time.perf_counter() - t0) / args.steps
    dev_ms = float(np.mean(per_loop))
    print(f"[prismatic] {len(loops)} loops over {n} atoms: {wall * 1e3:.2f} ms per sample "
          f"(device {dev_ms:.3f} ms per loop); slab atoms {sum(counts)}", flush=True)
    # parity on the workload: every loop against the oracle
    hx, hy, hz = (t.cpu().numpy() for t in (x, y, z))
    exact = True
    cpu = None
    t_cpu = 0.0
    for (st, sp), cnt in zip(loops, counts):
        a, sd, info = P.prismatic_slab_torch(ctx, st, sp, x, y, z, box)
        t1 = time.perf_counter()
        oa, os_, oi = oracle.prismatic_slab(st, sp, hx, hy, hz, box)
        t_cpu += time.perf_counter() - t1
        exact &= bool(np.array_equal(a.cpu().numpy(), oa) and np.array_equal(sd.cpu().numpy(), os_))
    print(f"[prismatic] identical to the oracle on all loops: {exact}; host scan {t_cpu:.2f} s "
          f"for {len(loops)} loops", flush=True)
    # the whole prismatic step: slabs, vacancy deletion and renumbering, copy
    # list (Code/main.cpp:586-687), then the copies appended after the sum
    # (:733-742), against the oracle's restatement of the same bookkeeping
    dloops = [{"state": st, "spacing": sp} for st, sp in loops]
    zero = torch.zeros_like(x)
    apply_ms = []
    for k in range(args.steps + 1):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        res = P.apply_slabs_torch(ctx, x, y, z, box, dloops)
        nk = res["x"].numel()
        outs, _ = P.append_copies_torch(ctx, res["x"], res["y"], res["z"], zero[:nk], zero[:nk],
                                        zero[:nk], res["copy_atom"], res["copy_shift"])
        torch.cuda.synchronize()
        if k:
            apply_ms.append((time.perf_counter() - t1) * 1e3)
    slabs = []
    for st, sp in loops:
        oa, os_, oi = oracle.prismatic_slab(st, sp, hx, hy, hz, box)
        slabs.append({"type": int(oi[0]), "atoms": oa, "side": os_, "grain": -1,
                      "shifts": np.array([oi[4:7], -oi[4:7]])})
    t1 = time.perf_counter()
    (ox, oy, oz, _), (oca, ocs) = oracle.prismatic_apply(hx, hy, hz, None, None, slabs)
    zk = np.zeros(ox.size)
    wants, _ = oracle.append_copies(ox, oy, oz, zk, zk, zk, None, oca, ocs)
    t_apply_cpu = time.perf_counter() - t1
    apply_exact = all(np.array_equal(o.cpu().numpy(), w) for o, w in zip(outs, wants))
    apply_dev = float(np.mean(apply_ms))
    print(f"[prismatic] whole step on the device (20 slabs + deletion + copies + append): "
          f"{apply_dev:.2f} ms; {hx.size} -> {outs[0].numel()} atoms "
          f"({hx.size - nk} removed, {oca.size} added); identical to the oracle: {apply_exact}; "
          f"oracle bookkeeping alone {t_apply_cpu * 1e3:.1f} ms", flush=True)
    if not args.no_cpu:
        cpu = {"value": len(loops) * n / t_cpu, "unit": "atom-loop tests/s", "cores": 1,
               "kind": "port",
               "sample": f"all {len(loops)} loops over all {n} atoms, the oracle's restatement of "
                         f"prismaticSlab on one host thread: {t_cpu:.2f} s",
               "ms_per_sample": t_cpu * 1e3}
    bytes_per_launch = 26.0 * n
    gbs = bytes_per_launch / (dev_ms * 1e-3) / 1e9
    _emit({"metric": "atom-loop slab tests/s (prismatic slab selection)",
           "value": len(loops) * n / wall, "unit": "atom-loop tests/s", "n_gpus": 1,
           "steps": args.steps, "warmup": 1, "ms_per_step": wall * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "FCC a0=3.615 100^3 (4,000,000 atoms), 20 prismatic {110} "
                                  "loops (circles, rounded heptagons, rounded rectangles, "
                                  "R 20-60 A), 2 planes each",
                      "slab_atoms": int(sum(counts)), "identical_to_oracle": exact,
                      "device_ms_per_loop": dev_ms,
                      "whole_step_ms": apply_dev, "atoms_after": int(outs[0].numel()),
                      "whole_step_identical_to_oracle": bool(apply_exact),
                      "oracle_bookkeeping_ms": t_apply_cpu * 1e3},
           "roofline": {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s",
                        "frac": gbs / 8000.0, "traffic": None,
                        "kernel": "one slab call: slab_flag_kernel + slab_scan_kernel + "
                                  "slab_scatter_kernel (HIP events around the call)",
                        "kernel_ms": dev_ms, "bytes_per_launch": bytes_per_launch},
           "cpu_baseline": cpu})
    ctx.close()


def run_apply(args):
    """The far-shell MLP alone: apply_fused_kernel over the 4M-atom sample (FCC
    100^3) for 200 loops -- 8e8 (atom, loop) pairs x 69,888 MFMA FLOP (layers
    2-4 on v_mfma_f32_32x32x16_bf16), with the weights of a fresh network (the
    throughput does not depend on their values).  Its parity with the forward
    GEMMs and the FP64 restatement is tested; no host baseline is timed here (the
    workload that uses it, c5, times the plain sum beside it)."""
    import torch
    from disco_amd import backend as B
    from disco_amd import surrogate as S
    ctx = B.Context(0)
    x, y, z = B.lattice_torch(ctx, "FCC", 100, 100, 100, 3.615)
    box = [361.5] * 3
    infos = [i for _, i in S.random_fcc_loops(np.random.default_rng(20245), 200)]
    sur = S.Surrogate(ctx, seed=3)
    u = [torch.zeros_like(x) for _ in range(3)]
    for _ in range(max(1, args.warmup)):
        sur.apply_torch(x, y, z, *u, infos, box)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ms = []
    for _ in range(args.steps):
        ev[0].record()
        sur.apply_torch(x, y, z, *u, infos, box)
        ev[1].record()
        torch.cuda.synchronize()
        ms.append(ev[0].elapsed_time(ev[1]))
    t = float(np.mean(ms))
    pairs = x.numel() * len(infos)
    tflops = pairs * 69888.0 / (t * 1e-3) / 1e12
    print(f"[apply] {pairs:.3e} (atom, loop) pairs in {t:.1f} ms: {tflops:.1f} TFLOP/s on MFMA",
          flush=True)
    _emit({"metric": "(atom, loop) pairs/s through the far-shell MLP", "value": pairs / (t * 1e-3),
           "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": t, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "bf16 (fp32 accumulate; layer 1 fp32)", "data": "synthetic",
           "config": {"workload": "FCC a0=3.615 100^3 (4,000,000 atoms) x 200 loops, MLP "
                                  "14 -> 128 -> 128 -> 128 -> 3 (random weights)"},
           "roofline": {"bound": "mfma", "achieved": tflops, "peak": 2500.0, "unit": "TFLOP/s",
                        "frac": tflops / 2500.0, "traffic": None,
                        "kernel": "apply_fused_kernel (HIP events around the call)",
                        "kernel_ms": t, "flop_per_pair": 69888},
           "cpu_baseline": None})
    sur.free()
    ctx.close()


def run_datagen(args):
    """Batched datagen over several batches of the 256-sample workload (seeds
    20244, 20245, ...) with the host's work hidden behind the device's
    (datagen.OverlappedBatches): while batch k runs on the device, the host draws
    batch k+1's loops, builds their contours with the C++ producer and packs its
    plan, and batch k+1's lattices are queued behind batch k.  Reported: the wall
    time per batch in the steady state against the device time per batch, and
    the host time the overlap hides."""
    from disco_amd import backend as B
    from disco_amd import datagen as D
    from disco_amd import geometry as G
    nb = max(2, getattr(args, "batches", 4))
    ctx = B.Context(0)
    G._nlib()
    pipe = D.OverlappedBatches(ctx, lambda k: G.config_c4_specs(seed=20244 + k), nb)
    t = pipe.run()
    for k in range(nb):
        print(f"[datagen] batch {k}: {t.atoms[k]} atoms, wall {t.wall_ms[k]:.1f} ms, device "
              f"{t.device_ms[k]:.1f} ms, host work of the next batch {t.host_ms[k]:.1f} ms",
              flush=True)
    # steady state: batches 1 .. nb-1 (batch 0's wall includes the first graph capture)
    w = float(np.mean(t.wall_ms[1:]))
    d = float(np.mean(t.device_ms[1:]))
    hh = float(np.mean(t.host_ms[:-1]))
    _emit({"metric": "ms per 256-sample batch, host datagen overlapped with the device sum",
           "value": w, "unit": "ms", "n_gpus": 1, "steps": nb - 1, "warmup": 1,
           "ms_per_step": w, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 (f64 image sums, f64 near cut)", "data": "synthetic",
           "config": {"workload": f"{nb} batches of 256 BCC samples (20-60 cells, 1-16 loops of "
                                  "16-128 segments), images +-1",
                      "device_ms_per_batch": d, "wall_over_device": w / d,
                      "host_ms_per_batch_hidden": hh, "atoms_per_batch": t.atoms,
                      "tuning": pipe.tuning, "walls_ms": t.wall_ms, "device_ms": t.device_ms},
           "roofline": None, "cpu_baseline": None})
    ctx.close()


def run(args):
    if not hasattr(args, "train_loops"):
        args.train_loops = 2000
    if not hasattr(args, "epochs"):
        args.epochs = 30
    {"c3": run_c3, "c4": run_c4, "c5": run_c5, "lattice": run_lattice,
     "writers": run_writers, "prismatic": run_prismatic, "apply": run_apply,
     "datagen": run_datagen}[args.workload](args)
