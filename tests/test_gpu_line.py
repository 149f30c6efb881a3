"""GPU parity of the line-integral sum (Code/barnett.hpp:44-107 summed as
Code/backend_cpu.cpp:38-63) through the C ABI, against the FP64 oracle and the
reference's own golden vectors.  Tolerances: FP64 1e-10 A; FP32C 1e-5 A (the
accuracy bar of the FP32 path)."""
import numpy as np
import pytest

import oracle
from conftest import golden
from loop_kinds import BOX, atoms_near, kind_loops
from disco_amd import backend as B
from disco_amd import geometry as G

pytestmark = pytest.mark.gpu
TOL64 = 1e-10
TOL32 = 1e-5


def _near_subsample(s, per_loop=120, rand=600, loops=6, seed=1):
    """Atoms within 12 A of some loop planes (inside R + 12) plus random atoms."""
    rng = np.random.default_rng(seed)
    out = [rng.choice(s.n, rand, replace=False)]
    for lp in s.loops[:loops]:
        c = np.array([lp.cx, lp.cy, lp.cz])
        n = np.array(lp.phys[6:9])
        R = np.max(np.hypot(np.hypot(lp.vx - c[0], lp.vy - c[1]), lp.vz - c[2]))
        h = (s.x - c[0]) * n[0] + (s.y - c[1]) * n[1] + (s.z - c[2]) * n[2]
        r2 = (s.x - c[0]) ** 2 + (s.y - c[1]) ** 2 + (s.z - c[2]) ** 2 - h * h
        cand = np.nonzero((np.abs(h) < 12) & (r2 < (R + 12) ** 2))[0]
        out.append(rng.choice(cand, min(per_loop, cand.size), replace=False))
    return np.unique(np.concatenate(out))


def test_golden_kroupa_cpu_c1(hip_ctx):
    """The reference's kroupaCPU output (4,096 C1 atoms, +-1 images)."""
    g = golden("cpu_line_c1.npz")
    loops = []
    for l in range(len(g["voff"]) - 1):
        a, b = g["voff"][l], g["voff"][l + 1]
        loops.append(B.LoopData(phys=tuple(g["phys"][10 * l:10 * l + 10]), line=True,
                                vx=g["vx"][a:b], vy=g["vy"][a:b], vz=g["vz"][a:b],
                                cx=g["c"][3 * l], cy=g["c"][3 * l + 1], cz=g["c"][3 * l + 2]))
    u64 = np.stack(B.line_sum(hip_ctx, g["x"], g["y"], g["z"], loops, g["box"], 1, B.PREC_FP64))
    u32 = np.stack(B.line_sum(hip_ctx, g["x"], g["y"], g["z"], loops, g["box"], 1, B.PREC_FP32C))
    e64, e32 = np.abs(u64 - g["u"]).max(), np.abs(u32 - g["u"]).max()
    print(f"C1 golden: fp64 {e64:.2e}  fp32c {e32:.2e}")
    assert e64 < TOL64
    assert e32 < TOL32


def test_golden_barnett_vectors(hip_ctx):
    """Single-loop, single-image fields of barnettLoop<double> (img = 0)."""
    g = golden("barnett_f64.npz")
    worst64 = worst32 = 0.0
    for l in range(len(g["voff"]) - 1):
        V = g["V"][g["voff"][l]:g["voff"][l + 1]]
        p = g["p"][g["poff"][l]:g["poff"][l + 1]]
        u = g["u"][g["poff"][l]:g["poff"][l + 1]]
        b, c = g["b"][l], g["c"][l]
        nb = np.linalg.norm(b)
        lp = B.LoopData(phys=(nb, nb, nb, *(b / nb), 0, 0, 1, 0), line=True, vx=V[:, 0],
                        vy=V[:, 1], vz=V[:, 2], cx=c[0], cy=c[1], cz=c[2])
        box = (1e4, 1e4, 1e4)
        u64 = np.stack(B.line_sum(hip_ctx, p[:, 0], p[:, 1], p[:, 2], [lp], box, 0, B.PREC_FP64), 1)
        u32 = np.stack(B.line_sum(hip_ctx, p[:, 0], p[:, 1], p[:, 2], [lp], box, 0, B.PREC_FP32C), 1)
        worst64 = max(worst64, np.abs(u64 - u).max())
        worst32 = max(worst32, np.abs(u32 - u).max())
    print(f"barnett vectors: fp64 {worst64:.2e}  fp32c {worst32:.2e}")
    assert worst64 < TOL64
    assert worst32 < TOL32


@pytest.mark.parametrize("prec,tol", [(B.PREC_FP64, TOL64), (B.PREC_FP32C, TOL32)])
def test_c2_near_subsample_vs_oracle(hip_ctx, prec, tol):
    """C2 geometry (50 loops, +-2 images) on atoms next to loop cuts."""
    s = G.config_c2()
    idx = _near_subsample(s)
    ref = np.stack(oracle.line_sum(s.x[idx], s.y[idx], s.z[idx], s.loops, s.box, s.img))
    u = np.stack(B.line_sum(hip_ctx, s.x[idx], s.y[idx], s.z[idx], s.loops, s.box, s.img, prec))
    err = np.abs(u - ref).max()
    print(f"C2 near subsample ({idx.size} atoms) prec {prec}: max|du| {err:.3e}")
    assert err < tol


def test_c1_full_fp32c_vs_fp64(hip_ctx):
    """All 128,000 C1 atoms: FP32C against the GPU FP64 sum (pinned above)."""
    s = G.config_c1()
    u64 = np.stack(B.line_sum(hip_ctx, s.x, s.y, s.z, s.loops, s.box, 1, B.PREC_FP64))
    u32 = np.stack(B.line_sum(hip_ctx, s.x, s.y, s.z, s.loops, s.box, 1, B.PREC_FP32C))
    err = np.abs(u32 - u64).max()
    print(f"C1 full: fp32c vs fp64 max {err:.3e} rms {np.sqrt(((u32 - u64) ** 2).mean()):.3e}")
    assert err < TOL32


def test_c2_full_size_properties(hip_ctx):
    """All 4,000,000 atoms of the 4M-atom sample (50 loops, +-2): FP32C within the
    tolerance of the FP64 sum on every atom, and additive over loops (the two
    halves of the loop set sum to the whole)."""
    import torch
    s = G.config_c2()
    dev = torch.device("cuda:0")
    x, y, z = (torch.from_numpy(a).to(dev) for a in (s.x, s.y, s.z))

    def run(loops, prec):
        u = [torch.zeros_like(x) for _ in range(3)]
        plan = B.Plan(hip_ctx, loops, s.box, s.img, prec)
        plan.run_torch(x, y, z, *u)
        torch.cuda.synchronize()
        plan.free()
        return torch.stack(u)

    u = run(s.loops, B.PREC_FP32C)
    assert torch.isfinite(u).all()
    u64 = run(s.loops, B.PREC_FP64)
    d = (u - u64).abs()
    err = d.max().item()
    worst = int(d.max(0).values.argmax().item())
    del d, u64
    half = len(s.loops) // 2
    parts = run(s.loops[:half], B.PREC_FP32C) + run(s.loops[half:], B.PREC_FP32C)
    add = (parts - u).abs().max().item()
    print(f"C2 full: fp32c vs fp64 over all {s.n} atoms {err:.3e} (atom {worst}); "
          f"additivity {add:.3e}")
    assert err < TOL32
    assert add < 1e-9


def test_plan_device_path_matches_host_path(hip_ctx):
    import torch
    s = G.config_c1()
    idx = np.arange(0, s.n, 13)
    x, y, z = s.x[idx], s.y[idx], s.z[idx]
    host = np.stack(B.line_sum(hip_ctx, x, y, z, s.loops, s.box, 1, B.PREC_FP32C))
    dev = [torch.tensor(a, device="cuda") for a in (x, y, z)]
    u = [torch.zeros_like(dev[0]) for _ in range(3)]
    plan = B.Plan(hip_ctx, s.loops, s.box, 1, B.PREC_FP32C)
    plan.run_torch(*dev, *u)
    torch.cuda.synchronize()
    got = np.stack([t.cpu().numpy() for t in u])
    assert np.array_equal(got, host)
    plan.run_torch(*dev, *u, accumulate=True)       # u += sum
    torch.cuda.synchronize()
    got2 = np.stack([t.cpu().numpy() for t in u])
    assert np.allclose(got2, 2 * host, rtol=0, atol=1e-14)
    assert plan.evals_per_atom() == 128 * 27
    plan.free()


def test_host_sum_accumulates_and_handles_empty(hip_ctx):
    s = G.config_c1()
    idx = np.arange(0, s.n, 1000)
    x, y, z = s.x[idx], s.y[idx], s.z[idx]
    base = [np.full(idx.size, 0.5), np.full(idx.size, -0.25), np.full(idx.size, 1.0)]
    u = B.line_sum(hip_ctx, x, y, z, s.loops, s.box, 1, B.PREC_FP64,
                   u=[a.copy() for a in base])
    ref = np.stack(oracle.line_sum(x, y, z, s.loops, s.box, 1))
    assert np.abs(np.stack(u) - np.stack(base) - ref).max() < TOL64
    # no loops: u unchanged; no atoms: nothing to do
    u0 = B.line_sum(hip_ctx, x, y, z, [], s.box, 1, B.PREC_FP64, u=[a.copy() for a in base])
    assert all(np.array_equal(a, b) for a, b in zip(u0, base))
    e = B.line_sum(hip_ctx, x[:0], y[:0], z[:0], s.loops, s.box, 1, B.PREC_FP32C)
    assert all(a.size == 0 for a in e)


def test_degenerate_and_duplicate_vertices(hip_ctx):
    """Repeated vertices (zero-length segments) add nothing, as in the reference;
    a contour collapsed to a segment has zero field."""
    s = G.config_c1()
    idx = np.arange(0, s.n, 777)
    x, y, z = s.x[idx], s.y[idx], s.z[idx]
    lp = s.loops[0]
    dup = B.LoopData(phys=lp.phys, line=True, vx=np.repeat(lp.vx, 2), vy=np.repeat(lp.vy, 2),
                     vz=np.repeat(lp.vz, 2), cx=lp.cx, cy=lp.cy, cz=lp.cz)
    a = np.stack(B.line_sum(hip_ctx, x, y, z, [lp], s.box, 1, B.PREC_FP64))
    b = np.stack(B.line_sum(hip_ctx, x, y, z, [dup], s.box, 1, B.PREC_FP64))
    assert np.abs(a - b).max() < 1e-13
    flat = B.LoopData(phys=lp.phys, line=True, vx=lp.vx[:2], vy=lp.vy[:2], vz=lp.vz[:2],
                      cx=lp.cx, cy=lp.cy, cz=lp.cz)
    c = np.stack(B.line_sum(hip_ctx, x, y, z, [flat], s.box, 1, B.PREC_FP64))
    assert np.abs(c).max() == 0.0


def test_invalid_arguments_raise(hip_ctx):
    s = G.config_c1()
    x = s.x[:10]
    with pytest.raises(B.DiscoError):
        B.line_sum(hip_ctx, x, x, x, s.loops, s.box, 1, prec=7)
    with pytest.raises(B.DiscoError):
        B.line_sum(hip_ctx, x, x, x, s.loops, s.box, 99, B.PREC_FP64)
    bad = B.LoopData(phys=s.loops[0].phys, line=True, vx=np.array([0., 10, np.nan, 0]),
                     vy=np.array([0., 0, 10, 10]), vz=np.array([0., 0, 0, 5]))
    with pytest.raises(B.DiscoError):
        B.line_sum(hip_ctx, x, x, x, [bad], s.box, 1, B.PREC_FP64)


@pytest.mark.parametrize("kind", ["glide", "prism", "mixed", "bent", "all"])
def test_loop_kinds_vs_oracle(hip_ctx, kind):
    """Each FP32 far-pass form (glide, prismatic and mixed b summed in the loop
    frame; a loop off its plane summed in global axes) and a plan holding all of
    them (one far kernel per form, accumulated), against the FP64 oracle, +-1
    images, random atoms and atoms near the loops."""
    rng = np.random.default_rng({"glide": 1, "prism": 2, "mixed": 3, "bent": 4, "all": 5}[kind])
    if kind == "all":
        loops = sum((kind_loops(k, rng, 2) for k in ("glide", "prism", "mixed", "bent")), [])
        rng.shuffle(loops)
    else:
        loops = kind_loops(kind, rng)
    x, y, z = atoms_near(loops, rng)
    ref = np.stack(oracle.line_sum(x, y, z, loops, BOX, 1))
    u64 = np.stack(B.line_sum(hip_ctx, x, y, z, loops, BOX, 1, B.PREC_FP64))
    u32 = np.stack(B.line_sum(hip_ctx, x, y, z, loops, BOX, 1, B.PREC_FP32C))
    e64, e32 = np.abs(u64 - ref).max(), np.abs(u32 - ref).max()
    print(f"{kind}: fp64 {e64:.2e}  fp32c {e32:.2e}  |u|max {np.abs(ref).max():.2f}")
    assert e64 < TOL64
    assert e32 < TOL32


@pytest.mark.parametrize("prec", [B.PREC_FP64, B.PREC_FP32C])
def test_per_loop_times_are_measured(hip_ctx, prec):
    """With loop times asked for, the host-buffer sum launches and times each loop
    on its own (Code/backend_cuda.cu:643-700): every real loop gets a positive
    time, a degenerate one 0, the times add up to at most the kernel time, and the
    sum equals the one-launch sum up to the order of the FP64 additions."""
    s = G.config_c1()
    idx = np.arange(0, s.n, 37)
    x, y, z = s.x[idx], s.y[idx], s.z[idx]
    lp = s.loops[0]
    dead = B.LoopData(phys=lp.phys, line=True, vx=lp.vx[:2].copy(), vy=lp.vy[:2].copy(),
                      vz=lp.vz[:2].copy(), cx=lp.cx, cy=lp.cy, cz=lp.cz)
    loops = [s.loops[0], dead, s.loops[1]]
    t = B.BackendTiming()
    timed = np.stack(B.line_sum(hip_ctx, x, y, z, loops, s.box, 1, prec, timing=t))
    one = np.stack(B.line_sum(hip_ctx, x, y, z, loops, s.box, 1, prec))
    assert t.loop_sec[0] > 0 and t.loop_sec[2] > 0 and t.loop_sec[1] == 0.0
    assert sum(t.loop_sec) <= t.kernels * 1.05 + 1e-3
    assert np.abs(timed - one).max() <= 1e-12 * max(1.0, np.abs(one).max())
