"""One sample through the whole device pipeline the driver runs
(Code/main.cpp:576-852): the lattice generated on the device, prismatic slabs
selected and applied (vacancy atoms deleted, interstitial copies listed), the
displacement sum over the surviving atoms, the copies appended, final positions
and both output files -- against the oracle's chain on the same inputs:
identical atoms and copies, the sum within 1e-10 A of the FP64 oracle, files
byte-identical to the oracle writers' given the same arrays."""
import numpy as np
import pytest
import torch

import oracle
from disco_amd import backend as B
from disco_amd import geometry as G
from disco_amd import prismatic as P
from disco_amd import writers as W

pytestmark = pytest.mark.gpu


def _host(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("kind", ["BCC", "FCC"])
def test_sample_pipeline(hip_ctx, tmp_path, kind):
    a0 = 2.8553 if kind == "BCC" else 3.615
    N = 12
    box = [N * a0] * 3
    x, y, z = B.lattice_torch(hip_ctx, kind, N, N, N, a0)
    hx, hy, hz = G.lattice(kind, N, N, N, a0)
    for t, h in zip((x, y, z), (hx, hy, hz)):
        assert np.array_equal(_host(t), h)
    # two glide loops, a vacancy loop and an interstitial loop
    if kind == "BCC":
        glide = [([1, 1, 1], [1, -1, 0]), ([1, -1, 1], [1, 1, 0])]
        prism = [([1, 1, 1], [1, 1, 1]), ([-1, -1, -1], [1, 1, 1])]
    else:
        glide = [([1, -1, 0], [1, 1, 1]), ([0, 1, -1], [1, 1, 1])]
        prism = [([1, 1, 0], [1, 1, 0]), ([-1, -1, 0], [1, 1, 0])]
    rng = np.random.default_rng(3 + (kind == "FCC"))
    loops = []
    for t, (m, n) in enumerate(glide + prism):
        c, sp = G.native_snap_to_mid_plane(list(rng.uniform(-0.3, 0.3, 3) * N * a0), n, a0, N, N,
                                           N, kind)
        if t % 2 == 0:
            lp = G.Loop("circle", c, m, n, a0, radius=float(rng.uniform(5, 9)))
        else:
            lp = G.Loop("polygon", c, m, n, a0,
                        vertices=G.native_regular_polygon(6, float(rng.uniform(5, 9)), m, n))
        loops.append((lp, sp))
    states = [(G.native_loop_state(lp), sp) for lp, sp in loops]

    # prismatic step on the device, and the oracle's
    res = P.apply_slabs_torch(hip_ctx, x, y, z, box,
                              [{"state": st, "spacing": sp} for st, sp in states])
    slabs = []
    for st, sp in states:
        if oracle.matter_type(st) == 0:
            slabs.append({"type": 0, "atoms": np.zeros(0, np.int32),
                          "side": np.zeros(0, np.int32)})
            continue
        a, s, info = oracle.prismatic_slab(st, sp, hx, hy, hz, box)
        slabs.append({"type": int(info[0]), "atoms": a, "side": s,
                      "shifts": np.array([info[4:7], -info[4:7]])})
    (ox, oy, oz, _), (oca, ocs) = oracle.prismatic_apply(hx, hy, hz, None, None, slabs)
    assert 0 < oca.size and ox.size < hx.size
    for t, w in zip((res["x"], res["y"], res["z"]), (ox, oy, oz)):
        assert np.array_equal(_host(t), w)
    assert np.array_equal(_host(res["copy_atom"]), oca)
    assert np.array_equal(_host(res["copy_shift"]), ocs)

    # the sum over the surviving atoms: every loop, images +-1, FP64
    data = [lp.line_data(0.01) for lp, _ in loops]
    u = [torch.zeros_like(res["x"]) for _ in range(3)]
    plan = B.Plan(hip_ctx, data, box, 1, B.PREC_FP64)
    plan.run_torch(res["x"], res["y"], res["z"], *u)
    plan.free()
    ou = oracle.line_sum(ox, oy, oz, data, box, 1)
    for t, w in zip(u, ou):
        assert np.abs(_host(t) - w).max() < 1e-10

    # the copies after the sum, then the output step
    outs, _ = P.append_copies_torch(hip_ctx, res["x"], res["y"], res["z"], *u, res["copy_atom"],
                                    res["copy_shift"])
    hu = [_host(t) for t in u]
    want, _ = oracle.append_copies(ox, oy, oz, *hu, None, oca, ocs)
    for t, w in zip(outs, want):
        assert np.array_equal(_host(t), w)
    assert outs[0].numel() == ox.size + oca.size
    name = str(tmp_path / "sample.xyz")
    f = W.write_outputs(hip_ctx, name, "both", *outs, box, "Fe", 55.845)
    torch.cuda.synchronize()
    fin = oracle.final_positions(*want, box)
    for got, w in zip(f, fin):
        assert np.array_equal(_host(got), w)
    Hp = np.diag(box)
    oracle.write_xyz(tmp_path / "o.xyz", *fin, Hp, "Fe")
    oracle.write_lammps(tmp_path / "o.lmp", *fin, Hp, "Fe", 55.845)
    assert open(name, "rb").read() == open(tmp_path / "o.xyz", "rb").read()
    assert open(name[:-3] + "lmp", "rb").read() == open(tmp_path / "o.lmp", "rb").read()
