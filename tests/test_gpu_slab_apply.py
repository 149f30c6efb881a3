"""Applying the prismatic slabs on the device (the driver's bookkeeping,
Code/main.cpp:586-687 and 733-742, as order-keeping compactions) against the
oracle's restatement: same surviving atoms in the same order, same labels,
same copies with the same shifts, same appended displacements -- bit for bit."""
import numpy as np
import pytest
import torch

import oracle
from disco_amd import backend as B
from disco_amd import geometry as G
from disco_amd import prismatic as P
from test_oracle_slab_apply import random_apply_case

pytestmark = pytest.mark.gpu


def _dev(a, dtype=torch.float64):
    return None if a is None else torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def _host(t):
    return None if t is None else t.cpu().numpy()


def _loops_for_device(states, slabs, extra=None):
    out = []
    for (st, spacing), s in zip(states, slabs):
        d = {"state": st, "spacing": spacing, "grain": s["grain"], "region": s["region"]}
        d.update(extra or {})
        out.append(d)
    return out


def _check_against_oracle(ctx, pos, grain, twin, box, slabs, states, rng, extra=None):
    x, y, z = (_dev(a) for a in pos)
    g, tw = _dev(grain, torch.int32), _dev(twin, torch.int32)
    res = P.apply_slabs_torch(ctx, x, y, z, box, _loops_for_device(states, slabs, extra), g, tw)
    (ox, oy, oz, og), (oca, ocs) = oracle.prismatic_apply(*pos, grain, twin, slabs)
    assert np.array_equal(_host(res["x"]), ox)
    assert np.array_equal(_host(res["y"]), oy)
    assert np.array_equal(_host(res["z"]), oz)
    if grain is not None:
        assert np.array_equal(_host(res["grain"]), og)
    assert np.array_equal(_host(res["copy_atom"]), oca)
    assert np.array_equal(_host(res["copy_shift"]), ocs)
    # after the sum: the copies appended with u + shift
    u = [rng.normal(size=ox.size) for _ in range(3)]
    outs, lab = P.append_copies_torch(ctx, res["x"], res["y"], res["z"], *(_dev(a) for a in u),
                                      res["copy_atom"], res["copy_shift"], res["grain"])
    wants, wlab = oracle.append_copies(ox, oy, oz, *u, og, oca, ocs)
    for t, w in zip(outs, wants):
        assert np.array_equal(_host(t), w)
    if grain is not None:
        assert np.array_equal(_host(lab), wlab)
    return res, oca.size


@pytest.mark.parametrize("kind", ["BCC", "FCC"])
@pytest.mark.parametrize("grains,twins", [(True, True), (True, False), (False, True),
                                          (False, False)])
def test_apply_equals_oracle(hip_ctx, kind, grains, twins):
    rng = np.random.default_rng(100 + 4 * (kind == "FCC") + 2 * grains + twins)
    pos, grain, twin, box, slabs, states = random_apply_case(rng, kind, grains=grains,
                                                             twins=twins)
    res, ncopies = _check_against_oracle(hip_ctx, pos, grain, twin, box, slabs, states, rng)
    assert res["x"].numel() < pos[0].size and ncopies > 0
    # the per-loop report: signed slab size after the filters
    for rep, s in zip(res["reports"], slabs):
        assert rep["type"] == s["type"]


def test_apply_frank_interstitial_with_rotated_grain(hip_ctx):
    """An FCC Frank interstitial loop: the copies carry frankCopyShift's vector
    (rotated into the crystal frame and back), from the host library."""
    rng = np.random.default_rng(9)
    a0, N, m, n = 3.615, 12, [-2 / 3] * 3, [1, 1, 1]
    x, y, z = G.lattice("FCC", N, N, N, a0)
    c, d = G.native_snap_to_mid_plane([1.1, -0.7, 0.4], n, a0, N, N, N, "FCC")
    st = G.native_loop_state(G.Loop("circle", c, m, n, a0, radius=12.0))
    box = [N * a0] * 3
    atoms, side, info = oracle.prismatic_slab(st, d, x, y, z, box)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    R = q   # the grain's orientation only enters the shift here
    sh = oracle.copy_shifts(info[4:7], True, np.array(st.n[:]), R, a0)
    slabs = [{"type": int(info[0]), "atoms": atoms, "side": side, "grain": -1, "region": 0,
              "shifts": sh}]
    _, ncopies = _check_against_oracle(hip_ctx, (x, y, z), None, None, box, slabs, [(st, d)],
                                       rng, {"frank": True, "R": R, "a0": a0})
    assert ncopies == atoms.size > 0


def test_apply_full_size_4m_atoms(hip_ctx):
    """FCC 100^3 (4,000,000 atoms), device lattice, 12 loops (vacancy and
    interstitial, some overlapping), three grains."""
    a0, cells = 3.615, 100
    x, y, z = B.lattice_torch(hip_ctx, "FCC", cells, cells, cells, a0)
    hx, hy, hz = (t.cpu().numpy() for t in (x, y, z))
    rng = np.random.default_rng(17)
    grain = (np.floor((hx / (cells * a0) + 0.5) * 3).astype(np.int32) % 3)
    box = [cells * a0] * 3
    slabs, loops = [], []
    for t in range(12):
        m, n = ([1, 1, 0], [1, 1, 0]) if t % 2 else ([-1, -1, 0], [1, 1, 0])
        if t % 3 == 0:
            m, n = ([0, 0, 2], [0, 0, 1]) if t % 2 else ([0, 0, -2], [0, 0, 1])
        c0 = list(rng.uniform(-0.35, 0.35, 3) * cells * a0)
        c, sp = G.native_snap_to_mid_plane(c0, n, a0, cells, cells, cells, "FCC")
        st = G.native_loop_state(G.Loop("circle", c, m, n, a0, radius=float(rng.uniform(20, 70))))
        atoms, side, info = oracle.prismatic_slab(st, sp, hx, hy, hz, box)
        gl = int(rng.integers(-1, 3))
        slabs.append({"type": int(info[0]), "atoms": atoms, "side": side, "grain": gl,
                      "region": 0, "shifts": oracle.copy_shifts(info[4:7], False, n, None, a0)})
        loops.append({"state": st, "spacing": sp, "grain": gl})
    res = P.apply_slabs_torch(hip_ctx, x, y, z, box, loops, _dev(grain, torch.int32))
    (ox, oy, oz, og), (oca, ocs) = oracle.prismatic_apply(hx, hy, hz, grain, None, slabs)
    assert 0 < res["x"].numel() < hx.size and oca.size > 1000
    for got, want in zip((res["x"], res["y"], res["z"], res["grain"]), (ox, oy, oz, og)):
        assert np.array_equal(_host(got), want)
    assert np.array_equal(_host(res["copy_atom"]), oca)
    assert np.array_equal(_host(res["copy_shift"]), ocs)


def test_apply_edge_cases(hip_ctx):
    a0, N = 2.8553, 6
    x, y, z = (_dev(a) for a in G.lattice("BCC", N, N, N, a0))
    box = [N * a0] * 3
    # no loops, and a glide loop: nothing changes, nothing is copied
    res = P.apply_slabs_torch(hip_ctx, x, y, z, box, [])
    assert res["x"] is x and res["copy_atom"].numel() == 0
    glide = G.native_loop_state(G.Loop("circle", [0.1, 0.2, 0.3], [1, -1, 0], [1, 1, 0], a0,
                                       radius=5.0))
    res = P.apply_slabs_torch(hip_ctx, x, y, z, box, [{"state": glide, "spacing": 1.0}])
    assert res["x"] is x and res["reports"] == [None]
    # a loop switched to displacement only
    c, sp = G.native_snap_to_mid_plane([0.0, 0.0, 0.0], [1, 1, 1], a0, N, N, N, "BCC")
    vac = G.native_loop_state(G.Loop("circle", c, [1, 1, 1], [1, 1, 1], a0, radius=5.0))
    res = P.apply_slabs_torch(hip_ctx, x, y, z, box,
                              [{"state": vac, "spacing": sp, "matter": False}])
    assert res["x"] is x
    # a filter that keeps nothing, and one that names atoms out of range
    atoms, side, _ = P.prismatic_slab_torch(hip_ctx, vac, sp, x, y, z, box)
    assert atoms.numel() > 0
    g = torch.zeros(x.numel(), dtype=torch.int32, device="cuda")
    a2, s2 = P.slab_filter_torch(hip_ctx, x.numel(), atoms.clone(), side.clone(), g, 1)
    assert a2.numel() == 0
    bad = atoms.clone()
    bad[0] = x.numel() + 5
    a3, _ = P.slab_filter_torch(hip_ctx, x.numel(), bad, side.clone(), g, 0)
    assert a3.numel() == atoms.numel() - 1
    # every atom removed: nothing survives, every copy is dropped
    removed = torch.ones(x.numel(), dtype=torch.uint8, device="cuda")
    ca = torch.arange(4, dtype=torch.int32, device="cuda")
    cs = torch.ones((4, 3), dtype=torch.float64, device="cuda")
    (xo, _, _, _), (cao, _) = P.drop_atoms_torch(hip_ctx, removed, x, y, z, None, ca, cs)
    assert xo.numel() == 0 and cao.numel() == 0
    # no copies to append
    outs, _ = P.append_copies_torch(hip_ctx, x, y, z, x, y, z,
                                    torch.empty(0, dtype=torch.int32, device="cuda"),
                                    torch.empty((0, 3), dtype=torch.float64, device="cuda"))
    assert torch.equal(outs[0], x)
    # invalid arguments are errors, not launches
    with pytest.raises(B.DiscoError):
        P.drop_atoms_torch(hip_ctx, removed[:0], x[:0], y[:0], z[:0], None, ca, cs)
