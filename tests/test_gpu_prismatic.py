"""Prismatic slab selection on the device (include/disco_prismatic.h) against
the oracle's restatement of prismaticSlab and the reference's golden slabs:
same atoms, same order, same sides, bit for bit."""
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, golden
from disco_amd import backend as B
from disco_amd import geometry as G
from disco_amd import prismatic as P
from test_oracle_prismatic import REF_CASES, hexagon, random_prismatic_loops

sys.path.insert(0, GOLDEN)
import make_prismatic as MP  # noqa: E402

pytestmark = pytest.mark.gpu


def _dev_lattice(ctx, kind, cells, a0):
    return B.lattice_torch(ctx, kind, cells, cells, cells, a0)


@pytest.mark.parametrize("k", range(len(MP.CASES)))
def test_slab_equals_golden(hip_ctx, k):
    g = golden("prismatic.npz")
    lp, shape, p, kind, a0, cells, spacing, box = MP.case(k)
    x, y, z = _dev_lattice(hip_ctx, kind, cells, a0)
    st = G.native_loop_state(lp)
    atoms, side, info = P.prismatic_slab_torch(hip_ctx, st, spacing, x, y, z, box)
    assert np.array_equal(atoms.cpu().numpy(), g[f"atoms{k}"])
    assert np.array_equal(side.cpu().numpy(), g[f"side{k}"].astype(np.int32))
    gi = g[f"info{k}"]
    assert (info["type"], info["planes"], info["height"], info["spacing"], *info["b"]) == \
        (int(gi[0]), int(gi[1]), gi[2], gi[3], gi[4], gi[5], gi[6])


def test_slab_equals_oracle_on_random_loops(hip_ctx):
    rng = np.random.default_rng(41)
    total = 0
    for lp, code, p, kind, a0, cells, spacing, box in random_prismatic_loops(rng, 80):
        x, y, z = _dev_lattice(hip_ctx, kind, cells, a0)
        st = G.native_loop_state(lp)
        atoms, side, info = P.prismatic_slab_torch(hip_ctx, st, spacing, x, y, z, box)
        hx, hy, hz = (t.cpu().numpy() for t in (x, y, z))
        oa, os_, oi = oracle.prismatic_slab(st, spacing, hx, hy, hz, box)
        assert np.array_equal(atoms.cpu().numpy(), oa), code
        assert np.array_equal(side.cpu().numpy(), os_), code
        assert info["count"] == oa.size and info["planes"] == int(oi[1])
        total += oa.size
    assert total > 2000


@pytest.mark.parametrize("shape", ["circle", "polygon", "rect"])
def test_slab_full_size_4m_atoms(hip_ctx, shape):
    """FCC 100^3 (4,000,000 atoms): a large interstitial loop on (110), 2 planes."""
    a0, cells = 3.615, 100
    x, y, z = _dev_lattice(hip_ctx, "FCC", cells, a0)
    m, n = [-1, -1, 0], [1, 1, 0]
    c, spacing = G.native_snap_to_mid_plane([40.0, -25.0, 33.0], n, a0, cells, cells, cells, "FCC")
    if shape == "circle":
        lp = G.Loop("circle", c, m, n, a0, radius=60.0)
    elif shape == "polygon":
        lp = G.Loop("polygon", c, m, n, a0, vertices=G.native_regular_polygon(7, 70.0, m, n),
                    corner_r=9.0)
    else:
        lp = G.Loop("rect", c, m, n, a0, half_m=80.0, half_p=45.0, corner_r=12.0)
    st = G.native_loop_state(lp)
    box = [cells * a0] * 3
    atoms, side, info = P.prismatic_slab_torch(hip_ctx, st, spacing, x, y, z, box)
    hx, hy, hz = (t.cpu().numpy() for t in (x, y, z))
    oa, os_, oi = oracle.prismatic_slab(st, spacing, hx, hy, hz, box)
    assert info["type"] == 2 and info["planes"] == 2
    assert oa.size > 1000
    assert np.array_equal(atoms.cpu().numpy(), oa)
    assert np.array_equal(side.cpu().numpy(), os_)


def test_slab_edge_cases(hip_ctx):
    import torch
    a0, cells = 2.8553, 10
    x, y, z = _dev_lattice(hip_ctx, "BCC", cells, a0)
    box = [cells * a0] * 3
    c, sp = G.native_snap_to_mid_plane([0, 0, 0], [1, 1, 1], a0, cells, cells, cells, "BCC")
    glide = G.native_loop_state(G.Loop("circle", c, [1, -1, 0], [1, 1, 1], a0, radius=8.0))
    atoms, side, info = P.prismatic_slab_torch(hip_ctx, glide, sp, x, y, z, box)
    assert info["type"] == 0 and info["count"] == 0 and atoms.numel() == 0
    vac = G.native_loop_state(G.Loop("circle", c, [1, 1, 1], [1, 1, 1], a0, radius=8.0))
    full, _, info = P.prismatic_slab_torch(hip_ctx, vac, sp, x, y, z, box)
    assert info["count"] > 10
    with pytest.raises(B.DiscoError) as e:                 # too small an output
        P.prismatic_slab_torch(hip_ctx, vac, sp, x, y, z, box, cap=5)
    assert e.value.code == 5
    again, _, info2 = P.prismatic_slab_torch(hip_ctx, vac, sp, x[:1000], y[:1000], z[:1000], box)
    hx, hy, hz = (t[:1000].cpu().numpy() for t in (x, y, z))
    assert np.array_equal(again.cpu().numpy(), oracle.prismatic_slab(vac, sp, hx, hy, hz, box)[0])
    c2, sp2 = G.native_snap_to_mid_plane([0, 0, 0], [1, 0, 0], a0, cells, cells, cells, "BCC")
    bad = G.native_loop_state(G.Loop("circle", c2, [0.5, 0, 0], [1, 0, 0], a0, radius=5.0))
    with pytest.raises(B.DiscoError) as e:
        P.prismatic_slab_torch(hip_ctx, bad, sp2, x, y, z, box)
    assert e.value.code == 1
    empty = torch.empty(0, dtype=torch.float64, device="cuda")
    a, s, info = P.prismatic_slab_torch(hip_ctx, vac, sp, empty, empty, empty, box)
    assert info["count"] == 0


@pytest.mark.parametrize("case", range(len(REF_CASES)))
def test_reference_unit_test_slabs(hip_ctx, case):
    """The reference test's checks on the device slab: matter type, |b.n| / d
    planes, slab size within 15 % of rho A |b.n|, every atom inside the slab
    bounds on its reported side -- and the slab equal to the oracle's."""
    kind, shape, m, n, rc, vdir, expect = REF_CASES[case]
    a0, N = 3.3024, 24
    x, y, z = B.lattice_torch(hip_ctx, kind, N, N, N, a0)
    c, d = G.native_snap_to_mid_plane([1.1, -0.7, 0.4], n, a0, N, N, N, kind)
    R = 18.0
    if shape == "circle":
        lp = G.Loop("circle", c, m, n, a0, radius=R)
        area = np.pi * R * R
    else:
        lp = G.Loop("polygon", c, m, n, a0, vertices=hexagon(R, m, n, vdir), corner_r=rc,
                    axis=vdir or [])
        area = 1.5 * np.sqrt(3) * R * R - rc * rc * (6 / np.sqrt(3) - np.pi)
    st = G.native_loop_state(lp)
    box = [N * a0] * 3
    atoms, side, info = P.prismatic_slab_torch(hip_ctx, st, d, x, y, z, box)
    atoms, side = atoms.cpu().numpy(), side.cpu().numpy()
    bn = abs(np.dot(st.bmag * np.array(st.m[:]), np.array(st.n[:])))
    assert info["type"] == expect == P.matter_type(st)
    assert info["planes"] == round(bn / d)
    rho = (2 if kind == "BCC" else 4) / a0 ** 3
    nominal = rho * area * bn
    assert abs(atoms.size - nominal) < 0.15 * nominal, (atoms.size, nominal)
    hx, hy, hz = (t.cpu().numpy() for t in (x, y, z))
    s = ((hx[atoms] - c[0]) * st.n[0] + (hy[atoms] - c[1]) * st.n[1] + (hz[atoms] - c[2]) * st.n[2])
    t_lo, t_hi = -np.floor(info["planes"] / 2), np.ceil(info["planes"] / 2)
    assert np.all(s / d >= t_lo - 1e-9) and np.all(s / d < t_hi)
    assert np.array_equal(s > 0, side > 0)
    oa, os_, _ = oracle.prismatic_slab(st, d, hx, hy, hz, box)
    assert np.array_equal(atoms, oa) and np.array_equal(side, os_)
