"""Applying the prismatic slabs, host side: the Burgers-vector helpers of the
host library (classifyBurgers, frankCopyShift, the copy shifts with the grain
rotation) against the compiled reference's own functions, the Frank-copy
checks of the reference's unit test, and the oracle's restatement of the
driver's bookkeeping (Code/main.cpp:593-687, 733-742) against invariants that
fix it: which atoms survive and in what order, which copies exist, where they
sit and what displacement they carry."""
import itertools

import numpy as np
import pytest

import oracle
from disco_amd import geometry as G
from test_oracle_prismatic import random_prismatic_loops

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference build not present")

A0 = 3.3024
T3 = 2.0 / 3.0


def _burgers_grid():
    comps = [0.0, T3 / 2, -T3 / 2, T3, -T3, 1.0, -1.0, 2.0, -2.0, 0.5, 4 / 3]
    ms = [list(m) for m in itertools.product(comps, repeat=3) if any(m)]
    ns = [[1, 1, 1], [1, -1, 1], [-1, -1, 1], [1, 1, 0], [0, 0, 1], [1, 1, 2], [2, 2, 2.0000001]]
    return ms, ns


def test_classify_burgers_known_answers():
    for f in (G.native_classify_burgers, oracle.classify_burgers):
        assert f("FCC", [1, 1, 0], [1, 1, 1]) == 0
        assert f("FCC", [T3, T3, T3], [1, 1, 1]) == 1
        assert f("FCC", [-T3, -T3, -T3], [2, 2, 2]) == 1
        assert f("FCC", [T3, T3, T3], [1, 1, 0]) == 3          # n not along m
        assert f("FCC", [T3 / 2, T3 / 2, -T3], [1, 1, 1]) == 2  # 1/3 [1 1 -2] in (111)
        assert f("FCC", [T3 / 2, T3 / 2, -T3], [1, 1, 0]) == 3  # not a {111} plane
        assert f("FCC", [1, 0, 0], [1, 0, 0]) == 3              # odd sum: not FCC
        assert f("BCC", [1, 1, 1], [1, 1, 0]) == 0
        assert f("BCC", [2, 0, 0], [1, 0, 0]) == 0
        assert f("BCC", [1, 0, 0], [1, 0, 0]) == 3
        assert f("BCC", [T3, T3, T3], [1, 1, 1]) == 3           # partials are FCC only


@needs_ref
def test_classify_burgers_matches_reference():
    ms, ns = _burgers_grid()
    seen = set()
    for kind in ("BCC", "FCC"):
        for m in ms:
            for n in ns:
                want = oracle.ref_classify_burgers(kind, m, n)
                assert G.native_classify_burgers(kind, m, n) == want, (kind, m, n)
                assert oracle.classify_burgers(kind, m, n) == want, (kind, m, n)
                seen.add(want)
    assert seen == {0, 1, 2, 3}


def _frank_inputs(rng, count):
    dirs = [np.array(d, float) for d in itertools.product([1, -1], repeat=3)]
    for t in range(count):
        d = dirs[t % 8]
        nhat = d / np.sqrt(3.0) if t % 5 else rng.normal(size=3)
        nhat = nhat / np.linalg.norm(nhat)
        a0 = float(rng.uniform(2.5, 4.5)) if t % 3 else A0
        if t % 4 == 0:
            v = rng.normal(size=3)                                   # any vector
        else:
            v = (1 if t % 2 else -1) * (a0 / 2) * T3 * d             # +-b of a Frank loop
        yield v, nhat, a0


@needs_ref
def test_frank_copy_shift_matches_reference():
    rng = np.random.default_rng(5)
    moved = 0
    for v, nhat, a0 in _frank_inputs(rng, 400):
        want = oracle.ref_frank_copy_shift(v, nhat, a0)
        assert np.array_equal(G.native_frank_copy_shift(v, nhat, a0), want), (v, nhat, a0)
        assert np.array_equal(oracle.frank_copy_shift(v, nhat, a0), want), (v, nhat, a0)
        moved += not np.array_equal(want, v)
    assert moved > 200


@needs_ref
def test_copy_shifts_match_reference_rotations():
    """The copy shifts = the reference's rotateToCrystal, frankCopyShift and
    rotateToLab composed (Code/main.cpp:633-650), for rotated and unrotated
    grains; +-b unchanged for a non-Frank loop."""
    rng = np.random.default_rng(6)
    for t in range(200):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        R = None if t % 4 == 0 else q
        d = np.array([1, 1, 1], float) * np.where(rng.random(3) < 0.5, -1, 1)
        nc = d / np.sqrt(3.0)                                  # crystal-frame normal
        n = nc if R is None else R @ nc                        # the loop's lab normal
        b = (A0 / 2) * T3 * (d if t % 2 else -d)
        b = b if R is None else R @ b
        frank = t % 3 != 0
        got = G.native_copy_shifts(b, frank, n, R, A0)
        assert np.array_equal(got, oracle.copy_shifts(b, frank, n, R, A0))
        for sd, sign in enumerate((1.0, -1.0)):
            v = sign * b
            if not frank:
                assert np.array_equal(got[sd], v)
                continue
            if R is None:
                want = oracle.ref_frank_copy_shift(v, n, A0)
            else:
                nh = oracle.ref_rotate(R, False, n)
                vc = oracle.ref_rotate(R, False, v)
                want = oracle.ref_rotate(R, True, oracle.ref_frank_copy_shift(vc, nh, A0))
            assert np.array_equal(got[sd], want), (t, sd)


def _is_fcc_site(p, N):
    q = (np.asarray(p) + 0.5 * A0 * N) / (0.5 * A0)
    r = np.round(q)
    return bool(np.all(np.abs(q - r) < 1e-6) and int(abs(r.sum())) % 2 == 0)


@pytest.mark.parametrize("m,n", [([-T3, -T3, -T3], [1, 1, 1]), ([T3, -T3, T3], [-1, 1, -1]),
                                 ([T3, T3, -T3], [-1, -1, 1])])
def test_frank_copies_reference_unit_test_checks(m, n):
    """Code/tests/test_prismatic.cpp:187-219 (checkFrankCopies, N = 12 cells
    here): every copy's shift is in the habit plane, |shift - v| = a0/sqrt6, and
    the copy lands neither on a site of the layer across the cut nor straight
    over its source."""
    N = 12
    x, y, z = G.lattice("FCC", N, N, N, A0)
    c, d = G.native_snap_to_mid_plane([1.1, -0.7, 0.4], n, A0, N, N, N, "FCC")
    lp = G.Loop("circle", c, m, n, A0, radius=15.0)
    assert G.native_classify_burgers("FCC", m, n) == 1
    st = G.native_loop_state(lp)
    atoms, side, info = oracle.prismatic_slab(st, d, x, y, z, [N * A0] * 3)
    assert int(info[0]) == 2 and int(info[1]) == 1 and atoms.size > 0
    nhat = np.array(st.n[:])
    shifts = G.native_copy_shifts(info[4:7], True, nhat, None, A0)
    for a, s in zip(atoms, side):
        v = s * info[4:7]
        w = shifts[0 if s > 0 else 1]
        assert abs(np.dot(w - v, nhat)) < 1e-9
        assert abs(np.linalg.norm(w - v) - A0 / np.sqrt(6.0)) < 1e-9
        p = np.array([x[a], y[a], z[a]]) + w
        assert not _is_fcc_site(p, N)
        assert not _is_fcc_site(p - v, N)


def random_apply_case(rng, kind="BCC", cells=10, nloops=6, grains=True, twins=True):
    """A lattice, grain and twin labels and several prismatic loops' oracle slabs
    (vacancy and interstitial, overlapping), as the driver would hold them."""
    a0 = 2.8553 if kind == "BCC" else 3.615
    x, y, z = G.lattice(kind, cells, cells, cells, a0)
    n = x.size
    grain = rng.integers(0, 3, n).astype(np.int32) if grains else None
    twin = (rng.random(n) < 0.2).astype(np.int32) if twins else None
    box = [cells * a0] * 3
    loops = [c for c in random_prismatic_loops(rng, 40) if c[3] == kind and c[5] <= cells]
    # a vacancy loop inside an interstitial one on the same plane: copies of
    # deleted atoms, which the renumbering drops
    c0 = list(rng.uniform(-0.3, 0.3, 3) * cells * a0)
    pair = [(G.Loop("circle", c0, [1, 1, 1], [1, 1, 1], a0, radius=0.2 * cells * a0), -1, 0),
            (G.Loop("circle", c0, [-1, -1, -1], [1, 1, 1], a0, radius=0.3 * cells * a0), -1, 0)]
    if kind == "FCC":
        pair = [(G.Loop("circle", c0, [1, 1, 0], [1, 1, 0], a0, radius=0.2 * cells * a0), -1, 0),
                (G.Loop("circle", c0, [-1, -1, 0], [1, 1, 0], a0, radius=0.3 * cells * a0), -1,
                 0)]
    chosen = [(c[0], int(rng.integers(-1, 3)) if grains else -1, int(rng.integers(0, 2)))
              for c in loops[:nloops]]
    slabs, states = [], []
    for lp, lg, reg in pair[:1] + chosen + pair[1:]:
        # the lattice of the case, not the loop's own cell count
        c, spacing = G.native_snap_to_mid_plane(lp.center, lp.n, a0, cells, cells, cells, kind)
        lp.center = c
        st = G.native_loop_state(lp)
        atoms, side, info = oracle.prismatic_slab(st, spacing, x, y, z, box)
        sh = G.native_copy_shifts(info[4:7], False, list(st.n), None, a0)
        slabs.append({"type": int(info[0]), "atoms": atoms, "side": side,
                      "grain": lg if grains else -1, "region": reg, "shifts": sh})
        states.append((st, spacing))
    return (x, y, z), grain, twin, box, slabs, states


def _filtered(s, grain, twin):
    keep = np.ones(len(s["atoms"]), bool)
    if s["grain"] >= 0:
        keep &= grain[s["atoms"]] == s["grain"]
    if twin is not None:
        keep &= twin[s["atoms"]] == s["region"]
    return s["atoms"][keep], s["side"][keep]


@pytest.mark.parametrize("grains,twins", [(True, True), (True, False), (False, False)])
def test_prismatic_apply_invariants(grains, twins):
    rng = np.random.default_rng(11 + 2 * grains + twins)
    (x, y, z), grain, twin, box, slabs, _ = random_apply_case(rng, grains=grains, twins=twins)
    assert {s["type"] for s in slabs} == {1, 2}
    (xo, yo, zo, go), (ca, cs) = oracle.prismatic_apply(x, y, z, grain, twin, slabs)
    n = x.size
    removed = np.zeros(n, bool)
    for s in slabs:
        if s["type"] == 1:
            removed[_filtered(s, grain, twin)[0]] = True
    kept = np.flatnonzero(~removed)                    # survivors in index order
    assert 0 < removed.sum() < n
    assert np.array_equal(xo, x[kept]) and np.array_equal(yo, y[kept])
    assert np.array_equal(zo, z[kept])
    if grain is not None:
        assert np.array_equal(go, grain[kept])
    new_index = np.full(n, -1)
    new_index[kept] = np.arange(kept.size)
    src, shift = [], []
    for s in slabs:                                    # loop order, then slab order
        if s["type"] != 2:
            continue
        a, sd = _filtered(s, grain, twin)
        src.append(a)
        shift.append(np.where((sd > 0)[:, None], s["shifts"][0], s["shifts"][1]))
    src, shift = np.concatenate(src), np.concatenate(shift)
    live = new_index[src] >= 0
    assert (~live).sum() > 0                           # some copies of deleted atoms
    assert np.array_equal(ca, new_index[src[live]])
    assert np.array_equal(cs, shift[live])
    # after the sum: each copy at its source's site with u + shift
    u = [rng.normal(size=kept.size) for _ in range(3)]
    (fx, fy, fz, fux, fuy, fuz), fg = oracle.append_copies(xo, yo, zo, *u, go, ca, cs)
    k = kept.size
    assert np.array_equal(fx[:k], xo) and np.array_equal(fux[:k], u[0])
    assert np.array_equal(fx[k:], xo[ca]) and np.array_equal(fz[k:], zo[ca])
    assert np.array_equal(fux[k:], u[0][ca] + cs[:, 0])
    assert np.array_equal(fuz[k:], u[2][ca] + cs[:, 2])
    if go is not None:
        assert np.array_equal(fg[k:], go[ca])


def test_prismatic_apply_without_loops_changes_nothing():
    x, y, z = G.lattice("BCC", 4, 4, 4, 2.8553)
    (xo, yo, zo, go), (ca, cs) = oracle.prismatic_apply(x, y, z, None, None, [])
    assert np.array_equal(xo, x) and go is None and ca.size == 0
    glide = {"type": 0, "atoms": np.arange(5, dtype=np.int32),
             "side": np.ones(5, np.int32), "grain": -1}
    (xo, _, _, _), (ca, _) = oracle.prismatic_apply(x, y, z, None, None, [glide])
    assert np.array_equal(xo, x) and ca.size == 0
