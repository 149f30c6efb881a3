"""Pins of the output-stage oracle (oracle/disco_writers.cpp) and the host
parts of the device writers, without a GPU.

The arithmetic is checked against exact rational arithmetic of the reference's
expression shapes as its host build fuses them (a*b + c*d + e*f ->
fma(e, f, fma(a, b, c*d)); a - b*c -> fma(-b, c, a)), and the text against
Python's own correctly rounded '%.8f' (an implementation independent of the C
library's)."""
from fractions import Fraction

import numpy as np
import pytest

import oracle
from disco_amd import writers as W


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def dot3(a, b):
    return fma(a[2], b[2], fma(a[0], b[0], a[1] * b[1]))


def special_values(rng, n):
    """Exact ties at the 8th decimal (odd multiples of 2^-9), their negatives,
    signed zeros, tiny values that round to zero, and ordinary coordinates."""
    ties = rng.integers(0, 5_000_000, n // 4) * 2 + 1
    v = np.concatenate([ties / 512.0, -(ties[: n // 8] / 512.0),
                        [0.0, -0.0, -1e-12, 1e-12, 4.9e-9, -4.9e-9, 5e-9, -5e-9, 1.5e-8,
                         2.5e-8, 9999999999.99, -9999999999.999999, 1e-300, -5e-324],
                        rng.uniform(-400, 400, n // 2),
                        rng.uniform(-1, 1, n // 8) * 10.0 ** rng.integers(-9, 10, n // 8)])
    return v


def test_fixed8_is_correctly_rounded_half_even():
    rng = np.random.default_rng(5)
    for v in special_values(rng, 4000):
        assert oracle.fixed8(v) == "%.8f" % v, v
    # the exact-tie cases round to even
    assert oracle.fixed8(1 / 512) == "0.00195312"
    assert oracle.fixed8(3 / 512) == "0.00585938"
    assert oracle.fixed8(-0.0) == "-0.00000000"
    assert oracle.fixed8(-1e-12) == "-0.00000000"
    assert oracle.fixed8(float("inf")) == "inf" and oracle.fixed8(-float("inf")) == "-inf"
    assert oracle.fixed8(float("nan")) == "nan"
    assert oracle.fixed8(-float("nan")) == "-nan"


def test_final_positions_exact_shapes_and_wrap():
    rng = np.random.default_rng(6)
    n = 300
    box = np.array([36.15, 28.553, 40.0])
    x = rng.uniform(-0.5, 0.5, (3, n)) * box[:, None]
    u = rng.normal(0, 3.0, (3, n))
    u[:, :20] = 0.0
    x[:, :20] = box[:, None] / 2       # exactly on the +L/2 face: not wrapped
    for U in (np.eye(3), np.eye(3) + rng.normal(0, 1e-3, (3, 3))):
        f = oracle.final_positions(*x, *u, box, U)
        for k in range(n):
            s = []
            for a in range(3):
                p = x[a, k] + u[a, k]
                h = box[a] / 2.0
                if p > h:
                    p -= box[a]
                elif p < -h:
                    p += box[a]
                s.append(p + h)
            for r in range(3):
                want = fma(U[r, 2], s[2], fma(U[r, 0], s[0], U[r, 1] * s[1]))
                assert f[r][k] == want, (k, r)
        if np.array_equal(U, np.eye(3)):
            # identity U leaves wrapped + L/2, in [0, L] after one wrap
            assert all(((f[a] >= 0) & (f[a] <= box[a])).all() for a in range(3)
                       if np.abs(u[a]).max() < box[a] / 2)


def _gram_schmidt_exact_shapes(Hp):
    a1, a2, a3 = Hp[:, 0], Hp[:, 1], Hp[:, 2]
    lx = float(np.sqrt(dot3(a1, a1)))
    e1 = [a / lx for a in a1]
    xy = dot3(a2, e1)
    r2 = [fma(-xy, e1[i], a2[i]) for i in range(3)]
    ly = float(np.sqrt(dot3(r2, r2)))
    e2 = [a / ly for a in r2]
    xz = dot3(a3, e1)
    r3 = [fma(-xz, e1[i], a3[i]) for i in range(3)]
    yz = dot3(r3, e2)
    r3p = [fma(-yz, e2[i], r3[i]) for i in range(3)]
    lz = float(np.sqrt(dot3(r3p, r3p)))
    e3 = [a / lz for a in r3p]
    return [lx, ly, lz, xy, xz, yz], [e1, e2, e3]


@pytest.mark.parametrize("seed", range(6))
def test_lammps_box_oracle_and_host_code(seed):
    rng = np.random.default_rng(seed)
    Hp = np.diag(rng.uniform(20, 400, 3)) + (rng.normal(0, 3, (3, 3)) if seed else 0)
    b, fr = oracle.lammps_box(Hp)
    want_b, want_e = _gram_schmidt_exact_shapes(Hp)
    assert list(b) == want_b
    assert fr.tolist() == want_e
    hb, hfr = W.lammps_box(Hp)              # the product's host code
    assert np.array_equal(hb, b) and np.array_equal(hfr, fr)
    if seed == 0:
        assert list(b[3:]) == [0.0, 0.0, 0.0] and np.array_equal(fr, np.eye(3))


def test_xyz_file_layout(tmp_path):
    rng = np.random.default_rng(7)
    n = 500
    f = [special_values(rng, 4 * n)[:n], rng.uniform(0, 50, n), rng.uniform(0, 50, n)]
    grain = rng.integers(0, 9, n).astype(np.int32)
    twin = rng.integers(-1, 2, n).astype(np.int32)
    Hp = np.diag([36.15, 50.0, 50.125])
    p = tmp_path / "a.xyz"
    oracle.write_xyz(p, *f, Hp, "Cu", grain, twin)
    lines = p.read_text().split("\n")
    assert lines[0] == str(n)
    assert lines[1] == ('Lattice="36.15 0 0  0 50 0  0 0 50.125" '
                        'Properties=species:S:1:pos:R:3:grain:I:1:twin:I:1')
    for k in range(n):
        want = "Cu " + " ".join("%.8f" % f[a][k] for a in range(3)) + f" {grain[k] + 1} {twin[k]}"
        assert lines[2 + k] == want
    assert lines[2 + n] == "" and len(lines) == n + 3
    oracle.write_xyz(p, *f, Hp, "Fe")
    assert p.read_text().split("\n")[1].endswith("Properties=species:S:1:pos:R:3")


def test_lammps_file_layout(tmp_path):
    rng = np.random.default_rng(8)
    n = 400
    f = [rng.uniform(0, 60, n) for _ in range(3)]
    for tilt in (False, True):
        Hp = np.diag([60.0, 61.0, 62.5]) + (np.array([[0, 1.5, 0.25], [0, 0, -0.5], [0, 0, 0]])
                                            if tilt else 0)
        p = tmp_path / "a.lmp"
        oracle.write_lammps(p, *f, Hp, "Fe", 55.845)
        lines = p.read_text().split("\n")
        b, fr = oracle.lammps_box(Hp)
        head = ["# LAMMPS data file generated by exec_cuda", "", f"{n} atoms", "",
                "1 atom types", "", "0.0 %.8f xlo xhi" % b[0], "0.0 %.8f ylo yhi" % b[1],
                "0.0 %.8f zlo zhi" % b[2]]
        if tilt:
            head.append("%.8f %.8f %.8f xy xz yz" % tuple(b[3:]))
        head += ["", "Masses", "", "1 55.845  # Fe", "", "Atoms  # atomic", ""]
        assert lines[:len(head)] == head
        body = lines[len(head):]
        for k in range(n):
            r = [f[a][k] for a in range(3)]
            proj = [fma(e[2], r[2], fma(e[0], r[0], e[1] * r[1])) for e in fr]
            assert body[k] == f"{k + 1} 1 " + " ".join("%.8f" % v for v in proj)
        assert body[n] == "" and len(body) == n + 1
