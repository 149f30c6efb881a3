"""The FP64 oracle against the reference's own outputs.

Golden vectors come from the reference's sources compiled here
(tests/golden/make_golden.py); where the compiled reference is present the
oracle is also compared with it directly on fresh random inputs."""
import numpy as np
import pytest

import oracle
from conftest import golden
from disco_amd import geometry as G
from disco_amd.backend import LoopData


def test_barnett_loop_golden():
    g = golden("barnett_f64.npz")
    worst = 0.0
    for l in range(len(g["voff"]) - 1):
        V = g["V"][g["voff"][l]:g["voff"][l + 1]]
        P = g["p"][g["poff"][l]:g["poff"][l + 1]]
        U = g["u"][g["poff"][l]:g["poff"][l + 1]]
        for p, u in zip(P, U):
            worst = max(worst, np.abs(oracle.barnett_loop(p, V, g["c"][l], g["b"][l]) - u).max())
    assert worst < 1e-12, worst


def test_barnett_loop_float_golden():
    """The float instantiation (the reference's --fp32 pair term)."""
    g = golden("barnett_f64.npz")
    gf = golden("barnett_f32.npz")
    worst = 0.0
    for l in range(len(g["voff"]) - 1):
        V = g["V"][g["voff"][l]:g["voff"][l + 1]]
        a, b = g["poff"][l], g["poff"][l + 1]
        for p, u in zip(g["p"][a:b], gf["u"][a:b]):
            worst = max(worst, np.abs(oracle.barnett_loop_f(p, V, g["c"][l], g["b"][l]) - u).max())
    # same float formula; contraction of products into FMAs may differ
    assert worst < 1e-4, worst


def test_kroupa_pair_golden():
    g = golden("kroupa_pair.npz")
    assert abs(oracle.k1() - float(g["k1"])) < 1e-17
    for rho, ph, u in zip(g["rho"], g["phys"], g["u"]):
        got = oracle.kroupa_pair(rho, ph)
        assert np.abs(got - u).max() <= 1e-14 * max(1.0, np.abs(u).max())
    assert np.all(g["u"][:9] == 0.0)           # |rho| < 1e-10 contributes nothing


def _line_loops(g):
    loops = []
    for l in range(len(g["voff"]) - 1):
        a, b = g["voff"][l], g["voff"][l + 1]
        loops.append(LoopData(phys=tuple(g["phys"][10 * l:10 * l + 10]), line=True,
                              vx=g["vx"][a:b], vy=g["vy"][a:b], vz=g["vz"][a:b],
                              cx=g["c"][3 * l], cy=g["c"][3 * l + 1], cz=g["c"][3 * l + 2]))
    return loops


def test_line_sum_golden_c1():
    g = golden("cpu_line_c1.npz")
    u = np.stack(oracle.line_sum(g["x"], g["y"], g["z"], _line_loops(g), g["box"], 1))
    assert np.abs(u - g["u"]).max() < 1e-12


def test_loopon_sum_golden_c3():
    g = golden("cpu_loopon_c3.npz")
    loops = [LoopData(phys=tuple(g["phys"][:10]), line=False, x=g["px"], y=g["py"], z=g["pz"])]
    u = np.stack(oracle.loopon_sum(g["x"], g["y"], g["z"], loops, g["box"], 1))
    # SIMD reduction order may differ from the reference build's
    assert np.abs(u - g["u"]).max() < 1e-12


def test_golden_inputs_match_the_workload_generator():
    """The fixture's atoms and loops are those the harness builds for C1."""
    g = golden("cpu_line_c1.npz")
    s = G.config_c1()
    assert np.array_equal(s.x[g["idx"]], g["x"]) and np.array_equal(s.z[g["idx"]], g["z"])
    for lp, ref in zip(s.loops, _line_loops(g)):
        assert np.array_equal(lp.vx, ref.vx) and np.array_equal(lp.vz, ref.vz)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build not present")
def test_oracle_vs_reference_build_random():
    """Fresh random loops and atoms, +-1 images: oracle == compiled kroupaCPU."""
    rng = np.random.default_rng(99)
    loops = []
    for k in range(3):
        c = rng.uniform(-20, 20, 3)
        lp = G.Loop("polygon", list(c), [1, 1, 1], [1, -1, 0], 2.8553,
                    vertices=G.regular_polygon(int(rng.integers(5, 40)), float(rng.uniform(5, 20)),
                                               [1, 1, 1], [1, -1, 0]))
        loops.append(lp.line_data())
    x, y, z = (rng.uniform(-40, 40, 300) for _ in range(3))
    box = (80.0, 85.0, 90.0)
    ref = np.stack(oracle.ref_kroupa_cpu_line(x, y, z, loops, box)[:3])
    got = np.stack(oracle.line_sum(x, y, z, loops, box, 1))
    assert np.abs(got - ref).max() < 1e-12
    lps = [G.Loop("circle", list(rng.uniform(-20, 20, 3)), [1, 1, 1], [1, -1, 0], 2.8553,
                  radius=8.0).loopon_data(0.9)]
    ref = np.stack(oracle.ref_kroupa_cpu_loopon(x, y, z, lps, box)[:3])
    got = np.stack(oracle.loopon_sum(x, y, z, lps, box, 1))
    assert np.abs(got - ref).max() < 1e-12
