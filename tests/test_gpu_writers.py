"""The output stage on the device (include/disco_output.h) against the oracle's
restatement of Code/main.cpp:797-852 and writeLAMMPS (:50-135): final positions
bit for bit, atom lines and whole files byte for byte."""
import numpy as np
import pytest

import oracle
from disco_amd import backend as B
from disco_amd import geometry as G
from disco_amd import writers as W
from test_oracle_writers import special_values

pytestmark = pytest.mark.gpu


def _dev(*arrays):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).to("cuda") for a in arrays]


def _host(ts):
    return [t.cpu().numpy() for t in ts]


def _body(path, skip):
    data = open(path, "rb").read()
    for _ in range(skip):
        data = data[data.index(b"\n") + 1:]
    return data


@pytest.mark.parametrize("randomU", [False, True])
def test_final_positions_bit_exact(hip_ctx, randomU):
    rng = np.random.default_rng(11)
    n = 200_003
    box = np.array([36.15, 28.553, 40.0])
    x = rng.uniform(-0.5, 0.5, (3, n)) * box[:, None]
    u = rng.normal(0, 2.0, (3, n))
    x[:, :64] = box[:, None] / 2
    u[:, :64] = 0.0
    x[:, 64:128] = -box[:, None] / 2
    u[:, 64:128] = -1e-300
    U = np.eye(3) + (rng.normal(0, 1e-3, (3, 3)) if randomU else 0)
    f = W.final_positions_torch(hip_ctx, *_dev(*x, *u), box, U if randomU else None)
    want = oracle.final_positions(*x, *u, box, U)
    for got, w in zip(_host(f), want):
        assert np.array_equal(got, w)


@pytest.mark.parametrize("n", [0, 1, 63, 256, 257, 4099, 100_000])
def test_xyz_lines_equal_oracle(hip_ctx, tmp_path, n):
    import torch
    rng = np.random.default_rng(n)
    f = [special_values(rng, 4 * n + 64)[rng.permutation(4 * n + 64)[:n]] for _ in range(3)]
    grain = rng.integers(0, 40, n).astype(np.int32)
    twin = rng.integers(-1, 2, n).astype(np.int32)
    p = tmp_path / "o.xyz"
    Hp = np.diag([10.0, 11.0, 12.0])
    for el, g, t in (("Fe", None, None), ("Cu", grain, twin), ("Xx_long_name", grain, None)):
        oracle.write_xyz(p, *f, Hp, el, g, t)
        dg = None if g is None else torch.from_numpy(g).to("cuda")
        dt = None if t is None else torch.from_numpy(t).to("cuda")
        got = W.format_text_torch(hip_ctx, "xyz", *_dev(*f), element=el, grain=dg, twin=dt)
        assert got == _body(p, 2)


@pytest.mark.parametrize("n", [1, 255, 257, 70_001])
def test_lammps_lines_equal_oracle(hip_ctx, tmp_path, n):
    rng = np.random.default_rng(100 + n)
    f = [rng.uniform(-5, 80, n) for _ in range(3)]
    f[0][: min(n, 50)] = special_values(rng, 200)[:min(n, 50)]
    Hp = np.diag([80.0, 81.0, 82.5]) + np.array([[0, 1.5, 0.25], [0, 0, -0.5], [0, 0, 0]])
    p = tmp_path / "o.lmp"
    oracle.write_lammps(p, *f, Hp, "Fe", 55.845)
    _, frame = W.lammps_box(Hp)
    got = W.format_text_torch(hip_ctx, "lammps", *_dev(*f), first_id=1, frame=frame)
    assert got == _body(p, 17)          # header: 16 lines with the tilt line, then a blank


def test_files_equal_oracle_across_chunks(hip_ctx, tmp_path):
    """1.2M atoms: two device chunks (2^20 atoms) and several 16 MiB staging pieces."""
    import torch
    rng = np.random.default_rng(12)
    n = 1_200_000
    f = [rng.uniform(0, 400, n) for _ in range(3)]
    grain = rng.integers(0, 12, n).astype(np.int32)
    twin = rng.integers(-1, 2, n).astype(np.int32)
    Hp = np.diag([400.0, 401.0, 402.5]) + np.array([[0, 2.0, 0], [0, 0, 0], [0, 0, 0]])
    d = _dev(*f)
    dg, dt = torch.from_numpy(grain).to("cuda"), torch.from_numpy(twin).to("cuda")
    W.write_xyz(hip_ctx, tmp_path / "g.xyz", *d, Hp, "Al", dg, dt)
    oracle.write_xyz(tmp_path / "o.xyz", *f, Hp, "Al", grain, twin)
    assert open(tmp_path / "g.xyz", "rb").read() == open(tmp_path / "o.xyz", "rb").read()
    W.write_lammps(hip_ctx, tmp_path / "g.lmp", *d, Hp, "Al", 26.9815385)
    oracle.write_lammps(tmp_path / "o.lmp", *f, Hp, "Al", 26.9815385)
    assert open(tmp_path / "g.lmp", "rb").read() == open(tmp_path / "o.lmp", "rb").read()


def test_write_outputs_after_the_sum(hip_ctx, tmp_path):
    """The driver's tail on a small sample: device lattice, FP64 sum, final
    positions and both files, against the oracle's sum and writers."""
    import torch
    s = G.config_c1()
    idx = np.arange(0, s.n, 37)
    x, y, z = s.x[idx], s.y[idx], s.z[idx]
    u = B.line_sum(hip_ctx, x, y, z, s.loops, s.box, 1, B.PREC_FP64)
    dx, dy, dz, dux, duy, duz = _dev(x, y, z, *u)
    name = str(tmp_path / "out.xyz")
    f = W.write_outputs(hip_ctx, name, "both", dx, dy, dz, dux, duy, duz, s.box, "Fe", 55.845)
    torch.cuda.synchronize()
    want = oracle.final_positions(x, y, z, *u, s.box)
    for got, w in zip(_host(f), want):
        assert np.array_equal(got, w)
    Hp = np.diag(s.box)
    oracle.write_xyz(tmp_path / "o.xyz", *want, Hp, "Fe")
    oracle.write_lammps(tmp_path / "o.lmp", *want, Hp, "Fe", 55.845)
    assert open(name, "rb").read() == open(tmp_path / "o.xyz", "rb").read()
    assert open(name[:-3] + "lmp", "rb").read() == open(tmp_path / "o.lmp", "rb").read()


def test_writer_errors(hip_ctx, tmp_path):
    big = _dev(np.array([1.0, 2e10, 3.0]), np.zeros(3), np.zeros(3))
    with pytest.raises(B.DiscoError) as e:
        W.format_text_torch(hip_ctx, "xyz", *big)
    assert e.value.code == 5
    ok = _dev(np.ones(4), np.ones(4), np.ones(4))
    with pytest.raises(B.DiscoError) as e:
        W.write_xyz(hip_ctx, tmp_path / "missing_dir" / "a.xyz", *ok, np.eye(3), "Fe")
    assert e.value.code == 7
    with pytest.raises(B.DiscoError):
        W.format_text_torch(hip_ctx, "xyz", *ok, element="X" * 16)
    with pytest.raises(B.DiscoError):
        W.final_positions_torch(hip_ctx, *ok, *ok, [1.0, 0.0, 1.0])
    # an empty write is the header alone
    empty = _dev(np.zeros(0), np.zeros(0), np.zeros(0))
    W.write_xyz(hip_ctx, tmp_path / "e.xyz", *empty, np.diag([1.0, 2.0, 3.0]), "Fe")
    oracle.write_xyz(tmp_path / "o.xyz", np.zeros(0), np.zeros(0), np.zeros(0),
                     np.diag([1.0, 2.0, 3.0]), "Fe")
    assert open(tmp_path / "e.xyz", "rb").read() == open(tmp_path / "o.xyz", "rb").read()
