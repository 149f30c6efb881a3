"""FP64 numpy restatement of the surrogate MLP (test infrastructure).

Layout of the flat parameter vector (as disco_hip_surrogate_get_params):
W1 (128 x 16), b1 (128), W2 (128 x 128), b2, W3 (128 x 128), b3, W4 (32 x 128),
b4 (32); rows 3..31 of W4 and b4[3:] are zero pads.  With ``bf16=True`` the
values the GPU kernels round to bfloat16 are rounded at the same points:
the hidden activations, W2..W4 and the backward deltas dY, dZ3, dZ2."""
import numpy as np

KIN, KH, KOUT = 16, 128, 32
SIZES = [("W1", (KH, KIN)), ("b1", (KH,)), ("W2", (KH, KH)), ("b2", (KH,)),
         ("W3", (KH, KH)), ("b3", (KH,)), ("W4", (KOUT, KH)), ("b4", (KOUT,))]


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    out, o = {}, 0
    for name, shp in SIZES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n], np.float64).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([d[name].reshape(-1) for name, _ in SIZES])


def _act(z):
    """Hard swish: z clip(z/6 + 1/2, 0, 1)."""
    return z * np.clip(z / 6.0 + 0.5, 0.0, 1.0)


def _act_grad(z):
    g = z / 6.0 + 0.5
    return np.where(g <= 0.0, 0.0, np.where(g >= 1.0, 1.0, z / 3.0 + 0.5))


def forward(p, X14, rnd=True):
    P = unpack(p)
    X = np.zeros((X14.shape[0], KIN))
    X[:, :14] = X14
    r = bf16 if rnd else (lambda a: np.asarray(a, np.float64))
    c = {"X": X}
    c["Z1"] = X @ P["W1"].T + P["b1"]
    c["A1"] = r(_act(c["Z1"]))
    c["Z2"] = c["A1"] @ r(P["W2"]).T + P["b2"]
    c["A2"] = r(_act(c["Z2"]))
    c["Z3"] = c["A2"] @ r(P["W3"]).T + P["b3"]
    c["A3"] = r(_act(c["Z3"]))
    c["Y"] = c["A3"] @ r(P["W4"]).T + P["b4"]
    return c["Y"][:, :3], c


def loss_grad(p, X14, T, rnd=True):
    P = unpack(p)
    r = bf16 if rnd else (lambda a: np.asarray(a, np.float64))
    Y3, c = forward(p, X14, rnd)
    M = X14.shape[0]
    d = Y3 - T
    loss = float((d ** 2).mean())
    dY = np.zeros((M, KOUT))
    dY[:, :3] = r(2.0 * d / (3.0 * M))
    g = {}
    g["W4"] = dY.T @ c["A3"]
    g["b4"] = dY.sum(0)
    dA = dY @ r(P["W4"])

    dZ = r(dA * _act_grad(c["Z3"]))
    g["W3"] = dZ.T @ c["A2"]
    g["b3"] = dZ.sum(0)
    dA = dZ @ r(P["W3"])
    dZ = r(dA * _act_grad(c["Z2"]))
    g["W2"] = dZ.T @ c["A1"]
    g["b2"] = dZ.sum(0)
    dA = dZ @ r(P["W2"])
    dZ1 = dA * _act_grad(c["Z1"])
    g["W1"] = dZ1.T @ c["X"]
    g["b1"] = dZ1.sum(0)
    g["W4"][3:] = 0.0
    g["b4"][3:] = 0.0
    return loss, pack(g)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in SIZES:
        if name.startswith("W"):
            a = np.sqrt(6.0 / shp[1])
            d[name] = rng.uniform(-a, a, size=shp)
        else:
            d[name] = rng.normal(scale=0.05, size=shp)
    d["W1"][:, 14:] = 0.0
    d["W4"][3:] = 0.0
    d["b4"][3:] = 0.0
    return pack(d).astype(np.float32)
