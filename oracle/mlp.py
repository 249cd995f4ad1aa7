"""numpy f64 restatement of the warm-start MLP (12 -> 64 -> 64 -> 64 -> 2,
GELU in its tanh form) and of its loss and gradient, for checking the HIP
training and inference kernels.

Test infrastructure: only tests/ import this module, as the checker.  The
parameter blob layout is the library's (nh_mlp_param_count): feature mean[12],
feature sd[12], W1[64x12], b1[64], W2[64x64], b2[64], W3[64x64], b3[64],
W4[2x64], b4[2], target mean[2], target sd[2]; W row-major (out x in).

The network predicts the correction to stage one's estimate: its target is
(log sigma^2, log phi) minus features 10 and 11, standardised; predictions add
those features back.  The loss is Huber (delta 1) on the standardised target.

Pinned by central finite differences of loss() (tests/test_oracle_mlp.py)
and by closed forms of GELU (gelu(0) = 0, gelu'(0) = 1/2, odd part).
"""
import numpy as np

IN, HID, OUT = 12, 64, 2
BASE = 10  # features 10, 11: stage one's (log sigma^2, log phi)
SHAPES = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, HID)), ("b3", (HID,)), ("W4", (OUT, HID)), ("b4", (OUT,))]
NET = sum(int(np.prod(s)) for _, s in SHAPES)
COUNT = 2 * IN + NET + 2 * OUT
C = np.sqrt(2.0 / np.pi)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    assert p.size == COUNT
    out = {"fm": p[:IN], "fs": p[IN:2 * IN]}
    at = 2 * IN
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[at:at + n].reshape(shape)
        at += n
    out["tm"] = p[at:at + OUT]
    out["ts"] = p[at + OUT:at + 2 * OUT]
    return out


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(C * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * C * (1.0 + 3 * 0.044715 * x * x)


def forward(params, features, rnd=None):
    """Standardised outputs and the per-layer cache.  rnd, if given, rounds
    weights and layer inputs (as the inference kernel does in BF16)."""
    q = unpack(params)
    r = rnd or (lambda a: a)
    x = (np.asarray(features, dtype=np.float64) - q["fm"]) / q["fs"]
    cache = [x]
    a = x
    for i in (1, 2, 3):
        z = r(a) @ r(q[f"W{i}"]).T + q[f"b{i}"]
        a = gelu(z)
        cache += [z, a]
    out = r(a) @ r(q["W4"]).T + q["b4"]
    return out, cache


def loss_grad(params, features, targets):
    """Mean Huber loss (delta 1) over the standardised corrections, and its
    gradient with respect to the network part of the blob (W1..b4 in blob
    order)."""
    q = unpack(params)
    f = np.asarray(features, dtype=np.float64)
    n = len(f)
    out, cache = forward(params, f)
    res = np.asarray(targets, dtype=np.float64) - f[:, BASE:BASE + OUT]
    t = (res - q["tm"]) / q["ts"]
    e = out - t
    d = np.clip(e, -1.0, 1.0) / n
    loss = np.sum(np.where(np.abs(e) <= 1.0, 0.5 * e * e, np.abs(e) - 0.5)) / n
    x, z1, a1, z2, a2, z3, a3 = cache
    g = {"W4": d.T @ a3, "b4": d.sum(0)}
    dz3 = (d @ q["W4"]) * gelu_grad(z3)
    g["W3"], g["b3"] = dz3.T @ a2, dz3.sum(0)
    dz2 = (dz3 @ q["W3"]) * gelu_grad(z2)
    g["W2"], g["b2"] = dz2.T @ a1, dz2.sum(0)
    dz1 = (dz2 @ q["W2"]) * gelu_grad(z1)
    g["W1"], g["b1"] = dz1.T @ x, dz1.sum(0)
    flat = np.concatenate([g[name].ravel() for name, _ in SHAPES])
    return loss, flat


def bf16(a):
    """Round to bfloat16 (nearest even), returned as f64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def predict(params, features, rnd=None):
    """(log sigma^2, log phi) per row."""
    q = unpack(params)
    f = np.asarray(features, dtype=np.float64)
    out, _ = forward(params, f, rnd)
    return out * q["ts"] + q["tm"] + f[:, BASE:BASE + OUT]


def random_params(seed, feature_scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(COUNT)
    p[:IN] = rng.normal(size=IN)
    p[IN:2 * IN] = feature_scale * rng.uniform(0.5, 2.0, size=IN)
    at = 2 * IN
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        fan = shape[-1] if len(shape) == 2 else 1
        p[at:at + n] = rng.normal(size=n) * (1.0 / np.sqrt(fan) if len(shape) == 2 else 0.1)
        at += n
    p[at:at + OUT] = rng.normal(size=OUT)
    p[at + OUT:] = rng.uniform(0.5, 2.0, size=OUT)
    return p
