"""numpy statement of the cost model of the LEARNED form (test reference).

8 -> 32 -> 32 -> 1 MLP, ReLU hidden layers, parameters rounded to BF16 (round
to nearest even) for every forward pass, FP32 arithmetic, output = feature 5 +
g * network output (a residual on the previous-step estimate, applied only
where that estimate is ambiguous or missing: g = 1 when feature 5 != feature
6, i.e. the group's bodies come from groups of different work, or feature
5 == 0, no history; else g = 0 and the output is the previous work itself),
loss mean((y - t)^2).  Parameter layout: W1[32][8] b1[32] W2[32][32] b2[32] W3[32] b3.
The backward pass below is the closed form of that loss; tests pin it against
central finite differences in float64."""
import numpy as np

F, H = 8, 32
SIZES = [("W1", (H, F)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H,)), ("b3", (1,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def bf16(a):
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(p):
    out, k = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[k:k + n]).reshape(shape)
        k += n
    return out


def forward(p, x, dtype=np.float32, round_bf16=True):
    q = unpack(bf16(p) if round_bf16 else p)
    x = np.asarray(x, dtype)
    z1 = x @ q["W1"].astype(dtype).T + q["b1"].astype(dtype)
    h1 = np.maximum(z1, 0)
    z2 = h1 @ q["W2"].astype(dtype).T + q["b2"].astype(dtype)
    h2 = np.maximum(z2, 0)
    # gated residual on feature 5 (previous step's scaled walk work)
    g = gate(x).astype(dtype)
    y = x[:, 5] + g * (h2 @ q["W3"].astype(dtype) + q["b3"].astype(dtype)[0])
    return y, (x, z1, h1, z2, h2)


def gate(x):
    """1 where the previous-step estimate is ambiguous (feature 5 != 6) or
    missing (feature 5 == 0), else 0."""
    x = np.asarray(x)
    return ((x[:, 5] != x[:, 6]) | (x[:, 5] == 0)).astype(np.float64)


def loss_and_grad(p, x, t, dtype=np.float32, round_bf16=True):
    q = unpack(bf16(p) if round_bf16 else p)
    y, (x, z1, h1, z2, h2) = forward(p, x, dtype, round_bf16)
    m = len(t)
    d = y - np.asarray(t, dtype)
    loss = float(np.mean(d * d))
    dy = (2.0 / m) * d * gate(x).astype(dtype)
    g = {}
    g["W3"] = h2.T @ dy
    g["b3"] = np.array([dy.sum()])
    d2 = np.outer(dy, q["W3"].astype(dtype)) * (z2 > 0)
    g["W2"] = d2.T @ h1
    g["b2"] = d2.sum(0)
    d1 = (d2 @ q["W2"].astype(dtype)) * (z1 > 0)
    g["W1"] = d1.T @ x
    g["b1"] = d1.sum(0)
    return loss, np.concatenate([np.ravel(g[name]) for name, _ in SIZES])


def init(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, np.float32)
    k = 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        if name.startswith("W"):
            fan = shape[1] if len(shape) > 1 else H
            p[k:k + n] = rng.uniform(-1, 1, n) * np.sqrt(6.0 / fan)
        k += n
    return p
