"""CPU oracle of the learned proposal's network (TEST INFRASTRUCTURE ONLY: used by tests/ as the
checker, never by the product).

Restates, in float64 numpy, what mcintegratorplusplus_amd/csrc/mci_mlp.hip and mci_lp_api.cpp
compute: the MLP d -> H -> H -> K (1 + 2 d) with SiLU, the diagonal Gaussian-mixture head
(outputs per component k: d means / 4 then d log-scale pre-activations, log-scale = o clamped to [-3, 3];
then K logits; input = x / 4), the mean negative
log-likelihood of targets y under q(. | x), and its gradient by backpropagation.  With
emulate=True the operands of every matrix product are rounded to bfloat16 (round to nearest
even) at the same points as on the GPU (input, weights, SiLU outputs, output gradient, SiLU
input gradients); with emulate=False everything is exact float64, which the tests pin against
central finite differences.
"""
import numpy as np

IN_SCALE, MU_SCALE, LS_MAX = 0.25, 4.0, 3.0
HALF_LOG_2PI = 0.5 * np.log(2.0 * np.pi)


def bf16(a):
    """float32 values rounded to bfloat16 (nearest even), as float64"""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def _f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def sizes(d, H, K):
    O = K * (1 + 2 * d)
    return [("W1", (d, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, O)), ("b3", (O,))]


def unpack(p, d, H, K):
    out, o = {}, 0
    for name, shp in sizes(d, H, K):
        n = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(t, d, H, K):
    return np.concatenate([t[name].ravel() for name, _ in sizes(d, H, K)])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, d, H, K, emulate=True):
    q = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    P = unpack(p, d, H, K)
    f = {}
    f["Xb"] = q(_f32(x) * IN_SCALE) if emulate else np.asarray(x, np.float64) * IN_SCALE
    f["Z1"] = f["Xb"] @ q(P["W1"]) + P["b1"]
    f["H1"] = q(silu(f["Z1"]))
    f["Z2"] = f["H1"] @ q(P["W2"]) + P["b2"]
    f["H2"] = q(silu(f["Z2"]))
    f["OUT"] = f["H2"] @ q(P["W3"]) + P["b3"]
    return f, P


def _lse(a, axis):
    m = np.max(a, axis=axis, keepdims=True)
    return (m + np.log(np.sum(np.exp(a - m), axis=axis, keepdims=True))).squeeze(axis)


def _split(out, d, K):
    """(logits [n][K], mean pre-activations [n][K][d], log-scale pre-activations [n][K][d])"""
    n = out.shape[0]
    comp = out[:, :2 * d * K].reshape(n, K, 2 * d)
    return out[:, 2 * d * K:2 * d * K + K], comp[:, :, :d], comp[:, :, d:]


def _join(g_a, g_mu, g_ls):
    n, K, d = g_mu.shape
    return np.concatenate([np.concatenate([g_mu, g_ls], axis=2).reshape(n, 2 * d * K), g_a], axis=1)


def mixture_logq(out, y, d, K):
    """log q(y | parameters in out), rows of out [n][K (1 + 2 d)]"""
    a, om, ol = _split(out, d, K)
    mu = MU_SCALE * om
    ls = np.clip(ol, -LS_MAX, LS_MAX)
    z = (np.asarray(y, np.float64)[:, None, :] - mu) * np.exp(-ls)
    comp = a - _lse(a, 1)[:, None] + np.sum(-ls - HALF_LOG_2PI - 0.5 * z * z, axis=2)
    return _lse(comp, 1)


def head(out, y, d, K):
    """per-row loss -log q(y | out) and d(mean loss)/d(out)"""
    n = out.shape[0]
    a, om, ol = _split(out, d, K)
    mu = MU_SCALE * om
    inside = np.abs(ol) < LS_MAX
    ls = np.clip(ol, -LS_MAX, LS_MAX)
    isg = np.exp(-ls)
    z = (np.asarray(y, np.float64)[:, None, :] - mu) * isg
    lsa = _lse(a, 1)
    comp = a - lsa[:, None] + np.sum(-ls - HALF_LOG_2PI - 0.5 * z * z, axis=2)
    L = _lse(comp, 1)
    gam = np.exp(comp - L[:, None])
    pi = np.exp(a - lsa[:, None])
    g_a = pi - gam
    g_mu = (-gam[:, :, None] * z * isg) * MU_SCALE
    g_ls = gam[:, :, None] * (1.0 - z * z) * inside
    return -L, _join(g_a, g_mu, g_ls) / n


def loss_grad(p, x, y, d, H, K, emulate=True):
    """(mean loss, flat gradient) with the GPU's rounding points when emulate"""
    q = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    f, P = forward(p, x, d, H, K, emulate)
    loss, dO = head(f["OUT"], y, d, K)
    dOb = q(dO)
    g = {"b3": dO.sum(axis=0), "W3": f["H2"].T @ dOb}
    dH2 = dOb @ q(P["W3"]).T
    dZ2 = dH2 * dsilu(f["Z2"])
    dZ2b = q(dZ2)
    g["b2"] = dZ2.sum(axis=0)
    g["W2"] = f["H1"].T @ dZ2b
    dH1 = dZ2b @ q(P["W2"]).T
    dZ1 = dH1 * dsilu(f["Z1"])
    dZ1b = q(dZ1)
    g["b1"] = dZ1.sum(axis=0)
    g["W1"] = f["Xb"].T @ dZ1b
    return float(np.mean(loss)), pack(g, d, H, K)
