"""FP64 restatement of the learned pose proposal network (TEST INFRASTRUCTURE).

The reference has no learned component; this is the CPU oracle of the project's own network, against
which the HIP training and inference kernels are checked.  Only tests/ may import it.

Network (all sizes from the caller):
    x  [B][Fin]                       input features (Fin = 2 * 812 low-resolution coefficients, zero-padded)
    a1 = x  W1^T + b1,  h1 = gelu(a1)  W1 [H][Fin]
    a2 = h1 W2^T + b2,  h2 = gelu(a2)  W2 [H][H]
    a3 = h2 W3^T + b3,  h3 = gelu(a3)  W3 [H][H]
    z  = h3 W4^T + b4                  W4 [O][H],  O = K * C + 2 (cell logits, then the 2 shift outputs)
gelu(a) = a / 2 (1 + erf(a / sqrt 2)).
Loss per image: cross entropy of softmax(z[:K*C]) against the target distribution over cells, plus
lam / 2 * |z[K*C:] - s|^2 for the shift target s; the batch loss is the mean over images.

bf16=True rounds every matrix-product operand to bfloat16 (round to nearest even) the way the GPU feeds its
MFMA units (weights, layer inputs, and in the backward pass the incoming gradients), keeping FP64
everywhere else; bf16=False is the plain FP64 network.
"""
import numpy as np
from scipy.special import erf

SQRT1_2 = 1.0 / np.sqrt(2.0)


def to_bf16(a):
    """Round FP64/FP32 values to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64).reshape(np.shape(a))


def gelu(a):
    return 0.5 * a * (1.0 + erf(a * SQRT1_2))


def gelu_grad(a):
    return 0.5 * (1.0 + erf(a * SQRT1_2)) + a * np.exp(-0.5 * a * a) / np.sqrt(2.0 * np.pi)


def forward(params, x, bf16=False):
    """params: dict W1..W4, b1..b4 (float64).  Returns (z, cache)."""
    q = to_bf16 if bf16 else (lambda v: v)
    cache = {"x": q(x)}
    h = cache["x"]
    for l in (1, 2, 3):
        a = h @ q(params[f"W{l}"]).T + params[f"b{l}"]
        cache[f"a{l}"] = a
        h = q(gelu(a))
        cache[f"h{l}"] = h
    z = h @ q(params["W4"]).T + params["b4"]
    return z, cache


def loss_and_grad_z(z, targets, shift, ncell, lam):
    """targets: [B][ncell] distributions; shift: [B][2].  Returns (mean loss, dL/dz)."""
    B = z.shape[0]
    lg = z[:, :ncell]
    m = lg.max(axis=1, keepdims=True)
    e = np.exp(lg - m)
    s = e.sum(axis=1, keepdims=True)
    logp = lg - m - np.log(s)
    ce = -(targets * logp).sum(axis=1)
    d = z[:, ncell:] - shift
    loss = float(np.mean(ce + 0.5 * lam * (d * d).sum(axis=1)))
    dz = np.empty_like(z)
    dz[:, :ncell] = (e / s - targets) / B
    dz[:, ncell:] = lam * d / B
    return loss, dz


def backward(params, cache, dz, bf16=False):
    q = to_bf16 if bf16 else (lambda v: v)
    g = {}
    dzq = q(dz)
    g["W4"] = dzq.T @ cache["h3"]
    g["b4"] = dz.sum(axis=0)
    dh = dzq @ q(params["W4"])
    for l in (3, 2, 1):
        da = dh * gelu_grad(cache[f"a{l}"])
        daq = q(da)
        inp = cache["x"] if l == 1 else cache[f"h{l - 1}"]
        g[f"W{l}"] = daq.T @ inp
        g[f"b{l}"] = da.sum(axis=0)
        if l > 1:
            dh = daq @ q(params[f"W{l}"])
    return g


def loss_and_grads(params, x, targets, shift, ncell, lam, bf16=False):
    z, cache = forward(params, x, bf16)
    loss, dz = loss_and_grad_z(z, targets, shift, ncell, lam)
    return loss, backward(params, cache, dz, bf16), z


def adam_step(params, grads, state, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (bias-corrected), in place; state holds m, v and the step count t."""
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    for k, gk in grads.items():
        m = state.setdefault("m_" + k, np.zeros_like(gk))
        v = state.setdefault("v_" + k, np.zeros_like(gk))
        m[...] = beta1 * m + (1 - beta1) * gk
        v[...] = beta2 * v + (1 - beta2) * gk * gk
        mh = m / (1 - beta1 ** t)
        vh = v / (1 - beta2 ** t)
        params[k] = params[k] - lr * mh / (np.sqrt(vh) + eps)
