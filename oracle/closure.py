"""CPU restatement of the learned closure's network (include/spm_closure.h):
126 (padded to 128) -> 128 -> 128 -> 128 -> 6 with exact GELU, its
mean-squared-error gradient and the Adam update.

TEST INFRASTRUCTURE ONLY: used by tests/ to check the HIP kernels
(kapsel_old_amd/csrc/closure.hip), never by the product.  The network has no
counterpart in the reference; the mathematics (GELU, backpropagation, Adam)
is pinned in tests/test_oracle_closure.py by finite differences, by the
error function of the C library and by PyTorch's bfloat16 conversion.

`bf16=True` rounds where the GPU path rounds: the normalised input and every
hidden activation going into a product, the weights, and in the backward
pass the output gradient and every dZ going into a product.
"""
import math

import numpy as np

IN, OUT, W, OUTP = 126, 6, 128, 16
NOUT = [W, W, W, OUTP]
OFF_W = [0, W * W + W, 2 * (W * W + W), 3 * (W * W + W)]
OFF_B = [W * W, 2 * W * W + W, 3 * W * W + 2 * W, 3 * (W * W + W) + W * OUTP]
NPARAM = 3 * (W * W + W) + W * OUTP + OUTP


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    return np.where(nan, a, out).astype(np.float32)


_erf = np.vectorize(math.erf, otypes=[np.float64])


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def dgelu(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def unpack(p):
    """Weights [in][out] and biases of the four layers from the flat vector."""
    Ws = [p[OFF_W[l]:OFF_W[l] + W * NOUT[l]].reshape(W, NOUT[l]) for l in range(4)]
    bs = [p[OFF_B[l]:OFF_B[l] + NOUT[l]] for l in range(4)]
    return Ws, bs


def init_params(seed):
    """Glorot-uniform weights of the right shapes (any deterministic values
    serve the tests; the library draws its own)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    nin, nout = [IN, W, W, W], [W, W, W, OUT]
    for l in range(4):
        lim = math.sqrt(6.0 / (nin[l] + nout[l]))
        w = np.zeros((W, NOUT[l]), dtype=np.float32)
        w[:nin[l], :nout[l]] = rng.uniform(-lim, lim, (nin[l], nout[l]))
        p[OFF_W[l]:OFF_W[l] + W * NOUT[l]] = w.ravel()
        p[OFF_B[l]:OFF_B[l] + nout[l]] = rng.uniform(-0.1, 0.1, nout[l])
    return p


def forward(p, x, shift, scale, oscale, bf16_round=True, dtype=np.float32):
    """Outputs y[n][6] (physical units) and the saved activations."""
    Ws, bs = unpack(np.asarray(p, dtype=dtype))
    rnd = bf16 if bf16_round else (lambda a: a)
    n = x.shape[0]
    h0 = np.zeros((n, W), dtype=dtype)
    h0[:, :IN] = (x - shift) * scale
    h = [rnd(h0).astype(dtype)]
    zs = []
    for l in range(3):
        z = h[-1] @ rnd(Ws[l]).astype(dtype) + bs[l]
        zs.append(z)
        h.append(rnd(gelu(z).astype(dtype)).astype(dtype))
    yn = (h[-1] @ rnd(Ws[3]).astype(dtype) + bs[3])[:, :OUT]
    return yn * oscale, (h, zs, yn)


def loss_and_grad(p, x, t, shift, scale, oscale, bf16_round=True, dtype=np.float32):
    """Mean squared error over the n x 6 normalised outputs and its gradient
    with respect to the flat parameter vector."""
    rnd = bf16 if bf16_round else (lambda a: a)
    Ws, _ = unpack(np.asarray(p, dtype=dtype))
    _, (h, zs, yn) = forward(p, x, shift, scale, oscale, bf16_round, dtype)
    n = x.shape[0]
    d = yn - t / oscale
    loss = float((d * d).sum() / (n * OUT))
    dz = np.zeros((n, OUTP), dtype=dtype)
    dz[:, :OUT] = 2.0 * d / (n * OUT)
    g = np.zeros(NPARAM, dtype=dtype)
    for l in range(3, -1, -1):
        g[OFF_B[l]:OFF_B[l] + NOUT[l]] = dz.sum(axis=0)
        g[OFF_W[l]:OFF_W[l] + W * NOUT[l]] = (h[l].T @ rnd(dz).astype(dtype)).ravel()
        if l == 0:
            break
        dh = rnd(dz).astype(dtype) @ rnd(Ws[l]).astype(dtype).T
        dz = (dh * dgelu(zs[l - 1])).astype(dtype)
    return loss, g


def adam(p, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t counted from 1), float32 as on the GPU."""
    p, m, v, g = (np.asarray(a, dtype=np.float32) for a in (p, m, v, g))
    m = np.float32(b1) * m + np.float32(1 - b1) * g
    v = np.float32(b2) * v + np.float32(1 - b2) * g * g
    c1 = np.float32(1.0 / (1.0 - np.float32(b1) ** t))
    c2 = np.float32(1.0 / (1.0 - np.float32(b2) ** t))
    p = p - np.float32(lr) * (m * c1) / (np.sqrt(v * c2) + np.float32(eps))
    return p.astype(np.float32), m.astype(np.float32), v.astype(np.float32)
