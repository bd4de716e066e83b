"""Pins the CPU restatement of the closure network (oracle/closure.py)
before it checks the HIP kernels:

  bf16 rounding     equals PyTorch's float32 -> bfloat16 conversion
  GELU              gelu(x) - gelu(-x) = x; gelu'(x) against central
                    differences
  loss gradient     the analytic backward pass against central differences
                    of the FP64 loss, on every layer's weights and biases
  Adam              the first step moves each weight by lr sign(g) (the
                    bias-corrected moments are g and g^2)
"""
import numpy as np
import pytest

from oracle import closure as oc


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    a = np.concatenate([rng.standard_normal(5000) * 10.0 ** rng.uniform(-8, 8, 5000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.0078125, 1.01171875, 3.0e38, -1e-40]]).astype(np.float32)
    want = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(oc.bf16(a), want)


def test_gelu_identities():
    x = np.linspace(-6, 6, 241)
    assert np.allclose(oc.gelu(x) - oc.gelu(-x), x, atol=1e-14)
    assert oc.gelu(np.array([0.0]))[0] == 0.0
    h = 1e-5
    fd = (oc.gelu(x + h) - oc.gelu(x - h)) / (2 * h)
    assert np.abs(fd - oc.dgelu(x)).max() < 1e-8


def test_gradient_against_finite_differences():
    rng = np.random.default_rng(2)
    n = 9
    p = oc.init_params(3).astype(np.float64)
    x = rng.standard_normal((n, oc.IN))
    t = rng.standard_normal((n, oc.OUT))
    shift, scale, oscale = rng.standard_normal(oc.IN) * 0.1, 1.0 + 0.2 * rng.random(oc.IN), 1.0 + rng.random(oc.OUT)
    loss, g = oc.loss_and_grad(p, x, t, shift, scale, oscale, bf16_round=False, dtype=np.float64)
    assert loss > 0
    Ws, bs = oc.unpack(np.arange(oc.NPARAM))
    picks = []
    for l in range(4):
        no = oc.NOUT[l]
        nin = oc.IN if l == 0 else oc.W
        nout = oc.OUT if l == 3 else oc.W
        for _ in range(6):
            picks.append(oc.OFF_W[l] + int(rng.integers(nin)) * no + int(rng.integers(nout)))
        picks.append(oc.OFF_B[l] + int(rng.integers(nout)))
    h = 1e-6
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = oc.loss_and_grad(pp, x, t, shift, scale, oscale, bf16_round=False, dtype=np.float64)
        lm, _ = oc.loss_and_grad(pm, x, t, shift, scale, oscale, bf16_round=False, dtype=np.float64)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 * max(1.0, abs(g).max()), (i, fd, g[i])
    # padding rows and columns get no gradient
    W1 = g[oc.OFF_W[0]:oc.OFF_W[0] + oc.W * oc.W].reshape(oc.W, oc.W)
    W4 = g[oc.OFF_W[3]:oc.OFF_W[3] + oc.W * oc.OUTP].reshape(oc.W, oc.OUTP)
    assert np.all(W1[oc.IN:] == 0) and np.all(W4[:, oc.OUT:] == 0)


def test_adam_first_step_is_a_signed_step():
    rng = np.random.default_rng(4)
    p = rng.standard_normal(1000).astype(np.float32)
    g = rng.standard_normal(1000).astype(np.float32)
    z = np.zeros(1000, np.float32)
    p1, m, v = oc.adam(p, z, z, g, 1, 1e-3)
    assert np.allclose(p1 - p, -1e-3 * np.sign(g), rtol=1e-4, atol=1e-9)
    assert np.allclose(m, 0.1 * g, rtol=1e-6) and np.allclose(v, 0.001 * g * g, rtol=1e-5)
