"""The closure network's HIP kernels (csrc/closure.hip) against the CPU
restatement oracle/closure.py, which rounds to bfloat16 where the GPU path
rounds: forward outputs, the loss gradient of every parameter, the Adam
step, training on a smooth synthetic map, and determinism.

Tolerances: FP32 accumulation in another order can flip the bfloat16
rounding of an activation now and then (one unit in the 8th bit), so
outputs are compared normwise to 1e-2 of their largest value and gradients
per layer to 2e-2 of the layer's largest entry; typical differences are
1e-5 and 1e-4."""
import numpy as np
import pytest

from oracle import closure as oc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Closure():
    from kapsel_old_amd.spm import Closure as _C
    return _C


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((n, oc.IN)) * 2.0 + 0.5).astype(np.float32)
    t = rng.standard_normal((n, oc.OUT)).astype(np.float32)
    shift = (rng.standard_normal(oc.IN) * 0.3).astype(np.float32)
    scale = (0.5 + rng.random(oc.IN)).astype(np.float32)
    oscale = (0.5 + rng.random(oc.OUT)).astype(np.float32)
    return x, t, shift, scale, oscale


@pytest.mark.parametrize("n", [1, 64, 150])
def test_forward_matches_oracle(Closure, n):
    x, _, shift, scale, oscale = _problem(n, 10 + n)
    p = oc.init_params(5)
    net = Closure(256)
    net.put_params(p)
    net.set_scaling(shift, scale, oscale)
    y = net.forward(x)
    want, _ = oc.forward(p, x, shift, scale, oscale)
    err = np.abs(y - want).max() / np.abs(want).max()
    print("forward n=%d normwise error %.2e" % (n, err))
    assert err < 1e-2
    assert np.allclose(net.get_params(), p)
    assert all(np.allclose(a, b) for a, b in zip(net.get_scaling(), (shift, scale, oscale)))


def test_large_batch_inference_kernel_equals_small_batch_kernel(Closure):
    """Batches from 2048 samples run the 64-sample LDS-staged kernel; its
    outputs equal, bit for bit, those of the 16-sample kernel that smaller
    batches run (the same matrix-instruction chain per output), and match
    the oracle."""
    n = 5000
    x, _, shift, scale, oscale = _problem(n, 77)
    p = oc.init_params(9)
    net = Closure(n)
    net.put_params(p)
    net.set_scaling(shift, scale, oscale)
    big = net.forward(x)
    small = np.concatenate([net.forward(x[i:i + 1000]) for i in range(0, n, 1000)])
    assert np.array_equal(big, small)
    want, _ = oc.forward(p, x, shift, scale, oscale)
    assert np.abs(big - want).max() / np.abs(want).max() < 1e-2


def test_gradient_matches_oracle(Closure):
    n = 150
    x, t, shift, scale, oscale = _problem(n, 3)
    p = oc.init_params(7)
    net = Closure(256)
    net.put_params(p)
    net.set_scaling(shift, scale, oscale)
    loss, g = net.gradient(x, t)
    lw, gw = oc.loss_and_grad(p, x, t, shift, scale, oscale)
    assert abs(loss - lw) < 1e-3 * lw
    for l in range(4):
        for off, size in ((oc.OFF_W[l], oc.W * oc.NOUT[l]), (oc.OFF_B[l], oc.NOUT[l])):
            a, b = g[off:off + size], gw[off:off + size]
            err = np.abs(a - b).max() / np.abs(b).max()
            print("layer %d block at %d: normwise gradient error %.2e" % (l + 1, off, err))
            assert err < 2e-2
    # padding stays without gradient
    W1 = g[:oc.W * oc.W].reshape(oc.W, oc.W)
    assert np.all(W1[oc.IN:] == 0)


def test_adam_step_matches_oracle(Closure):
    n = 100
    x, t, shift, scale, oscale = _problem(n, 4)
    p = oc.init_params(8)
    net = Closure(128)
    net.put_params(p)
    net.set_scaling(shift, scale, oscale)
    _, g = net.gradient(x, t)
    net.set_data(x, t)
    net.train(1, n, lr=1e-3)
    z = np.zeros_like(p)
    want, _, _ = oc.adam(p, z, z, g, 1, 1e-3)
    got = net.get_params()
    assert np.abs(got - want).max() < 1e-6 * max(1.0, np.abs(p).max())
    assert np.abs(got - p).max() > 5e-4


def test_training_fits_a_smooth_map_and_is_deterministic(Closure):
    rng = np.random.default_rng(11)
    n = 2048
    x = rng.standard_normal((n, oc.IN)).astype(np.float32)
    A = rng.standard_normal((oc.IN, oc.OUT)) / np.sqrt(oc.IN)
    t = (np.sin(x @ A) + 0.5 * (x @ A) ** 2 / 4).astype(np.float32)
    runs = []
    for _ in range(2):
        net = Closure(512, seed=3)
        net.set_scaling(np.zeros(oc.IN), np.ones(oc.IN), np.ones(oc.OUT))
        net.set_data(x, t)
        loss = net.train(600, 256, lr=1e-3)
        runs.append((loss, net.get_params()))
    loss = runs[0][0]
    print("loss first %.4f last %.5f" % (loss[:8].mean(), loss[-8:].mean()))
    assert loss[-8:].mean() < 0.1 * loss[:8].mean()
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])


def test_rejections(Closure):
    from kapsel_old_amd.spm import SpmError
    net = Closure(64)
    with pytest.raises(SpmError):
        net.forward(np.zeros((65, oc.IN), np.float32))
    p = oc.init_params(1)
    p[oc.OFF_W[0] + oc.IN * oc.W] = 1.0  # a padding row of W1
    with pytest.raises(SpmError):
        net.put_params(p)
    with pytest.raises(SpmError):
        net.set_scaling(np.zeros(oc.IN), np.ones(oc.IN), np.zeros(oc.OUT))


# ---------------------------------------------------------------------
# coupling to the SPM step
# ---------------------------------------------------------------------
@pytest.fixture(scope="module")
def SPM():
    from kapsel_old_amd.spm import SPM as _S
    return _S


def _trilinear(field, pts):
    """Periodic trilinear interpolation of a real [n][n][n] field at pts."""
    n = field.shape[0]
    out = np.zeros(len(pts))
    for k, x in enumerate(pts):
        i0 = np.floor(x).astype(int)
        f = x - i0
        for c in range(8):
            cb = np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1])
            w = np.prod(np.where(cb == 1, f, 1 - f))
            i = (i0 + cb) % n
            out[k] += w * field[i[0], i[1], i[2]]
    return out


def test_features_match_numpy(SPM):
    """spm_closure_features against numpy: trilinear u at x_p + a o, phi
    there as the clamped sum of the particles' profiles (the oracle's Phi),
    the slip velocity from f_hydro, and the nearest-neighbour gap vector."""
    from kapsel_old_amd import abi
    from oracle.oracle import Oracle
    from tests.helpers import particles, smooth_field
    n, a, npart = 32, 3.0, 10
    prm = abi.make_params(n, a=a, n_particles=npart, rotation=1)
    p = particles(npart, (n, n, n), a, seed=5, vel=0.02, omega=0.005, min_factor=2.2)
    g, o = SPM(prm), Oracle(prm)
    for obj in (g, o):
        obj.put_particles(p)
    for d in range(3):
        u = smooth_field(o.shape, n, seed=30 + d, kmax=3, scale=0.05)
        g.put_field("u%d" % d, u)
        o.put_field("u%d" % d, u)
    g.Calc_f_hydro_correct_precision()
    x = g.closure_features()
    q = g.get_particles()
    offs = np.array([(k // 9 - 1, (k // 3) % 3 - 1, k % 3 - 1) for k in range(27)])
    us = [g.get_field("u%d" % d)[:, :, :n] for d in range(3)]
    for m in range(npart):
        pts = (q["x"][m] + a * offs) % n
        for d in range(3):
            want = _trilinear(us[d], pts) - q["v"][m][d]
            assert np.abs(x[m, 3 + 3 * np.arange(27) + d] - want).max() < 1e-6
        phi = []
        for pt in pts:
            r = q["x"] - pt
            r -= n * np.rint(r / n)
            phi.append(min(1.0, sum(o.phi_profile(d) for d in np.linalg.norm(r, axis=1))))
        assert np.abs(x[m, 84:111] - np.array(phi)).max() < 1e-6
        slip = -q["f_hydro"][m] * prm.dt / (prm.rho * 4.0 / 3.0 * np.pi * a ** 3)
        assert np.allclose(x[m, :3], slip, rtol=1e-6, atol=1e-9)
        r = q["x"] - q["x"][m]
        r -= n * np.rint(r / n)
        dist = np.linalg.norm(r, axis=1)
        dist[m] = np.inf
        k, k2 = [int(i) for i in np.argsort(dist, kind="stable")[:2]]
        assert np.allclose(x[m, 111:114], r[k] * (1 - 2 * a / dist[k]), rtol=1e-6, atol=1e-6)
        assert np.allclose(x[m, 114:117], q["v"][k] - q["v"][m], rtol=1e-6, atol=1e-9)
        assert np.allclose(x[m, 117:120], q["omega"][k] - q["omega"][m], rtol=1e-6, atol=1e-9)
        assert np.allclose(x[m, 120:123], r[k2] * (1 - 2 * a / dist[k2]), rtol=1e-6, atol=1e-6)
        assert np.allclose(x[m, 123:126], q["v"][k2] - q["v"][m], rtol=1e-6, atol=1e-9)


def test_hook_adds_the_network_output(SPM, Closure):
    """A network whose only nonzero parameters are the last bias outputs a
    constant c: with it attached, f_hydro and torque_hydro of a step exceed
    those of the same step without it by c exactly (to rounding), and the
    fused and literal schedules agree."""
    from kapsel_old_amd import abi
    from tests.helpers import particles
    n, a, npart = 32, 3.0, 12
    prm = abi.make_params(n, a=a, n_particles=npart, rotation=1, kbt=0.0)
    p = particles(npart, (n, n, n), a, seed=8, vel=0.02, omega=0.005, min_factor=2.2)
    c = np.array([0.01, -0.02, 0.03, 0.004, -0.005, 0.006], np.float32)
    par = np.zeros(oc.NPARAM, np.float32)
    par[oc.OFF_B[3]:oc.OFF_B[3] + 6] = c
    net = Closure(64)
    net.put_params(par)
    out = {}
    for name, attach, sched in (("plain", False, 0), ("closure", True, 0), ("closure_lit", True, 1)):
        g = SPM(prm)
        g.set_schedule(sched)
        g.put_particles(p)
        g.init_fluid()
        g.set_particle_motion(True)
        if attach:
            g.set_closure(net)
        g.step(2)
        q = g.get_particles()
        out[name] = (q["f_hydro"].copy(), q["torque_hydro"].copy(), q["x"].copy(), q["v"].copy())
        g.close()
    f0, t0 = out["plain"][:2]
    f1, t1 = out["closure"][:2]
    assert np.abs(f1 - f0 - c[:3]).max() < 1e-12 and np.abs(t1 - t0 - c[3:]).max() < 1e-12
    assert np.abs(out["closure_lit"][0] - f1).max() < 1e-10 * np.abs(f1).max()
    # prescribed motion: positions and velocities held
    assert np.array_equal(out["closure"][2], p["x"]) and np.array_equal(out["closure"][3], p["v"])


def test_closure_lowers_coarse_drag_error(SPM, Closure):
    """End to end at a small size: coarse 32^3 (a = 3) and fine 128^3 (a =
    12) resistance-problem pairs of random suspensions at volume fractions
    0.1 to 0.3 train the network (48 lattice symmetries, Adam on the GPU);
    on a held-out suspension at 0.2 the coarse drag error against the fine
    judge drops to less than half with the network."""
    from kapsel_old_amd import closure as cl
    cfg = [(phi, 40 + k) for k in range(2) for phi in (0.1, 0.2, 0.3)]
    x, t, fine, coarse = cl.dataset(cfg, n_grid=32, steps=400)
    xa, ta = cl.augment(x, t)
    net, loss = cl.train_closure(xa, ta, steps=4000, batch=512, lr=1e-3, max_batch=1024)
    xe, te, fe, ce = cl.dataset([(0.2, 77)], n_grid=32, steps=400)
    y = net.forward(xe)
    e0 = np.median(cl.drag_error(ce[:, :3], fe[:, :3]))
    e1 = np.median(cl.drag_error(ce[:, :3] + y[:, :3], fe[:, :3]))
    print("32^3 vs 128^3: %d training samples, median drag error %.4f without, %.4f with the closure"
          % (len(x), e0, e1))
    assert e1 < 0.5 * e0


def test_features_turn_with_the_lattice_symmetries(SPM):
    """The inputs of a configuration mapped by a lattice symmetry S (positions
    S x mod L, velocities S v, angular velocities det(S) S w, the flow
    u'(S x) = S u(x)) equal closure.transform of the original inputs: the
    permutation of the 27 samples, vectors turned with S and the relative
    angular velocity (a pseudovector) with det(S) S.  This pins the
    augmentation that multiplies the training set by 48."""
    from kapsel_old_amd import abi
    from kapsel_old_amd import closure as cl
    from tests.helpers import particles, smooth_field
    n, a, npart = 32, 3.0, 10
    prm = abi.make_params(n, a=a, n_particles=npart, rotation=1)
    p = particles(npart, (n, n, n), a, seed=8, vel=0.02, omega=0.005, min_factor=2.2)
    g = SPM(prm)
    shape = (n, n, g.get_field("u0").shape[2])
    us = [smooth_field(shape, n, seed=60 + d, kmax=3, scale=0.05) for d in range(3)]
    S = np.array([[0, 1, 0], [0, 0, -1], [-1, 0, 0]], dtype=float)
    det = np.rint(np.linalg.det(S))

    def features(parts, fields):
        h = SPM(prm)
        h.put_particles(parts)
        for d in range(3):
            h.put_field("u%d" % d, fields[d])
        h.Calc_f_hydro_correct_precision()
        x = h.closure_features()
        h.close()
        return x

    x0 = features(p, us)
    # the mapped configuration
    q = p.copy()
    q["x"] = (p["x"] @ S.T) % n
    q["x_nopbc"] = q["x"]
    q["v"] = p["v"] @ S.T
    q["omega"] = det * (p["omega"] @ S.T)
    coords = np.indices((n, n, n)).reshape(3, -1)
    moved = ((S @ coords) % n).astype(int)
    U = np.stack([u[:, :, :n].reshape(-1) for u in us])
    U2 = S @ U
    vs = []
    for d in range(3):
        v = np.zeros(shape)
        v[moved[0], moved[1], moved[2]] = U2[d]
        vs.append(v)
    x1 = features(q, vs)
    xt, _ = cl.transform(x0, np.zeros((npart, 6), np.float32), S)
    err = np.abs(x1 - xt).max() / np.abs(xt).max()
    print("inputs of the mapped configuration vs transformed inputs: rel err %.2e" % err)
    assert err < 1e-5
