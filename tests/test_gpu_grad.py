"""GPU adjoint (gg_grad, gg_geodesic_acceleration_vjp) against the oracle.

* VJP of the Kerr RHS vs w^T J of the oracle's dual-number Jacobian.
* gg_grad's loss, gradient and per-pixel dI vs the oracle's forward-mode
  dual-number gradient (step sizes differentiated through the controller,
  and with frozen_steps held fixed).
* gg_grad vs central finite differences of the loss the render computes
  (h = 1e-6 over every ray, h = 1e-5 over the rays whose step decisions do
  not flip between theta - h and theta + h): 1e-4 relative.  With
  frozen_steps, vs central differences of the oracle's frozen-step loss.
* The result does not depend on the checkpoint interval K (bitwise: the
  recomputed segments repeat the forward steps exactly).
"""
import math

import numpy as np
import pytest

import gargantua_amd as gg
from gargantua_amd import configs

from .gradcheck import central_fd, masked_gradient

pytestmark = pytest.mark.gpu

N = 40


def _scene(n=N, **kw):
    return configs.grad_512_params(n, n, **kw)


def _target(oracle, n=N):
    t = _scene(n, spin=0.7, inclination_deg=60.0, q=2.5, r_in=configs.r_isco(1.0, 0.7))
    return oracle.grad(t, np.zeros(n * n))["intensity"]


@pytest.mark.parametrize("spin", [0.0, 0.5, 0.9, 0.998])
def test_rhs_vjp_matches_oracle_jacobian(renderer, oracle, spin):
    rng = np.random.default_rng(5)
    n = 2048
    r = np.exp(rng.uniform(math.log(2.5), math.log(500.0), n))
    th = rng.uniform(0.05, math.pi - 0.05, n)
    pos = np.stack([rng.uniform(0, 10, n), r, th, rng.uniform(0, 2 * math.pi, n)], 1)
    vel = rng.normal(size=(n, 4)) * np.array([1.0, 1.0, 0.05, 0.05])
    w = rng.normal(size=(n, 4))
    metric = gg.Kerr(1.0, spin)
    g = renderer.geodesic_acceleration_vjp(metric, pos, vel, w)
    J = oracle.rhs_jacobian(metric, pos, vel)
    want = np.einsum("nm,nmk->nk", w, J)   # (r, th, v0..v3, a)
    got = np.concatenate([g["pos_bar"][:, 1:3], g["vel_bar"], g["spin_bar"][:, None]], 1)
    scale = np.max(np.abs(J), axis=(1, 2)) * np.max(np.abs(w), axis=1)
    err = np.max(np.abs(got - want), axis=1) / scale
    print(f"a={spin}: vjp max rel {err.max():.2e}")
    assert err.max() < 1e-11
    assert np.all(g["pos_bar"][:, [0, 3]] == 0.0)


@pytest.mark.parametrize("frozen", [0, 1])
def test_grad_matches_oracle_dual_gradient(renderer, oracle, frozen):
    target = _target(oracle)
    g0 = _scene()
    g0.frozen_steps = frozen
    o = oracle.grad(g0, target, want_dI=True)
    a = renderer.grad(g0, target, checkpoint_every=8, want_dI=True)
    st = a["stats"]
    print("stats", st)
    print("gpu", a["grad"], "oracle", o["grad"])
    disk = a["intensity"] > 0
    # disk-class rays include hits whose redshift factor is 0 (I = 0)
    assert st["disk_rays"] >= int(np.count_nonzero(disk))
    assert abs(int(np.count_nonzero(disk)) - int(np.count_nonzero(o["intensity"] > 0))) <= 2
    np.testing.assert_allclose(a["intensity"], o["intensity"], rtol=1e-6, atol=1e-12 * o["intensity"].max())
    assert abs(a["loss"] - o["loss"]) <= 1e-6 * o["loss"]
    rel = np.abs(a["grad"] - o["grad"]) / np.abs(o["grad"])
    print("rel", rel)
    assert np.all(rel < 1e-4)
    # per-pixel tangents where both traced the same ray to the disk
    both = disk & (o["intensity"] > 0)
    scale = np.abs(o["dI"][both]).max(axis=0)
    err = np.abs(a["dI"][both] - o["dI"][both]) / scale
    close = np.all(err < 1e-6, axis=1)
    print(f"dI within 1e-6 of the column scale on {close.mean():.4f} of {both.sum()} disk rays")
    assert close.mean() >= 0.99
    assert np.all(a["dI"][~disk] == 0.0)


@pytest.mark.parametrize("checkpoint_every", [0, 32])
def test_grad_matches_central_differences_of_the_loss(renderer, oracle, checkpoint_every):
    """The gradient of the loss the render computes, step sizes included, against
    central differences of the oracle's loss on the 40 x 40 gradient scene."""
    target = _target(oracle)
    a = renderer.grad(_scene(), target, checkpoint_every=checkpoint_every, want_dI=True)
    fd6 = central_fd(oracle, _scene, target, 1e-6)
    rel6 = np.abs(a["grad"] - fd6["fd"]) / np.abs(fd6["fd"])
    print("adjoint", a["grad"], "fd(1e-6)", fd6["fd"], "rel", rel6, "flips", fd6["flips"])
    assert np.all(rel6 < 1e-4)
    fd5 = central_fd(oracle, _scene, target, 1e-5)
    ad5 = masked_gradient(a["dI"], a["intensity"], target, fd5["mask"])
    rel5 = np.abs(ad5 - fd5["fd_smooth"]) / np.abs(fd5["fd_smooth"])
    print("fd(1e-5) over smooth rays", fd5["fd_smooth"], "rel", rel5, "flips", fd5["flips"])
    assert np.all(rel5 < 1e-4)
    assert np.all(fd5["flips"] <= 0.01 * N * N)


def test_grad_matches_frozen_step_central_differences(renderer, oracle):
    target = _target(oracle)
    g0 = _scene()
    g0.frozen_steps = 1
    a = renderer.grad(g0, target, checkpoint_every=32)
    fields = ["spin", "inclination", "emissivity_q", "r_inner"]
    h = 1e-5
    fd = np.zeros(4)
    for k, f in enumerate(fields):
        gp, gm = _scene(), _scene()
        setattr(gp, f, getattr(g0, f) + h)
        setattr(gm, f, getattr(g0, f) - h)
        fd[k] = (oracle.loss_frozen(g0, gp, target)[0] - oracle.loss_frozen(g0, gm, target)[0]) / (2 * h)
    rel = np.abs(a["grad"] - fd) / np.abs(fd)
    print("adjoint", a["grad"], "fd", fd, "rel", rel)
    assert np.all(rel < 1e-4)


@pytest.mark.parametrize("frozen", [0, 1])
def test_grad_independent_of_checkpoint_interval(renderer, oracle, frozen):
    target = _target(oracle)
    g0 = _scene()
    g0.frozen_steps = frozen
    ref = renderer.grad(g0, target, checkpoint_every=1, want_dI=True)
    assert ref["stats"]["mode"] == 0
    for K in (3, 8, 32, 128, 100000 // 2, 0):
        a = renderer.grad(g0, target, checkpoint_every=K, want_dI=True)
        _assert_same(a, ref, K)
        st = a["stats"]
        if K == 0:
            # the render recorded every step: no checkpoint, no re-trace
            assert st["mode"] == 2
            assert st["checkpoints"] == 0 and st["attempts"] == 0 and st["segment_bytes"] == 0
            assert st["record_bytes"] % 4224 == 0 and st["record_bytes"] >= 256 * st["steps"]
        elif K == 100000 // 2:
            # one segment per ray: the re-trace stores every step record
            assert st["mode"] == 1
            assert st["checkpoints"] == st["disk_rays"]
            assert st["record_bytes"] == 256 * st["steps"]
            assert st["segment_bytes"] == 0
        else:
            assert st["checkpoint_every"] == K
            assert st["checkpoints"] <= ref["stats"]["checkpoints"]


def _assert_same(a, ref, label):
    assert a["loss"] == ref["loss"], label
    assert np.array_equal(a["grad"], ref["grad"]), label
    assert np.array_equal(a["dI"], ref["dI"]), label
    assert a["stats"]["steps"] == ref["stats"]["steps"], label


def test_grad_automatic_mode_follows_the_memory_budget(renderer, oracle):
    """checkpoint_every = 0 under budgets that fit the render's records (mode
    2), only the disk rays' records (mode 1), or neither (mode 0, K = 32):
    the same gradient bit for bit."""
    target = _target(oracle)
    g0 = _scene()
    ref = renderer.grad(g0, target, checkpoint_every=1, want_dI=True)
    try:
        a2 = renderer.grad(g0, target, checkpoint_every=0, want_dI=True)
        assert a2["stats"]["mode"] == 2
        need1 = 256 * ref["stats"]["steps"] + 72 * ref["stats"]["disk_rays"]
        assert a2["stats"]["record_bytes"] > need1
        renderer.set_grad_memory(need1)
        a1 = renderer.grad(g0, target, checkpoint_every=0, want_dI=True)
        assert a1["stats"]["mode"] == 1
        renderer.set_grad_memory(1)
        a0 = renderer.grad(g0, target, checkpoint_every=0, want_dI=True)
        assert a0["stats"]["mode"] == 0 and a0["stats"]["checkpoint_every"] == 32
        for lab, a in (("mode 2", a2), ("mode 1", a1), ("mode 0", a0)):
            _assert_same(a, ref, lab)
    finally:
        renderer.set_grad_memory(0)


class _DeviceArray:
    """float64 values copied to device memory with the HIP runtime the library uses."""

    dtype = "float64"

    def __init__(self, a):
        import ctypes as C
        self.hip = C.CDLL("libamdhip64.so.7")
        self.n = a.size
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(a.nbytes)) == 0
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def data_ptr(self):
        return self.p.value

    def numel(self):
        return self.n

    def is_contiguous(self):
        return True

    def free(self):
        self.hip.hipFree(self.p)


def test_grad_learned_order_and_device_target(renderer, oracle):
    """The render under the learned ray order (queued by the previous call's
    attempt counts) and a target read in place from device memory give the
    same gradient bit for bit."""
    target = _target(oracle)
    g0 = _scene()
    ref = renderer.grad(g0, target, checkpoint_every=1, want_dI=True)
    t_dev = _DeviceArray(np.ascontiguousarray(target, dtype=np.float64))
    try:
        renderer.set_launch(order=1)
        for _ in range(2):   # the second call runs in the learned order
            a = renderer.grad(g0, t_dev, checkpoint_every=0, want_dI=True, want_image=False)
            assert a["stats"]["mode"] == 2 and a["intensity"] is None
            _assert_same(a, ref, "learned order")
    finally:
        renderer.set_launch()
        t_dev.free()


def test_grad_mixed_forward(renderer, oracle):
    """The mixed-precision forward render (FP32 stages where the trace
    kernel's policy allows, the render's own records for the reverse sweep):
    loss and gradient within 1e-6 of the FP64 forward's; it needs the
    recorded mode, so an explicit checkpoint interval is rejected."""
    target = _target(oracle)
    a = renderer.grad(_scene(), target, want_dI=True)
    b = renderer.grad(_scene(precision="mixed"), target, want_dI=True)
    assert b["stats"]["mode"] == 2
    rel = np.abs(b["grad"] - a["grad"]) / np.abs(a["grad"])
    print("fp64", a["grad"], "mixed", b["grad"], "rel", rel)
    assert np.all(rel < 1e-6)
    assert abs(b["loss"] - a["loss"]) <= 1e-6 * a["loss"]
    with pytest.raises(gg.GargantuaError):
        renderer.grad(_scene(precision="mixed"), target, checkpoint_every=8)


def test_grad_record_pool_grows(oracle):
    """A fresh context sized by a small scene meets a larger one: the record
    pool overflows, is resized from the count of chunks the render asked for,
    and the render runs again; the gradient equals a fresh context's."""
    import gargantua_amd as gg
    target = _target(oracle)
    g0 = _scene()
    r1 = gg.Renderer(0)
    r2 = gg.Renderer(0)
    try:
        small = configs.grad_512_params(8, 8)
        r1.grad(small, np.zeros(64), checkpoint_every=0)
        a = r1.grad(g0, target, checkpoint_every=0, want_dI=True)
        b = r2.grad(g0, target, checkpoint_every=0, want_dI=True)
        assert a["stats"]["mode"] == 2 and b["stats"]["mode"] == 2
        _assert_same(a, b, "grown pool")
    finally:
        r1.close()
        r2.close()


def test_grad_image_is_the_fp64_render(renderer, oracle):
    g0 = _scene()
    img = renderer.grad_image(g0)
    a = renderer.grad(g0, np.zeros(N * N))
    assert np.array_equal(img, a["intensity"])
    assert a["loss"] == pytest.approx(np.mean(img ** 2), rel=1e-12)


def test_grad_zero_at_target(renderer):
    g0 = _scene()
    img = renderer.grad_image(g0)
    a = renderer.grad(g0, img)
    assert a["loss"] == 0.0
    assert np.all(a["grad"] == 0.0)


def test_grad_rejects_bad_arguments(renderer):
    g0 = _scene()
    with pytest.raises(gg.GargantuaError):
        renderer.grad(g0, np.zeros(N * N), checkpoint_every=65537)
    bad = _scene()
    bad.spin = 1.0
    with pytest.raises(gg.GargantuaError):
        renderer.grad(bad, np.zeros(N * N))
    with pytest.raises(gg.GargantuaError):
        renderer.geodesic_acceleration_vjp(gg.Schwarzschild(1.0), np.ones((1, 4)), np.ones((1, 4)), np.ones((1, 4)))


def test_gradient_descent_recovers_parameters(renderer):
    """200 Adam steps from theta0 = (0.5, 45 deg, 3.0, r_ISCO + 0.5) on the
    image at (0.7, 60 deg, 2.5, r_ISCO(0.7)), mixed forward render: the loss
    falls by orders of magnitude and the well-determined parameters
    (inclination, q) approach their true values closely.  Spin and r_in
    move the image in nearly the same way and converge last."""
    n = 96
    tgt = _scene(n, spin=0.7, inclination_deg=60.0, q=2.5, r_in=configs.r_isco(1.0, 0.7))
    truth = np.array([tgt.spin, tgt.inclination, tgt.emissivity_q, tgt.r_inner])
    target = renderer.grad_image(tgt)
    g0 = _scene(n, precision="mixed")
    theta0 = np.array([g0.spin, g0.inclination, g0.emissivity_q, g0.r_inner])
    fit = renderer.fit(g0, target, 200, log=lambda it, loss, th: print(f"step {it}: loss {loss:.3e} theta {th}"))
    theta = fit["theta"][-1]
    err0, err = np.abs(theta0 - truth), np.abs(theta - truth)
    print("loss", fit["loss"][0], "->", fit["loss"][-1], "|theta - truth|", err0, "->", err)
    # measured: loss 3.3e-5 -> 1.5e-9; |theta - truth| (0.2, 0.26, 0.5, 1.34) -> (0.031, 0.0030, 0.0007, 0.021)
    assert fit["loss"][-1] < 1e-3 * fit["loss"][0]
    assert err[1] < 0.05 * err0[1] and err[2] < 0.01 * err0[2]
    assert err[0] < 0.25 * err0[0] and err[3] < 0.05 * err0[3]
