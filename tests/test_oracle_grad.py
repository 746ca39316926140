"""Pins of the oracle's derivatives (forward-mode dual numbers, FP64).

* The RHS Jacobian d a^mu / d(r, theta, v, spin) against central finite
  differences of the oracle's own geodesic_acceleration.
* The loss gradient dL/d(spin, inclination, q, r_in) of the differentiable
  thin-disk render, with the step sizes carrying derivatives through the
  error estimate and the controller, against central finite differences of
  the loss the render computes (h = 1e-6 over every ray; h = 1e-5 over the
  rays whose step decisions do not flip between theta - h and theta + h).
* With frozen_steps set, the gradient of the discrete map with every ray's
  step sizes and accept/reject pattern held at the unperturbed run's values
  against central finite differences (h = 1e-5) of the loss replayed with
  that step sequence.
"""
import math

import numpy as np
import pytest

import gargantua_amd as gg
from gargantua_amd import configs

from .gradcheck import central_fd, masked_gradient


def _points(rng, n):
    r = np.exp(rng.uniform(math.log(3.0), math.log(200.0), n))
    th = rng.uniform(0.2, math.pi - 0.2, n)
    pos = np.stack([rng.uniform(0, 10, n), r, th, rng.uniform(0, 2 * math.pi, n)], 1)
    vel = rng.normal(size=(n, 4)) * np.array([1.0, 1.0, 0.05, 0.05])
    return pos, vel


@pytest.mark.parametrize("spin", [0.0, 0.5, 0.9])
def test_rhs_jacobian_matches_central_differences(oracle, spin):
    rng = np.random.default_rng(11)
    pos, vel = _points(rng, 64)
    metric = gg.Kerr(1.0, spin)
    J = oracle.rhs_jacobian(metric, pos, vel)
    fd = np.zeros_like(J)
    for k, (col, h) in enumerate([(1, 1e-6), (2, 1e-7)]):
        dp = np.zeros(4)
        dp[col] = 1.0
        hs = h * (pos[:, col:col + 1] if col == 1 else 1.0)
        a_p = oracle.geodesic_acceleration(metric, pos + hs * dp, vel)
        a_m = oracle.geodesic_acceleration(metric, pos - hs * dp, vel)
        fd[:, :, k] = (a_p - a_m) / (2 * hs)
    for c in range(4):
        dv = np.zeros(4)
        dv[c] = 1.0
        h = 1e-6
        a_p = oracle.geodesic_acceleration(metric, pos, vel + h * dv)
        a_m = oracle.geodesic_acceleration(metric, pos, vel - h * dv)
        fd[:, :, 2 + c] = (a_p - a_m) / (2 * h)
    h = 1e-6
    fd[:, :, 6] = (oracle.geodesic_acceleration(gg.Kerr(1.0, spin + h), pos, vel) -
                   oracle.geodesic_acceleration(gg.Kerr(1.0, spin - h), pos, vel)) / (2 * h)
    scale = np.max(np.abs(J), axis=(1, 2), keepdims=True)
    err = np.abs(J - fd) / scale
    print(f"a={spin}: max rel err {err.max():.2e}")
    assert err.max() < 1e-6
    # a = -Gamma v v is quadratic in v: J_v v = 2 a (Euler's theorem on homogeneous functions)
    acc = oracle.geodesic_acceleration(metric, pos, vel)
    np.testing.assert_allclose(np.einsum("nmk,nk->nm", J[:, :, 2:6], vel), 2 * acc, rtol=1e-10,
                               atol=1e-12 * np.abs(acc).max())


def _small_grad_scene(n=24, **kw):
    g = configs.grad_512_params(n, n, **kw)
    return g


def _target24(oracle):
    target_scene = _small_grad_scene(spin=0.7, inclination_deg=60.0, q=2.5, r_in=configs.r_isco(1.0, 0.7))
    return oracle.grad(target_scene, np.zeros(24 * 24))["intensity"]


def test_dual_gradient_matches_central_differences_of_the_loss(oracle):
    target = _target24(oracle)
    g0 = _small_grad_scene()
    out = oracle.grad(g0, target, want_dI=True)
    # h = 1e-6: no decision flips on this scene; every ray enters
    fd6 = central_fd(oracle, _small_grad_scene, target, 1e-6)
    rel6 = np.abs(out["grad"] - fd6["fd"]) / np.abs(fd6["fd"])
    print("dual", out["grad"], "fd(1e-6)", fd6["fd"], "rel", rel6, "flips", fd6["flips"])
    assert np.all(rel6 < 1e-4)
    # h = 1e-5: the rays whose step decisions stay put
    fd5 = central_fd(oracle, _small_grad_scene, target, 1e-5)
    ad5 = masked_gradient(out["dI"], out["intensity"], target, fd5["mask"])
    rel5 = np.abs(ad5 - fd5["fd_smooth"]) / np.abs(fd5["fd_smooth"])
    print("fd(1e-5) smooth", fd5["fd_smooth"], "rel", rel5, "flips", fd5["flips"])
    assert np.all(rel5 < 1e-4)
    assert np.all(fd5["flips"] <= 0.01 * 24 * 24)
    # the step sizes' derivative is not negligible: holding them fixed misses the loss's slope
    g0.frozen_steps = 1
    frozen = oracle.grad(g0, target)
    assert frozen["loss"] == out["loss"]
    assert np.abs(frozen["grad"][:2] - fd6["fd"][:2]).max() > 1e-3 * np.abs(fd6["fd"][:2]).max()


def test_dual_gradient_matches_frozen_step_central_differences(oracle):
    target = _target24(oracle)
    g0 = _small_grad_scene()
    g0.frozen_steps = 1
    out = oracle.grad(g0, target)
    assert out["loss"] > 0.0
    fields = ["spin", "inclination", "emissivity_q", "r_inner"]
    h = 1e-5
    fd = np.zeros(4)
    for k, f in enumerate(fields):
        gp, gm = _small_grad_scene(), _small_grad_scene()
        setattr(gp, f, getattr(g0, f) + h)
        setattr(gm, f, getattr(g0, f) - h)
        lp, _ = oracle.loss_frozen(g0, gp, target)
        lm, _ = oracle.loss_frozen(g0, gm, target)
        fd[k] = (lp - lm) / (2 * h)
    rel = np.abs(out["grad"] - fd) / np.abs(fd)
    print("dual", out["grad"], "fd", fd, "rel", rel)
    assert np.all(rel < 1e-4)
    # frozen replay with unperturbed parameters reproduces the plain loss exactly
    l0, _ = oracle.loss_frozen(g0, g0, target)
    assert l0 == out["loss"]


def test_gradient_vanishes_at_the_target(oracle):
    g = _small_grad_scene()
    target = oracle.grad(g, np.zeros(24 * 24))["intensity"]
    out = oracle.grad(g, target)
    assert out["loss"] == 0.0
    assert np.all(out["grad"] == 0.0)


@pytest.fixture(scope="module")
def host_vjp(tmp_path_factory):
    """The device VJP source compiled for the host with g++."""
    import ctypes as C
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    so = str(tmp_path_factory.mktemp("vjp") / "libhostvjp.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I", os.path.join(root, "gargantua_amd", "csrc"),
                    "-o", so, os.path.join(root, "tests", "host_vjp.cpp")], check=True)
    lib = C.CDLL(so)
    lib.host_kerr_rhs_vjp.argtypes = [C.c_double, C.c_double, C.c_uint64] + [C.c_void_p] * 4
    lib.host_kerr_rhs_vjp.restype = None
    lib.host_sincos_near.argtypes = [C.c_double, C.c_double, C.c_double, C.c_void_p]
    lib.host_sincos_near.restype = C.c_int
    return lib


def test_sincos_near_matches_libm(host_vjp):
    """The adjoint's stage angles: sin and cos of theta0 + d by angle addition
    with the Taylor series of sin d, cos d, against libm at theta0 + d (d
    within 1/16), and the refusal beyond."""
    rng = np.random.default_rng(5)
    out = np.zeros(2)
    worst = 0.0
    for _ in range(20000):
        th0 = rng.uniform(0.0, math.pi)
        d = rng.uniform(-0.0625, 0.0625) * rng.choice([1.0, 1e-3, 1e-8])
        assert host_vjp.host_sincos_near(math.sin(th0), math.cos(th0), d, out.ctypes.data) == 1
        worst = max(worst, abs(out[0] - math.sin(th0 + d)), abs(out[1] - math.cos(th0 + d)))
    print(f"max abs error {worst:.2e}")
    assert worst < 5e-16
    assert host_vjp.host_sincos_near(math.sin(1.0), math.cos(1.0), 0.0, out.ctypes.data) == 1
    assert out[0] == math.sin(1.0) and out[1] == math.cos(1.0)
    for d in (0.0626, -0.07, 1.0, float("nan")):
        assert host_vjp.host_sincos_near(0.5, 0.8, d, out.ctypes.data) == 0


@pytest.mark.parametrize("spin", [0.0, 0.3, 0.9, 0.998])
def test_hand_reverse_vjp_matches_dual_jacobian(oracle, host_vjp, spin):
    """The adjoint kernel's reverse-mode RHS VJP (host build of the same
    source) equals w^T J of the oracle's forward-mode Jacobian."""
    rng = np.random.default_rng(21)
    n = 4096
    r = np.exp(rng.uniform(math.log(1.5), math.log(1000.0), n))
    th = rng.uniform(0.02, math.pi - 0.02, n)
    pos = np.stack([np.zeros(n), r, th, np.zeros(n)], 1)
    vel = rng.normal(size=(n, 4)) * np.array([1.0, 1.0, 0.05, 0.05])
    w = rng.normal(size=(n, 4))
    metric = gg.Kerr(1.0, spin)
    out = np.zeros((n, 7))
    host_vjp.host_kerr_rhs_vjp(1.0, spin, n, pos.ctypes.data, vel.ctypes.data, w.ctypes.data, out.ctypes.data)
    J = oracle.rhs_jacobian(metric, pos, vel)
    want = np.einsum("nm,nmk->nk", w, J)
    scale = np.max(np.abs(J), axis=(1, 2)) * np.max(np.abs(w), axis=1)
    err = np.max(np.abs(out - want), axis=1) / scale
    print(f"a={spin}: max rel {err.max():.2e}")
    assert err.max() < 1e-11
