"""Pins of the oracle's derivatives (forward-mode dual numbers, FP64).

* The RHS Jacobian d a^mu / d(r, theta, v, spin) against central finite
  differences of the oracle's own geodesic_acceleration.
* The loss gradient dL/d(spin, inclination, q, r_in) of the differentiable
  thin-disk render against central finite differences (h = 1e-5) of the loss
  with every ray's step sizes and accept/reject pattern replayed from the
  unperturbed run (the discrete map the dual numbers differentiate).
"""
import math

import numpy as np
import pytest

import gargantua_amd as gg
from gargantua_amd import configs


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


def test_dual_gradient_matches_frozen_step_central_differences(oracle):
    target_scene = _small_grad_scene(spin=0.7, inclination_deg=60.0, q=2.5, r_in=configs.r_isco(1.0, 0.7))
    target = oracle.grad(target_scene, np.zeros(24 * 24))["intensity"]
    g0 = _small_grad_scene()
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
    return lib


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
