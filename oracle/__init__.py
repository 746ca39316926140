"""ctypes front end of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's CPU
baseline use it as the checker; the product (gargantua_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from gargantua_amd import _abi
from gargantua_amd import Scene, alloc_outputs, outputs_struct

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liboracle.so")

_D = C.c_double
_P = C.c_void_p
_SIGS = [
    ("oracle_metric_tensor", C.c_int, [C.POINTER(_abi.gg_metric), _P, _P]),
    ("oracle_christoffel", C.c_int, [C.POINTER(_abi.gg_metric), _P, _P]),
    ("oracle_geodesic_acceleration", C.c_int, [C.POINTER(_abi.gg_metric), C.c_uint64, _P, _P, _P]),
    ("oracle_constants_of_motion", C.c_int, [C.POINTER(_abi.gg_metric), C.c_uint64, _P, _P]),
    ("oracle_rhs_jacobian", C.c_int, [C.POINTER(_abi.gg_metric), C.c_uint64, _P, _P, _P]),
    ("oracle_event_horizon", C.c_int, [C.POINTER(_abi.gg_metric), C.POINTER(_D)]),
    ("oracle_orbital_four_velocity", C.c_int, [C.POINTER(_abi.gg_metric), _P, _P]),
    ("oracle_redshift_factor", _D, [C.POINTER(_abi.gg_metric), _P, _P, _D]),
    ("oracle_rk45_step", C.c_int, [C.POINTER(_abi.gg_metric), C.POINTER(_abi.gg_integrator), _P, C.POINTER(_D),
                                   C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("oracle_rk4_step", C.c_int, [C.POINTER(_abi.gg_metric), _D, _D, _D, _P]),
    ("oracle_pixel_direction", None, [C.POINTER(_abi.gg_camera), _D, _D, _P]),
    ("oracle_from_camera", C.c_int, [C.POINTER(_abi.gg_scene), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    ("oracle_disk_density", _D, [C.POINTER(_abi.gg_disk), _P]),
    ("oracle_disk_emission", _D, [C.POINTER(_abi.gg_disk), _P]),
    ("oracle_disk_color", None, [_D, _P]),
    ("oracle_encode_srgb", None, [_P, _P]),
    ("oracle_sky", None, [_P, _P]),
    ("oracle_set_sky_hdri", C.c_int, [C.c_uint32, C.c_uint32, _P, C.c_float]),
    ("oracle_set_hittables", C.c_int, [C.c_uint32, _P]),
    ("oracle_set_user_metric", None, [_P, C.c_int, C.c_double]),
    ("oracle_inverse4", C.c_int, [_P, _P]),
    ("oracle_scene_hit", C.c_int, [_P, _P, _P]),
    ("oracle_g4_static", C.c_double, [_P, _P, C.c_double]),
    ("oracle_hdri_sample", None, [C.c_uint32, C.c_uint32, _P, C.c_float, _P, _P]),
    ("oracle_splitmix64", C.c_uint64, [C.c_uint64, C.c_uint64]),
    ("oracle_render_rays", C.c_int, [C.POINTER(_abi.gg_scene), C.c_uint64, _P, C.POINTER(_abi.gg_outputs), C.c_int]),
    ("oracle_render", C.c_int, [C.POINTER(_abi.gg_scene), C.POINTER(_abi.gg_outputs), C.c_int]),
    ("oracle_trace_states", C.c_int, [C.POINTER(_abi.gg_scene), C.c_uint64, _P, C.POINTER(_abi.gg_outputs),
                                      C.c_int]),
    ("oracle_last_rhs", C.c_uint64, []),
    ("oracle_last_reflections", C.c_uint64, []),
    ("oracle_r_isco", _D, [_D, _D]),
    ("oracle_grad", C.c_int, [C.POINTER(_abi.gg_grad_scene), _P, C.POINTER(_D), _P, _P, _P, C.c_int]),
    ("oracle_loss_frozen", C.c_int, [C.POINTER(_abi.gg_grad_scene), C.POINTER(_abi.gg_grad_scene), _P,
                                     C.POINTER(_D), _P, C.c_int]),
    ("oracle_grad_scene_to_scene", C.c_int, [C.POINTER(_abi.gg_grad_scene), C.POINTER(_abi.gg_scene)]),
    ("oracle_grad_decisions", C.c_int, [C.POINTER(_abi.gg_grad_scene), _P, C.c_int]),
]


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None:
        assert a.size == n
    return a


class Oracle:
    def __init__(self, threads: int = 0):
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        self.lib = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(self.lib, name)
            f.restype = res
            f.argtypes = args
        self.threads = threads
        self.n_hittables = 0

    # ---- metric level ----
    def metric_tensor(self, metric, pos):
        g = np.zeros(16)
        p = _f64(pos, 4)
        self.lib.oracle_metric_tensor(C.byref(metric.to_c()), p.ctypes.data, g.ctypes.data)
        return g.reshape(4, 4)

    def christoffel(self, metric, pos):
        g = np.zeros(64)
        p = _f64(pos, 4)
        self.lib.oracle_christoffel(C.byref(metric.to_c()), p.ctypes.data, g.ctypes.data)
        return g.reshape(4, 4, 4)

    def geodesic_acceleration(self, metric, pos, vel):
        pos = _f64(pos).reshape(-1, 4)
        vel = _f64(vel).reshape(-1, 4)
        acc = np.zeros_like(pos)
        self.lib.oracle_geodesic_acceleration(C.byref(metric.to_c()), len(pos), pos.ctypes.data, vel.ctypes.data,
                                              acc.ctypes.data)
        return acc

    def constants_of_motion(self, metric, states):
        """[n, 4]: E, L_z, Carter Q, mu^2 of states [n, 9]."""
        states = _f64(states).reshape(-1, 9)
        out = np.zeros((len(states), 4))
        self.lib.oracle_constants_of_motion(C.byref(metric.to_c()), len(states), states.ctypes.data, out.ctypes.data)
        return out

    def rhs_jacobian(self, metric, pos, vel):
        """[n, 4, 7]: d a^mu / d(r, theta, v^t, v^r, v^theta, v^phi, spin) (dual numbers)."""
        pos = _f64(pos).reshape(-1, 4)
        vel = _f64(vel).reshape(-1, 4)
        jac = np.zeros((len(pos), 4, 7))
        self.lib.oracle_rhs_jacobian(C.byref(metric.to_c()), len(pos), pos.ctypes.data, vel.ctypes.data,
                                     jac.ctypes.data)
        return jac

    def event_horizon(self, metric):
        rh = C.c_double()
        ok = self.lib.oracle_event_horizon(C.byref(metric.to_c()), C.byref(rh))
        return rh.value if ok else None

    def orbital_four_velocity(self, metric, pos):
        u = np.zeros(4)
        p = _f64(pos, 4)
        ok = self.lib.oracle_orbital_four_velocity(C.byref(metric.to_c()), p.ctypes.data, u.ctypes.data)
        return u if ok else None

    def redshift_factor(self, metric, pos, p, sqrt_neg_gtt_obs):
        x = _f64(pos, 4)
        k = _f64(p, 4)
        return self.lib.oracle_redshift_factor(C.byref(metric.to_c()), x.ctypes.data, k.ctypes.data,
                                               sqrt_neg_gtt_obs)

    # ---- integrator level ----
    def rk45_step(self, metric, integ, state, h):
        s = _f64(state, 9).copy()
        hh = C.c_double(h)
        acc = C.c_int()
        nr = C.c_int()
        res = self.lib.oracle_rk45_step(C.byref(metric.to_c()), C.byref(integ.to_c()), s.ctypes.data, C.byref(hh),
                                        C.byref(acc), C.byref(nr))
        return res, s, hh.value, bool(acc.value), nr.value

    def rk4_step(self, metric, step, state, max_radius=100.0, horizon_buffer=0.01):
        s = _f64(state, 9).copy()
        res = self.lib.oracle_rk4_step(C.byref(metric.to_c()), step, max_radius, horizon_buffer, s.ctypes.data)
        return res, s

    def pixel_direction(self, camera, u, v):
        d = np.zeros(3)
        self.lib.oracle_pixel_direction(C.byref(camera.to_c()), u, v, d.ctypes.data)
        return d

    def from_camera(self, scene: Scene, px, py, width=None, height=None):
        s = np.zeros(9)
        sc = scene.to_c()
        self.lib.oracle_from_camera(C.byref(sc), px, py, width or scene.width * scene.samples_per_axis,
                                    height or scene.height * scene.samples_per_axis, s.ctypes.data)
        return s

    # ---- disk & shading ----
    def disk_density(self, disk, pos):
        p = _f64(pos, 4)
        return self.lib.oracle_disk_density(C.byref(disk.to_c()), p.ctypes.data)

    def disk_emission(self, disk, pos):
        p = _f64(pos, 4)
        return self.lib.oracle_disk_emission(C.byref(disk.to_c()), p.ctypes.data)

    def disk_color(self, intensity):
        out = np.zeros(3)
        self.lib.oracle_disk_color(intensity, out.ctypes.data)
        return out

    def encode_srgb(self, lin):
        lin = _f64(lin, 3)
        out = np.zeros(3, np.uint8)
        self.lib.oracle_encode_srgb(lin.ctypes.data, out.ctypes.data)
        return out

    def sky(self, d):
        d = _f64(d, 3)
        out = np.zeros(3)
        self.lib.oracle_sky(d.ctypes.data, out.ctypes.data)
        return out

    def hdri_sample(self, pixels, d, intensity=1.0):
        """HdriSky::sample for pixels [height, width, 3] (f32, row 0 = +z pole)."""
        px = np.ascontiguousarray(pixels, dtype=np.float32)
        d = _f64(d, 3)
        out = np.zeros(3)
        self.lib.oracle_hdri_sample(px.shape[1], px.shape[0], px.ctypes.data, intensity, d.ctypes.data,
                                    out.ctypes.data)
        return out

    def set_sky_hdri(self, pixels=None, intensity=1.0):
        """Sky of render() (None: the procedural sky)."""
        if pixels is None:
            self.lib.oracle_set_sky_hdri(0, 0, None, 1.0)
            return
        px = np.ascontiguousarray(pixels, dtype=np.float32)
        assert self.lib.oracle_set_sky_hdri(px.shape[1], px.shape[0], px.ctypes.data, intensity) == 0

    def splitmix64(self, seed, n):
        return self.lib.oracle_splitmix64(seed, n)

    def r_isco(self, mass, spin):
        return self.lib.oracle_r_isco(mass, spin)

    # ---- whole paths ----
    def set_user_metric(self, metric=None):
        """The tensor of GG_METRIC_USER scenes: metric.source (the device code a Renderer compiles) built as
        host C++ with g++ (-ffp-contract=off), or None to clear."""
        if metric is None:
            self.lib.oracle_set_user_metric(None, 0, 0.0)
            self._user_lib = None
            return
        import hashlib
        import subprocess
        import tempfile
        text = ("#include <cmath>\nusing namespace std;\n#define __device__\n#define __host__\n"
                "namespace gg {\n" + metric.source + "\n}\n"
                "extern \"C\" void gg_user_metric_host(const double *x, double m, double a, double *g) "
                "{ gg::gg_user_metric(x, m, a, g); }\n")
        cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared"]
        # keyed by the whole translation unit and the compile line, in a
        # directory private to this oracle (mkdtemp: mode 0700)
        tag = hashlib.sha256((text + "\0" + " ".join(cmd)).encode()).hexdigest()[:32]
        if getattr(self, "_user_dir", None) is None:
            self._user_dir = tempfile.mkdtemp(prefix="gg_oracle_user_metrics_")
        so = os.path.join(self._user_dir, f"um_{tag}.so")
        if not os.path.exists(so):
            src = os.path.join(self._user_dir, f"um_{tag}.cpp")
            with open(src, "w") as f:
                f.write(text)
            tmp = f"{so}.tmp"
            subprocess.run(cmd + ["-o", tmp, src], check=True)
            os.replace(tmp, so)
        self._user_lib = C.CDLL(so)
        fn = C.cast(self._user_lib.gg_user_metric_host, C.c_void_p)
        rh = metric.event_horizon
        self.lib.oracle_set_user_metric(fn, 0 if rh is None else 1, 0.0 if rh is None else float(rh))

    def inverse4(self, m):
        a = np.ascontiguousarray(np.asarray(m, dtype=np.float64).reshape(16))
        out = np.zeros(16)
        ok = self.lib.oracle_inverse4(a.ctypes.data, out.ctypes.data)
        return out.reshape(4, 4) if ok else None

    def set_hittables(self, world=None):
        """Scene geometry of render / trace_states (None: none); returns the status code."""
        if world is None or len(world) == 0:
            self.lib.oracle_set_hittables(0, None)
            self.n_hittables = 0
            return 0
        arr = world.to_c()
        rc = self.lib.oracle_set_hittables(len(world), C.cast(arr, C.c_void_p))
        self.n_hittables = len(world) if rc == 0 else 0
        return rc

    def scene_hit(self, origin, direction):
        """Closest hit of origin + t direction, t in (0, 1): None or (object, point, normal, shade_hit colour)."""
        out = np.zeros(12)
        o, d = _f64(origin, 3), _f64(direction, 3)
        if not self.lib.oracle_scene_hit(o.ctypes.data, d.ctypes.data, out.ctypes.data):
            return None
        return int(out[1]), out[2:5].copy(), out[5:8].copy(), out[8:11].copy()

    def g4_static(self, metric, point, g_tt_obs):
        p = _f64(point, 3)
        return self.lib.oracle_g4_static(C.byref(metric.to_c()), p.ctypes.data, g_tt_obs)

    def render(self, scene: Scene, rgb: bool = False, state: bool = True) -> dict:
        o = alloc_outputs(scene.n_rays, state=state, rgb_pixels=scene.width * scene.height if rgb else 0,
                          color=self.n_hittables > 0)
        sc = scene.to_c()
        rc = self.lib.oracle_render(C.byref(sc), C.byref(outputs_struct(o)), self.threads)
        if rc:
            raise RuntimeError(f"oracle_render failed: {rc}")
        o["rhs"] = self.lib.oracle_last_rhs()
        o["reflections"] = self.lib.oracle_last_reflections()
        return o

    def render_rays(self, scene: Scene, idx, state: bool = True) -> dict:
        idx = np.ascontiguousarray(idx, dtype=np.uint64)
        o = alloc_outputs(len(idx), state=state, color=self.n_hittables > 0)
        sc = scene.to_c()
        rc = self.lib.oracle_render_rays(C.byref(sc), len(idx), idx.ctypes.data, C.byref(outputs_struct(o)),
                                         self.threads)
        if rc:
            raise RuntimeError(f"oracle_render_rays failed: {rc}")
        o["rhs"] = self.lib.oracle_last_rhs()
        o["reflections"] = self.lib.oracle_last_reflections()
        return o

    def trace_states(self, scene: Scene, states) -> dict:
        states = _f64(states).reshape(-1, 9)
        o = alloc_outputs(len(states), color=self.n_hittables > 0)
        sc = scene.to_c()
        rc = self.lib.oracle_trace_states(C.byref(sc), len(states), states.ctypes.data, C.byref(outputs_struct(o)),
                                          self.threads)
        if rc:
            raise RuntimeError(f"oracle_trace_states failed: {rc}")
        return o

    def grad(self, gscene, target, want_dI=False):
        n = gscene.width * gscene.height
        target = _f64(target, n)
        loss = C.c_double()
        grad = np.zeros(4)
        inten = np.zeros(n)
        dI = np.zeros(4 * n) if want_dI else None
        rc = self.lib.oracle_grad(C.byref(gscene), target.ctypes.data, C.byref(loss), grad.ctypes.data,
                                  inten.ctypes.data, dI.ctypes.data if want_dI else None, self.threads)
        if rc:
            raise RuntimeError(f"oracle_grad failed: {rc}")
        out = {"loss": loss.value, "grad": grad, "intensity": inten}
        if want_dI:
            out["dI"] = dI.reshape(n, 4)
        return out

    def grad_render(self, gscene):
        """The oracle render of the scene oracle_grad differentiates (cls, intensity, attempts, accepted, state)."""
        sc = _abi.gg_scene()
        rc = self.lib.oracle_grad_scene_to_scene(C.byref(gscene), C.byref(sc))
        if rc:
            raise RuntimeError(f"oracle_grad_scene_to_scene failed: {rc}")
        o = alloc_outputs(gscene.width * gscene.height, state=True)
        rc = self.lib.oracle_render(C.byref(sc), C.byref(outputs_struct(o)), self.threads)
        if rc:
            raise RuntimeError(f"oracle_render failed: {rc}")
        return o

    def grad_decisions(self, gscene):
        """[N] uint64: hash of each ray's sequence of step decisions in the render oracle_grad differentiates."""
        out = np.zeros(gscene.width * gscene.height, dtype=np.uint64)
        rc = self.lib.oracle_grad_decisions(C.byref(gscene), out.ctypes.data, self.threads)
        if rc:
            raise RuntimeError(f"oracle_grad_decisions failed: {rc}")
        return out

    def loss_frozen(self, ref, gscene, target):
        n = gscene.width * gscene.height
        target = _f64(target, n)
        loss = C.c_double()
        inten = np.zeros(n)
        rc = self.lib.oracle_loss_frozen(C.byref(ref), C.byref(gscene), target.ctypes.data, C.byref(loss),
                                         inten.ctypes.data, self.threads)
        if rc:
            raise RuntimeError(f"oracle_loss_frozen failed: {rc}")
        return loss.value, inten
