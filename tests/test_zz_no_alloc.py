"""gg_render_batch_noalloc, the batch render that allocates nothing (its
scene tables and counters live in a static device array per GPU), against
gg_render_batch on the same scenes and device output buffers: bit for bit in
FP64, to the batch tests' mixed-precision bar otherwise.  Two calls in a row,
the second with fewer and smaller scenes (the array keeps the first call's
contents), a call at the maximum batch size, and the refusals."""
import ctypes as C

import numpy as np
import pytest

import gargantua_amd as gg
from gargantua_amd import _abi
from tests.test_gpu_batch import assert_same

pytestmark = pytest.mark.gpu

KEYS = ("cls", "r", "theta", "phi", "intensity", "attempts", "accepted")
_DT = {"cls": np.uint8, "attempts": np.uint32, "accepted": np.uint32}


class _DevOutputs:
    """One scene's per-ray output planes in device memory, from the HIP
    runtime the library uses."""

    def __init__(self, n):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.n = n
        self.ptr = {}
        for k in KEYS:
            p = C.c_void_p()
            nbytes = n * np.dtype(_DT.get(k, np.float64)).itemsize
            assert self.hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
            assert self.hip.hipMemset(p, 0, C.c_size_t(nbytes)) == 0
            self.ptr[k] = p

    def struct(self, **extra):
        return gg.outputs_struct({**{k: p.value for k, p in self.ptr.items()}, **extra})

    def host(self):
        o = {}
        for k, p in self.ptr.items():
            a = np.zeros(self.n, _DT.get(k, np.float64))
            assert self.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, C.c_size_t(a.nbytes), 2) == 0
            o[k] = a
        return o

    def free(self):
        for p in self.ptr.values():
            self.hip.hipFree(p)
        self.ptr = {}


def _call(renderer, fn, scenes, precision, outs):
    arr = (_abi.gg_scene * len(scenes))(*[s.to_c() for s in scenes])
    oarr = (_abi.gg_outputs * len(scenes))(*[o.struct() for o in outs])
    st = _abi.gg_stats()
    rc = getattr(renderer.lib, fn)(renderer.ctx, arr, len(scenes), gg.PRECISIONS[precision], oarr, C.byref(st))
    return rc, st


def _render_both(renderer, scenes, precision):
    a = [_DevOutputs(s.n_rays) for s in scenes]
    b = [_DevOutputs(s.n_rays) for s in scenes]
    try:
        rc, st = _call(renderer, "gg_render_batch_noalloc", scenes, precision, a)
        assert rc == 0, renderer.lib.gg_last_error().decode()
        rc2, st2 = _call(renderer, "gg_render_batch", scenes, precision, b)
        assert rc2 == 0, renderer.lib.gg_last_error().decode()
        assert st.rays == st2.rays == sum(s.n_rays for s in scenes)
        if precision != "mixed":
            assert st.attempts == st2.attempts and st.rhs_evals == st2.rhs_evals
        return [o.host() for o in a], [o.host() for o in b]
    finally:
        for o in a + b:
            o.free()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_render_batch_on_two_calls(renderer, precision):
    first = gg.configs.batch_64(seed=7, n=6, resolutions=(24, 40, 64))
    second = gg.configs.batch_64(seed=11, n=3, resolutions=(16, 24))
    assert sum(s.n_rays for s in second) < sum(s.n_rays for s in first)
    for tag, scenes in (("first", first), ("second", second)):
        print(f"{precision} {tag} call: {len(scenes)} scenes, {sum(s.n_rays for s in scenes)} rays", flush=True)
        got, want = _render_both(renderer, scenes, precision)
        for i, (g, w) in enumerate(zip(got, want)):
            assert_same(g, w, precision, what=f"{tag} call, scene {i}: ")


def test_noalloc_at_the_maximum_batch(renderer):
    n_max = 64
    scenes = gg.configs.batch_64(seed=3, n=n_max, resolutions=(8, 12))
    print(f"fp64 maximum batch: {n_max} scenes, {sum(s.n_rays for s in scenes)} rays", flush=True)
    got, want = _render_both(renderer, scenes, "fp64")
    for i, (g, w) in enumerate(zip(got, want)):
        assert_same(g, w, "fp64", what=f"scene {i}: ")


def test_noalloc_refusals(renderer):
    scenes = gg.configs.batch_64(seed=3, n=65, resolutions=(8,))
    outs = [_DevOutputs(s.n_rays) for s in scenes]
    try:
        rc, _ = _call(renderer, "gg_render_batch_noalloc", scenes, "fp64", outs)
        assert rc == gg.GG_ERR_INVALID_ARG
        assert "at most 64 scenes" in renderer.lib.gg_last_error().decode()
        sc = scenes[:1]
        arr = (_abi.gg_scene * 1)(sc[0].to_c())
        # host output buffers would need staging
        host = gg.alloc_outputs(sc[0].n_rays, state=False)
        oarr = (_abi.gg_outputs * 1)(gg.outputs_struct(host))
        rc = renderer.lib.gg_render_batch_noalloc(renderer.ctx, arr, 1, gg.PRECISIONS["fp64"], oarr, None)
        assert rc == gg.GG_ERR_INVALID_ARG
        assert "device memory only" in renderer.lib.gg_last_error().decode()
        # so would rgb
        rgb = _DevOutputs(3 * sc[0].n_rays)
        try:
            oarr = (_abi.gg_outputs * 1)(outs[0].struct(rgb=rgb.ptr["cls"].value))
            rc = renderer.lib.gg_render_batch_noalloc(renderer.ctx, arr, 1, gg.PRECISIONS["fp64"], oarr, None)
            assert rc == gg.GG_ERR_UNSUPPORTED
        finally:
            rgb.free()
    finally:
        for o in outs:
            o.free()
    # the entry point still works after the refusals
    got, want = _render_both(renderer, sc, "fp64")
    assert_same(got[0], want[0], "fp64")
