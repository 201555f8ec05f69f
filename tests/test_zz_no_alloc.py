"""The allocation-free entry point scopi_hip_se_closest_points_noalloc (device buffers of the caller,
scratch in the library's static device array) against scopi_hip_se_closest_points: the same kernels,
so every output is bitwise equal, on two calls in a row, the second with fewer pairs (the scratch
array keeps the first call's contents); and the refusal of more pairs than the array holds.
The test (the caller) allocates its device buffers with hipMalloc of the HIP runtime the library
itself links, so every pointer belongs to the runtime that launches the kernels."""
import ctypes as C
import sys

import numpy as np
import pytest

from scopi_amd import hip, scenes

pytestmark = pytest.mark.gpu

OUT = (("dij", 1, np.float64), ("nij", 3, np.float64), ("pi", 3, np.float64), ("pj", 3, np.float64),
       ("u0", 4, np.float64), ("u", 4, np.float64), ("info", 1, np.int32), ("nfev", 1, np.int32))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_pairs(n, seed):
    rng = np.random.default_rng(seed)
    qa = rng.normal(size=(n, 4)); qa /= np.linalg.norm(qa, axis=1, keepdims=True)
    qb = rng.normal(size=(n, 4)); qb /= np.linalg.norm(qb, axis=1, keepdims=True)
    ra = rng.uniform(0.008, 0.015, (n, 3)); rb = rng.uniform(0.008, 0.015, (n, 3))
    ea = rng.uniform(0.4, 1.6, (n, 2)); eb = rng.uniform(0.4, 1.6, (n, 2))
    ca = rng.uniform(-1, 1, (n, 3))
    cb = ca + np.array([0.035, 0.0, 0.0]) + rng.uniform(-0.01, 0.003, (n, 3))
    return ca, qa, ra, ea, cb, qb, rb, eb


@pytest.fixture(scope="module")
def ctx():
    g = hip.HipScene(scenes.superellipsoids_3d(2, 2, 2, seed=1))
    yield g
    g.close()


def hip_runtime():
    """The HIP runtime libscopi_hip.so is linked against (already loaded: the same handle)."""
    from scopi_amd import hip as h
    h.lib()
    with open("/proc/self/maps") as fh:
        paths = {ln.split()[-1] for ln in fh if "libamdhip64.so" in ln and "/torch/" not in ln}
    rt = C.CDLL(sorted(paths)[0])
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return rt


class DeviceBuffer:
    def __init__(self, rt, host):
        self.rt, self.host = rt, np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert rt.hipMalloc(C.byref(self.ptr), max(self.host.nbytes, 8)) == 0
        assert rt.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0  # host to device

    def read(self):
        out = np.empty_like(self.host)
        assert self.rt.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        self.rt.hipFree(self.ptr)


def test_noalloc_matches_se_closest_points(ctx):
    rt = hip_runtime()
    for n, seed in ((200, 3), (37, 4)):
        say(f"[noalloc] {n} pairs (seed {seed})")
        args = random_pairs(n, seed)
        ref = ctx.se_closest_points(*args)
        ins = [DeviceBuffer(rt, np.asarray(a, np.float64)) for a in args]
        outs = {k: DeviceBuffer(rt, np.full((n, w) if w > 1 else (n,), -7, t)) for k, w, t in OUT}
        try:
            assert rt.hipDeviceSynchronize() == 0
            rc = ctx.se_closest_points_noalloc(n, [b.ptr.value for b in ins] + [outs[k].ptr.value for k, _, _ in OUT])
            assert rc == 0, ctx.last_error()
            assert rt.hipDeviceSynchronize() == 0
            for k, _, _ in OUT:
                got = outs[k].read()
                assert np.array_equal(got, ref[k]), (k, got[:3], ref[k][:3])
        finally:
            for b in ins + list(outs.values()):
                b.free()
        say(f"  bitwise equal; info {np.bincount(ref['info'])}")


def test_noalloc_refuses_more_pairs_than_its_scratch(ctx):
    rc = ctx.se_closest_points_noalloc(hip.NOALLOC_MAX_PAIRS + 1, [0] * 16)
    assert rc == -3  # SCOPI_HIP_ERR_CAPACITY, before any device work
    assert "SCOPI_HIP_NOALLOC_MAX_PAIRS" in ctx.last_error()
