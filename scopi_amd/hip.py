"""ctypes binding of libscopi_hip.so (include/scopi_hip.h).

This is the product path's Python face: it loads the in-tree HIP library and
raises if it is missing or if a call fails -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libscopi_hip.so")

SHAPE_SPHERE, SHAPE_PLANE, SHAPE_SEGMENT, SHAPE_SUPERELLIPSOID = 0, 1, 2, 3
NOFRICTION, FRICTION, FRICTION_FIXED_POINT, VISCOUS, VISCOUS_FRICTION = 0, 1, 2, 3, 4
KDTREE, BRUTE_FORCE = 0, 1
VAP_FIXED, VAP_FPD, VAP_PROJECTION = 0, 1, 2
APGD, PGD = 0, 1
FP64, FP32_REFINE = 0, 1

# every symbol include/scopi_hip.h declares
EXPORTS = [
    "scopi_hip_default_params", "scopi_hip_create", "scopi_hip_destroy", "scopi_hip_last_error",
    "scopi_hip_upload_state", "scopi_hip_download_state", "scopi_hip_upload_velocities",
    "scopi_hip_compute_contacts", "scopi_hip_download_contacts", "scopi_hip_upload_contacts",
    "scopi_hip_set_a_priori_velocity", "scopi_hip_solve", "scopi_hip_download_solution",
    "scopi_hip_update_and_move", "scopi_hip_displacement_obstacles", "scopi_hip_run",
    "scopi_hip_set_kernel_timing", "scopi_hip_kernel_timing", "scopi_hip_set_tuning", "scopi_hip_se_closest_points",
    "scopi_hip_closest_points", "scopi_hip_portable_math", "scopi_hip_newton_probe",
    "scopi_hip_batch_create", "scopi_hip_batch_destroy", "scopi_hip_batch_last_error", "scopi_hip_batch_add_scene",
    "scopi_hip_batch_remove_scene", "scopi_hip_batch_step", "scopi_hip_batch_size", "scopi_hip_batch_download_scene",
    "scopi_hip_batch_scene_info", "scopi_hip_batch_download_contacts", "scopi_hip_batch_download_solution",
    "scopi_hip_scene_footprint", "scopi_hip_set_box", "scopi_hip_set_projection_velocity",
    "scopi_hip_set_contact_gamma", "scopi_hip_download_contact_gamma", "scopi_hip_se_closest_points_noalloc",
    "scopi_hip_batch_set_tuning",
]
NOALLOC_MAX_PAIRS = 256  # SCOPI_HIP_NOALLOC_MAX_PAIRS
TUNE_REDUCE_BLOCKS, TUNE_BT_GROUP, TUNE_GRAPH_ROUNDS = 1, 2, 3
TUNE_QUEUE_MAX_CONTACTS, TUNE_QUEUE_BLOCKS, TUNE_QUEUE_THREADS, TUNE_ROW_ITERS = 4, 5, 6, 7


class Params(C.Structure):
    _fields_ = [
        ("problem", C.c_int32), ("contact_method", C.c_int32), ("vap", C.c_int32), ("method", C.c_int32),
        ("precision", C.c_int32), ("dynamic_descent", C.c_int32), ("dmax", C.c_double),
        ("kd_tree_radius", C.c_double), ("alpha", C.c_double), ("tolerance", C.c_double), ("mu", C.c_double),
        ("max_ite", C.c_int64), ("fp32_tolerance", C.c_double), ("warm_start", C.c_int32),
        ("general_jacobian", C.c_int32), ("fixed_point_tol", C.c_double), ("fixed_point_max_iter", C.c_int64),
        ("gamma", C.c_double), ("gamma_min", C.c_double), ("gamma_tol", C.c_double),
    ]


class SolveInfo(C.Structure):
    _fields_ = [
        ("n_contacts", C.c_int64), ("rows", C.c_int64), ("iterations", C.c_int64), ("iterations_fp32", C.c_int64),
        ("backtracks", C.c_int64), ("q_applies", C.c_int64), ("converged", C.c_int32), ("status", C.c_int32),
        ("residual", C.c_double), ("outer_iterations", C.c_int64), ("fixed_point_error", C.c_double),
        ("step", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


class ScopiHipError(RuntimeError):
    pass


def lib():
    """Load libscopi_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ScopiHipError(f"{LIB_PATH} is missing: build it with `python -m scopi_amd.build`")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    I32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    I64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
    Dn = C.c_void_p  # nullable double*
    L.scopi_hip_default_params.argtypes = [C.POINTER(Params)]
    L.scopi_hip_create.argtypes = [C.c_int32, C.c_int64, C.c_int64, C.POINTER(P)]
    L.scopi_hip_destroy.argtypes = [P]
    L.scopi_hip_last_error.argtypes = [P]
    L.scopi_hip_last_error.restype = C.c_char_p
    L.scopi_hip_upload_state.argtypes = [P, C.c_int64, C.c_int64, I32, D, D, D, D, D, D, D, D, D, D, D]
    L.scopi_hip_download_state.argtypes = [P, Dn, Dn, Dn, Dn]
    L.scopi_hip_upload_velocities.argtypes = [P, Dn, Dn]
    L.scopi_hip_compute_contacts.argtypes = [P, C.POINTER(Params), C.POINTER(C.c_int64)]
    L.scopi_hip_download_contacts.argtypes = [P, I64, I64, D, D, D, D]
    L.scopi_hip_upload_contacts.argtypes = [P, C.c_int64, I64, I64, D, D, D, D, Dn]
    L.scopi_hip_set_a_priori_velocity.argtypes = [P, C.POINTER(Params), C.c_double]
    L.scopi_hip_solve.argtypes = [P, C.POINTER(Params), C.c_double, C.POINTER(SolveInfo)]
    L.scopi_hip_download_solution.argtypes = [P, Dn, Dn, Dn]
    L.scopi_hip_update_and_move.argtypes = [P, C.c_double]
    L.scopi_hip_displacement_obstacles.argtypes = [P, C.c_double]
    L.scopi_hip_run.argtypes = [P, C.POINTER(Params), C.c_double, C.c_int64, C.c_int64, C.POINTER(SolveInfo)]
    L.scopi_hip_set_kernel_timing.argtypes = [P, C.c_int32]
    L.scopi_hip_kernel_timing.argtypes = [P, D]
    L.scopi_hip_set_tuning.argtypes = [P, C.c_int32, C.c_int32]
    L.scopi_hip_se_closest_points.argtypes = [P, C.c_int64] + [D] * 12 + [C.c_void_p] * 4
    L.scopi_hip_se_closest_points_noalloc.argtypes = [P, C.c_void_p, C.c_int64] + [C.c_void_p] * 16
    L.scopi_hip_portable_math.argtypes = [P, C.c_int32, C.c_int64, D, C.c_void_p, D]
    L.scopi_hip_newton_probe.argtypes = [P, C.c_int32, D, D]
    L.scopi_hip_closest_points.argtypes = ([P, C.c_int64, I32, D, D, D, C.c_void_p, I32, D, D, D, C.c_void_p] +
                                           [D] * 4 + [C.c_void_p] * 4)
    L.scopi_hip_set_box.argtypes = [P, D, D, I32]
    L.scopi_hip_set_projection_velocity.argtypes = [P, D, D]
    L.scopi_hip_set_contact_gamma.argtypes = [P, Dn]
    L.scopi_hip_download_contact_gamma.argtypes = [P, Dn]
    L.scopi_hip_batch_create.argtypes = [C.POINTER(Params), C.c_double, C.POINTER(P)]
    L.scopi_hip_batch_destroy.argtypes = [P]
    L.scopi_hip_batch_last_error.argtypes = [P]
    L.scopi_hip_batch_last_error.restype = C.c_char_p
    L.scopi_hip_batch_add_scene.argtypes = [P, C.c_int64, C.c_int64, I32, D, D, D, D, D, D, D, D, D, D, D,
                                            C.POINTER(C.c_int64)]
    L.scopi_hip_batch_remove_scene.argtypes = [P, C.c_int64]
    L.scopi_hip_batch_step.argtypes = [P, C.c_int64]
    L.scopi_hip_batch_set_tuning.argtypes = [P, C.c_int32, C.c_int32]
    L.scopi_hip_batch_size.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p]
    L.scopi_hip_batch_download_scene.argtypes = [P, C.c_int64, Dn, Dn, Dn, Dn]
    L.scopi_hip_batch_scene_info.argtypes = [P, C.c_int64, C.c_void_p, C.c_void_p]
    L.scopi_hip_batch_download_contacts.argtypes = [P, C.c_int64, C.POINTER(C.c_int64)] + [C.c_void_p] * 6
    L.scopi_hip_batch_download_solution.argtypes = [P, C.c_int64, Dn, Dn, Dn]
    L.scopi_hip_scene_footprint.argtypes = [C.c_int64, C.c_int64, C.POINTER(Params)]
    L.scopi_hip_scene_footprint.restype = C.c_int64
    for name in [n for n in dir(L) if n.startswith("scopi_hip_") and n not in (
            "scopi_hip_last_error", "scopi_hip_batch_last_error", "scopi_hip_scene_footprint")]:
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def params(**kw) -> Params:
    p = Params()
    lib().scopi_hip_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        setattr(p, k, v)
    return p


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class HipScene:
    """One scene resident on the current GPU (a scopi_hip_ctx)."""

    def __init__(self, scene, capacity_contacts=0):
        self._L = lib()
        self._h = C.c_void_p()
        rc = self._L.scopi_hip_create(scene.dim, scene.n, capacity_contacts, C.byref(self._h))
        if rc != 0:
            raise ScopiHipError(f"scopi_hip_create failed ({rc})")
        self.dim = scene.dim
        self.n = scene.n
        self.n_inactive = scene.n_inactive
        self.nc = 0
        self.upload(scene)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.scopi_hip_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc < 0:
            msg = self._L.scopi_hip_last_error(self._h).decode()
            raise ScopiHipError(f"{what} failed ({rc}): {msg}")
        return rc

    def upload(self, scene):
        f = lambda a: np.ascontiguousarray(a, np.float64)
        self._check(self._L.scopi_hip_upload_state(
            self._h, scene.n, scene.n_inactive, np.ascontiguousarray(scene.shape, np.int32), f(scene.pos), f(scene.q),
            f(scene.size), f(scene.squareness), f(scene.v), f(scene.vd), f(scene.omega), f(scene.domega), f(scene.f),
            f(scene.m), f(scene.J)), "upload_state")
        self.n = scene.n
        self.n_inactive = scene.n_inactive

    def state(self):
        n = self.n
        pos, q, v, om = np.zeros((n, 3)), np.zeros((n, 4)), np.zeros((n, 3)), np.zeros((n, 3))
        self._check(self._L.scopi_hip_download_state(self._h, _ptr(pos), _ptr(q), _ptr(v), _ptr(om)), "download_state")
        return dict(pos=pos, q=q, v=v, omega=om)

    def upload_velocities(self, v=None, omega=None):
        v = None if v is None else np.ascontiguousarray(v, np.float64)
        omega = None if omega is None else np.ascontiguousarray(omega, np.float64)
        self._check(self._L.scopi_hip_upload_velocities(self._h, _ptr(v), _ptr(omega)), "upload_velocities")

    def compute_contacts(self, prm: Params) -> int:
        nc = C.c_int64()
        self._check(self._L.scopi_hip_compute_contacts(self._h, C.byref(prm), C.byref(nc)), "compute_contacts")
        self.nc = nc.value
        return nc.value

    def contacts(self):
        nc = self.nc
        ci = np.zeros(max(nc, 1), np.int64); cj = np.zeros(max(nc, 1), np.int64); dij = np.zeros(max(nc, 1))
        nij = np.zeros((max(nc, 1), 3)); pi = np.zeros((max(nc, 1), 3)); pj = np.zeros((max(nc, 1), 3))
        self._check(self._L.scopi_hip_download_contacts(self._h, ci, cj, dij, nij, pi, pj), "download_contacts")
        return dict(ij=np.stack([ci[:nc], cj[:nc]], axis=1), dij=dij[:nc], nij=nij[:nc], pi=pi[:nc], pj=pj[:nc])

    def upload_contacts(self, ij, dij, nij, pi, pj, mu=None):
        nc = int(len(dij))
        ij = np.asarray(ij, np.int64).reshape(nc, 2)
        f = lambda a, w: np.ascontiguousarray(np.asarray(a, np.float64).reshape(max(nc, 1) if nc else 1, w)
                                               if nc else np.zeros((1, w)))
        ci = np.ascontiguousarray(ij[:, 0]) if nc else np.zeros(1, np.int64)
        cj = np.ascontiguousarray(ij[:, 1]) if nc else np.zeros(1, np.int64)
        mu_a = None if mu is None else np.ascontiguousarray(mu, np.float64)
        self._check(self._L.scopi_hip_upload_contacts(self._h, nc, ci, cj, np.ascontiguousarray(dij, np.float64)
                                                      if nc else np.zeros(1), f(nij, 3), f(pi, 3), f(pj, 3),
                                                      _ptr(mu_a)), "upload_contacts")
        self.nc = nc

    def set_a_priori_velocity(self, prm: Params, dt):
        self._check(self._L.scopi_hip_set_a_priori_velocity(self._h, C.byref(prm), float(dt)), "set_a_priori_velocity")

    def solve(self, prm: Params, dt):
        info = SolveInfo()
        self._check(self._L.scopi_hip_solve(self._h, C.byref(prm), float(dt), C.byref(info)), "solve")
        self.last_rows = info.rows
        return info.as_dict()

    def set_contact_gamma(self, gamma):
        """gamma of every current contact (Viscous problems), in the contact list's order."""
        g = np.ascontiguousarray(np.broadcast_to(np.asarray(gamma, np.float64), (max(self.nc, 1),)))
        self._check(self._L.scopi_hip_set_contact_gamma(self._h, _ptr(g)), "set_contact_gamma")

    def contact_gamma(self):
        g = np.zeros(max(self.nc, 1))
        self._check(self._L.scopi_hip_download_contact_gamma(self._h, _ptr(g)), "download_contact_gamma")
        return g[: self.nc]

    def solution(self):
        na = self.n - self.n_inactive
        u = np.zeros((max(na, 1), 3)); w = np.zeros((max(na, 1), 3))
        rows = getattr(self, "last_rows", 0)
        lam = np.zeros(max(rows, 1))
        self._check(self._L.scopi_hip_download_solution(self._h, _ptr(u), _ptr(w), _ptr(lam)), "download_solution")
        return dict(u=u[:na], w=w[:na], lam=lam[:rows])

    def update_and_move(self, dt):
        self._check(self._L.scopi_hip_update_and_move(self._h, float(dt)), "update_and_move")

    def displacement_obstacles(self, dt):
        self._check(self._L.scopi_hip_displacement_obstacles(self._h, float(dt)), "displacement_obstacles")

    def run(self, prm: Params, dt, total_it, initial_iter=0):
        info = SolveInfo()
        self._check(self._L.scopi_hip_run(self._h, C.byref(prm), float(dt), int(total_it), int(initial_iter),
                                          C.byref(info)), "run")
        if total_it > initial_iter:
            self.nc = info.n_contacts  # contacts of the last step stay on the device
            self.last_rows = info.rows
        return info.as_dict()

    def set_projection_velocity(self, u, w):
        """vap_projection::set_u_w: u, w are Na x 3."""
        na = self.n - self.n_inactive
        u = np.ascontiguousarray(np.reshape(u, (max(na, 1) if na else 1, 3)) if na else np.zeros((1, 3)), np.float64)
        w = np.ascontiguousarray(np.reshape(w, (max(na, 1) if na else 1, 3)) if na else np.zeros((1, 3)), np.float64)
        self._check(self._L.scopi_hip_set_projection_velocity(self._h, u, w), "set_projection_velocity")

    def set_box(self, lo, hi, periodic):
        """BoxDomain(lo, hi).with_periodicity(axes with periodic[d] true) of the scene."""
        lo = np.ascontiguousarray(lo, np.float64)
        hi = np.ascontiguousarray(hi, np.float64)
        per = np.ascontiguousarray([1 if p else 0 for p in periodic], np.int32)
        self._check(self._L.scopi_hip_set_box(self._h, lo, hi, per), "set_box")

    def set_kernel_timing(self, on=True):
        self._check(self._L.scopi_hip_set_kernel_timing(self._h, 1 if on else 0), "set_kernel_timing")

    def set_tuning(self, key, value):
        self._check(self._L.scopi_hip_set_tuning(self._h, int(key), int(value)), "set_tuning")

    def se_closest_points(self, pos_i, q_i, size_i, sq_i, pos_j, q_j, size_j, sq_j):
        """closest_points(superellipsoid<3>, superellipsoid<3>) of n independent pairs on the device."""
        f = lambda a, w: np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1, w))
        args = [f(pos_i, 3), f(q_i, 4), f(size_i, 3), f(sq_i, 2), f(pos_j, 3), f(q_j, 4), f(size_j, 3), f(sq_j, 2)]
        n = args[0].shape[0]
        if any(a.shape[0] != n for a in args):
            raise ValueError("se_closest_points: every array needs one row per pair")
        m = max(n, 1)
        dij = np.zeros(m); nij = np.zeros((m, 3)); pi = np.zeros((m, 3)); pj = np.zeros((m, 3))
        u0 = np.zeros((m, 4)); u = np.zeros((m, 4)); info = np.zeros(m, np.int32); nfev = np.zeros(m, np.int32)
        self._check(self._L.scopi_hip_se_closest_points(self._h, n, *args, dij, nij, pi, pj, _ptr(u0), _ptr(u),
                                                        _ptr(info), _ptr(nfev)), "se_closest_points")
        return dict(dij=dij[:n], nij=nij[:n], pi=pi[:n], pj=pj[:n], u0=u0[:n], u=u[:n], info=info[:n],
                    nfev=nfev[:n])

    def se_closest_points_noalloc(self, n, ptrs, stream=None):
        """scopi_hip_se_closest_points_noalloc on device buffers: `ptrs` holds the 16 device addresses
        (pos_i, q_i, size_i, sq_i, pos_j, q_j, size_j, sq_j, dij, nij, pi, pj, u0, u, info, nfev; the last
        four may be 0), `stream` a hipStream_t handle (None: the context's stream).  Asynchronous;
        returns the status code."""
        return int(self._L.scopi_hip_se_closest_points_noalloc(self._h, stream, n, *[p or None for p in ptrs]))

    def last_error(self):
        return self._L.scopi_hip_last_error(self._h).decode()

    def closest_points(self, shape_i, pos_i, q_i, size_i, sq_i, shape_j, pos_j, q_j, size_j, sq_j):
        """closest_points_dispatcher::dispatch of n independent pairs (the context's dimension) on the device."""
        f = lambda a, w: np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1, w))
        si = np.ascontiguousarray(shape_i, np.int32); sj = np.ascontiguousarray(shape_j, np.int32)
        n = len(si)
        args = [f(pos_i, 3), f(q_i, 4), f(size_i, 3), f(sq_i, 2), f(pos_j, 3), f(q_j, 4), f(size_j, 3), f(sq_j, 2)]
        m = max(n, 1)
        dij = np.zeros(m); nij = np.zeros((m, 3)); pi = np.zeros((m, 3)); pj = np.zeros((m, 3))
        u0 = np.zeros((m, 4)); u = np.zeros((m, 4)); info = np.zeros(m, np.int32); nfev = np.zeros(m, np.int32)
        self._check(self._L.scopi_hip_closest_points(
            self._h, n, si, args[0], args[1], args[2], _ptr(args[3]), sj, args[4], args[5], args[6], _ptr(args[7]),
            dij, nij, pi, pj, _ptr(u0), _ptr(u), _ptr(info), _ptr(nfev)), "closest_points")
        return dict(dij=dij[:n], nij=nij[:n], pi=pi[:n], pj=pj[:n], u0=u0[:n], u=u[:n], info=info[:n],
                    nfev=nfev[:n])

    def portable_math(self, fn, x, y=None):
        """The device's portable elementary function `fn` (cos, sin, asin, log, exp, pow) elementwise."""
        code = {"cos": 0, "sin": 1, "asin": 2, "log": 3, "exp": 4, "pow": 5}[fn]
        x = np.ascontiguousarray(x, np.float64)
        yy = None if y is None else np.ascontiguousarray(y, np.float64)
        out = np.zeros(max(len(x), 1))
        self._check(self._L.scopi_hip_portable_math(self._h, code, len(x), x if len(x) else np.zeros(1), _ptr(yy),
                                                    out), "portable_math")
        return out[:len(x)]

    def newton_probe(self, kind, se, other, u, A, b):
        """F, J at u of a sphere (0) / plane (1) - superellipsoid pair and the solve of A x = b, on the device.
        se = (centre, q, semi-axes, (e, n)); other = (centre, q, radius)."""
        v = np.concatenate([np.ravel(a) for a in (se[0], se[1], se[2], se[3], other[0], other[1], [other[2]], u,
                                                  np.ravel(A), b)]).astype(np.float64)
        out = np.zeros(24)
        self._check(self._L.scopi_hip_newton_probe(self._h, kind, np.ascontiguousarray(v), out), "newton_probe")
        return out[:4], out[4:20].reshape(4, 4), out[20:]

    def kernel_timing(self):
        out = np.zeros(6)
        self._check(self._L.scopi_hip_kernel_timing(self._h, out), "kernel_timing")
        return dict(bt_ms=out[0], bt_launches=int(out[1]), row_ms=out[2], row_launches=int(out[3]),
                    trial_ms=out[4], trial_launches=int(out[5]))


def scene_footprint(n, n_inactive, prm: Params) -> int:
    """Device bytes a scene may hold in a batch (the admission estimate of a serving loop)."""
    return int(lib().scopi_hip_scene_footprint(int(n), int(n_inactive), C.byref(prm)))


class HipBatch:
    """Independent 3D sphere scenes stepped together on the current GPU (a scopi_hip_batch).

    Each scene follows ScopiSolver::run's semantics on its own; the batch shares one set of
    kernel launches per step among all of them, and a scene's results do not depend on which
    other scenes share the batch.
    """

    def __init__(self, prm: Params, dt):
        self._L = lib()
        self._h = C.c_void_p()
        self._prm = prm
        rc = self._L.scopi_hip_batch_create(C.byref(prm), float(dt), C.byref(self._h))
        if rc != 0:
            raise ScopiHipError(f"scopi_hip_batch_create failed ({rc})")
        self._scenes = {}

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.scopi_hip_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc < 0:
            msg = self._L.scopi_hip_batch_last_error(self._h).decode()
            raise ScopiHipError(f"{what} failed ({rc}): {msg}")
        return rc

    def add(self, scene) -> int:
        f = lambda a: np.ascontiguousarray(a, np.float64)
        sid = C.c_int64()
        self._check(self._L.scopi_hip_batch_add_scene(
            self._h, scene.n, scene.n_inactive, np.ascontiguousarray(scene.shape, np.int32), f(scene.pos), f(scene.q),
            f(scene.size), f(scene.squareness), f(scene.v), f(scene.vd), f(scene.omega), f(scene.domega), f(scene.f),
            f(scene.m), f(scene.J), C.byref(sid)), "batch_add_scene")
        self._scenes[sid.value] = (scene.n, scene.n_inactive)
        return sid.value

    def remove(self, sid):
        self._check(self._L.scopi_hip_batch_remove_scene(self._h, int(sid)), "batch_remove_scene")
        self._scenes.pop(sid, None)

    def step(self, n_steps=1):
        self._check(self._L.scopi_hip_batch_step(self._h, int(n_steps)), "batch_step")

    def set_tuning(self, key, value):
        self._check(self._L.scopi_hip_batch_set_tuning(self._h, int(key), int(value)), "batch_set_tuning")

    def size(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self._L.scopi_hip_batch_size(self._h, C.byref(a), C.byref(b), C.byref(c)), "batch_size")
        return dict(scenes=a.value, particles=b.value, contacts=c.value)

    def state(self, sid):
        n, _ = self._scenes[sid]
        pos, q, v, om = np.zeros((n, 3)), np.zeros((n, 4)), np.zeros((n, 3)), np.zeros((n, 3))
        self._check(self._L.scopi_hip_batch_download_scene(self._h, int(sid), _ptr(pos), _ptr(q), _ptr(v), _ptr(om)),
                    "batch_download_scene")
        return dict(pos=pos, q=q, v=v, omega=om)

    def info(self, sid):
        steps = C.c_int64()
        info = SolveInfo()
        self._check(self._L.scopi_hip_batch_scene_info(self._h, int(sid), C.byref(steps), C.byref(info)),
                    "batch_scene_info")
        d = info.as_dict()
        d["steps"] = steps.value
        return d

    def contacts(self, sid):
        nc = C.c_int64()
        self._check(self._L.scopi_hip_batch_download_contacts(self._h, int(sid), C.byref(nc), None, None, None, None,
                                                              None, None), "batch_download_contacts")
        m = max(nc.value, 1)
        ci = np.zeros(m, np.int64); cj = np.zeros(m, np.int64); dij = np.zeros(m)
        nij = np.zeros((m, 3)); pi = np.zeros((m, 3)); pj = np.zeros((m, 3))
        self._check(self._L.scopi_hip_batch_download_contacts(self._h, int(sid), C.byref(nc), _ptr(ci), _ptr(cj),
                                                              _ptr(dij), _ptr(nij), _ptr(pi), _ptr(pj)),
                    "batch_download_contacts")
        k = nc.value
        return dict(ij=np.stack([ci[:k], cj[:k]], axis=1), dij=dij[:k], nij=nij[:k], pi=pi[:k], pj=pj[:k])

    def solution(self, sid):
        n, ni = self._scenes[sid]
        na = n - ni
        rows = self.info(sid)["rows"]
        u = np.zeros((max(na, 1), 3)); w = np.zeros((max(na, 1), 3)); lam = np.zeros(max(rows, 1))
        self._check(self._L.scopi_hip_batch_download_solution(self._h, int(sid), _ptr(u), _ptr(w), _ptr(lam)),
                    "batch_download_solution")
        return dict(u=u[:na], w=w[:na], lam=lam[:rows])
