"""ctypes wrapper of the FP64 CPU restatement (liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg are the only callers.  The product path never imports this.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

NOFRICTION, FRICTION, FRICTION_FIXED_POINT, VISCOUS, VISCOUS_FRICTION = 0, 1, 2, 3, 4
KDTREE, BRUTE_FORCE = 0, 1
VAP_FIXED, VAP_FPD, VAP_PROJECTION = 0, 1, 2
APGD, PGD = 0, 1


class Params(C.Structure):
    _fields_ = [
        ("problem", C.c_int32), ("contact_method", C.c_int32), ("vap", C.c_int32), ("method", C.c_int32),
        ("dmax", C.c_double), ("kd_tree_radius", C.c_double), ("alpha", C.c_double), ("max_ite", C.c_int64),
        ("tolerance", C.c_double), ("dynamic_descent", C.c_int32), ("warm_start", C.c_int32), ("mu", C.c_double),
        ("fixed_point_tol", C.c_double), ("fixed_point_max_iter", C.c_double),
        ("gamma", C.c_double), ("gamma_min", C.c_double), ("gamma_tol", C.c_double),
        ("objective_perturbation", C.c_double),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("n_contacts", C.c_int64), ("iterations", C.c_int64), ("backtracks", C.c_int64), ("q_applies", C.c_int64),
        ("converged", C.c_int64), ("total_iterations", C.c_int64), ("total_q_applies", C.c_int64),
        ("fixed_point_iterations", C.c_int64), ("t_contacts", C.c_double), ("t_vap", C.c_double), ("t_solve", C.c_double), ("t_move", C.c_double),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        I32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        I64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        L.oracle_scene_new.restype = P
        L.oracle_scene_free.argtypes = [P]
        L.oracle_scene_clone.argtypes = [P]
        L.oracle_scene_clone.restype = P
        L.oracle_default_params.argtypes = [C.POINTER(Params)]
        L.oracle_set_params.argtypes = [P, C.POINTER(Params)]
        L.oracle_scene_set.argtypes = [P, C.c_int32, C.c_int64, C.c_int64, I32] + [D] * 11
        L.oracle_get_state.argtypes = [P, D, D, D, D]
        L.oracle_set_box.argtypes = [P, D, D, I32]
        L.oracle_set_projection_velocity.argtypes = [P, D, D]
        L.oracle_run.argtypes = [P, C.c_double, C.c_int64, C.c_int64]
        L.oracle_compute_contacts.argtypes = [P]
        L.oracle_compute_contacts.restype = C.c_int64
        L.oracle_n_contacts.argtypes = [P]
        L.oracle_n_contacts.restype = C.c_int64
        L.oracle_get_contacts.argtypes = [P, I64, D, D, D, D]
        L.oracle_amat.argtypes = [P, D, D]
        L.oracle_atmat.argtypes = [P, D, D]
        L.oracle_solve.argtypes = [P, C.c_double, D, D, D]
        L.oracle_gradient.argtypes = [P, C.c_double, D, D]
        L.oracle_set_contacts.argtypes = [P, C.c_int64, I64, D, D, D, D]
        L.oracle_set_a_priori_velocity.argtypes = [P, C.c_double]
        L.oracle_set_forces.argtypes = [P, D]
        L.oracle_update_contact_properties.argtypes = [P, C.c_double]
        L.oracle_lambda_size.argtypes = [P]
        L.oracle_lambda_size.restype = C.c_int64
        L.oracle_get_contact_gamma.argtypes = [P, D]
        L.oracle_set_contact_gamma.argtypes = [P, D]
        L.oracle_get_stats.argtypes = [P, C.POINTER(Stats)]
        L.oracle_closest_points.argtypes = [C.c_int32, C.c_int32, D, D, D, D, C.c_int32, D, D, D, D, D, D, D, D]
        L.oracle_se_closest_points.argtypes = [D] * 8 + [D, D, D, D, D, D, I32, I32, I32]
        L.oracle_se_residual.argtypes = [D] * 8 + [D, D, C.c_void_p]
        L.oracle_se_point_normal.argtypes = [D, D, D, D, C.c_double, C.c_double, D, D]
        L.oracle_set_se_math.argtypes = [C.c_int32]
        L.oracle_set_minpack_variant.argtypes = [C.c_int32]
        L.oracle_set_threads.argtypes = [C.c_int32]
        L.oracle_set_jacobian_ad.argtypes = [C.c_int32]
        L.oracle_dogleg_fresh_wa4_count.restype = C.c_int64
        L.oracle_gen_segment_demo.argtypes = [C.c_int32, C.c_double, D, D, D]
        L.oracle_gen_box_spheres_demo.argtypes = [C.c_int32, C.c_double, D, D, D, D, D]
        L.oracle_gen_uniform.argtypes = [C.c_int32, C.c_double, C.c_double, D]
        L.oracle_minpack_dogleg.argtypes = [C.c_int32, D, D, D, C.c_double, C.c_double, D, I32]
        L.oracle_newton_closest_points.argtypes = [C.c_int32, D, D, D, D, C.c_int32, D, D, D, D, D, D, D, D, D, D, I32]
        L.oracle_newton_residual.argtypes = [C.c_int32] + [D] * 8 + [D, C.c_void_p]
        L.oracle_lu_solve4.argtypes = [D, D, D]
        L.oracle_se2_closest_points.argtypes = [D, D, D, C.c_double, D, D, D, C.c_double, D, D, D, D, D, I32]
        L.oracle_se2_residual.argtypes = [D, D, D, C.c_double, D, D, D, C.c_double, D, C.c_void_p, C.c_void_p]
        L.oracle_math.argtypes = [C.c_int32, C.c_double, C.c_double, C.POINTER(C.c_double)]
        L.oracle_quaternion.argtypes = [C.c_double, D, D]
        L.oracle_gen_2d_case_spheres.argtypes = [C.c_int32, D, D, D, D]
        _lib = L
    return _lib


def default_params(**kw) -> Params:
    p = Params()
    lib().oracle_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Oracle:
    """One scene of the CPU restatement (ScopiSolver state + slot parameters)."""

    def __init__(self, scene, **params):
        self._L = lib()
        self._h = self._L.oracle_scene_new()
        self.dim = scene.dim
        self.n = scene.n
        self.n_inactive = scene.n_inactive
        self.problem = params.get("problem", NOFRICTION)
        self.set_params(**params)
        rc = self._L.oracle_scene_set(
            self._h, scene.dim, scene.n, scene.n_inactive, np.ascontiguousarray(scene.shape, np.int32),
            *(np.ascontiguousarray(getattr(scene, k), np.float64) for k in (
                "pos", "q", "size", "squareness", "v", "vd", "omega", "domega", "f", "m", "J")))
        if rc != 0:
            raise ValueError("oracle_scene_set failed")

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.oracle_scene_free(self._h)
            self._h = None

    def clone(self):
        """A copy of this scene, contact history and parameters included."""
        o = Oracle.__new__(Oracle)
        o._L, o.dim, o.n, o.n_inactive, o.problem = self._L, self.dim, self.n, self.n_inactive, self.problem
        o.params = default_params(**{k: getattr(self.params, k) for k, _ in Params._fields_})
        o._h = self._L.oracle_scene_clone(self._h)
        return o

    def set_params(self, **kw):
        self.params = default_params(**kw)
        self.problem = self.params.problem
        self._L.oracle_set_params(self._h, C.byref(self.params))

    def set_box(self, lo, hi, periodic):
        """BoxDomain(lo, hi).with_periodicity(axes with periodic[d] true)."""
        lo = np.ascontiguousarray(lo, np.float64)
        hi = np.ascontiguousarray(hi, np.float64)
        per = np.ascontiguousarray([1 if p else 0 for p in periodic], np.int32)
        assert lib().oracle_set_box(self._h, lo, hi, per) == 0

    def set_projection_velocity(self, u, w):
        """vap_projection::set_u_w (Na x 3 each)."""
        na = self.n - self.n_inactive
        assert lib().oracle_set_projection_velocity(self._h, np.ascontiguousarray(np.reshape(u, (na, 3)), np.float64),
                                                    np.ascontiguousarray(np.reshape(w, (na, 3)), np.float64)) == 0

    def run(self, dt, total_it, initial_iter=0):
        self._L.oracle_run(self._h, float(dt), int(total_it), int(initial_iter))

    def state(self):
        pos = np.zeros((self.n, 3)); q = np.zeros((self.n, 4)); v = np.zeros((self.n, 3)); om = np.zeros((self.n, 3))
        self._L.oracle_get_state(self._h, pos, q, v, om)
        return dict(pos=pos, q=q, v=v, omega=om)

    def compute_contacts(self):
        return int(self._L.oracle_compute_contacts(self._h))

    def contacts(self):
        nc = int(self._L.oracle_n_contacts(self._h))
        ij = np.zeros((nc, 2), np.int64); dij = np.zeros(nc); nij = np.zeros((nc, 3))
        pi = np.zeros((nc, 3)); pj = np.zeros((nc, 3))
        if nc:
            self._L.oracle_get_contacts(self._h, ij, dij, nij, pi, pj)
        return dict(ij=ij, dij=dij, nij=nij, pi=pi, pj=pj)

    def set_contacts(self, ct):
        """Replace the contact list with records {ij, dij, nij, pi, pj} (e.g. the device's)."""
        nc = len(ct["dij"])
        f = lambda a, w: np.ascontiguousarray(np.reshape(a, (nc, w)) if nc else np.zeros((1, w)), np.float64)
        ij = np.ascontiguousarray(np.reshape(ct["ij"], (nc, 2)) if nc else np.zeros((1, 2)), np.int64)
        dij = np.ascontiguousarray(ct["dij"] if nc else np.zeros(1), np.float64)
        assert self._L.oracle_set_contacts(self._h, nc, ij, dij, f(ct["nij"], 3), f(ct["pi"], 3), f(ct["pj"], 3)) == 0

    def amat(self, u):
        nc = int(self._L.oracle_n_contacts(self._h))
        out = np.zeros(3 * nc)
        self._L.oracle_amat(self._h, np.ascontiguousarray(u, np.float64), out)
        return out

    def atmat(self, f):
        out = np.zeros(6 * (self.n - self.n_inactive))
        self._L.oracle_atmat(self._h, np.ascontiguousarray(f, np.float64), out)
        return out

    def set_a_priori_velocity(self, dt):
        self._L.oracle_set_a_priori_velocity(self._h, float(dt))

    def solve(self, dt):
        nc = int(self._L.oracle_n_contacts(self._h))
        na = self.n - self.n_inactive
        lam = np.zeros(max(1, 4 * nc))  # the largest layout: four rows per contact
        u = np.zeros(max(1, 3 * na)); w = np.zeros(max(1, 3 * na))
        self._L.oracle_solve(self._h, float(dt), lam, u, w)
        nl = int(self._L.oracle_lambda_size(self._h))
        return dict(lam=lam[:nl], u=u[: 3 * na].reshape(na, 3), w=w[: 3 * na].reshape(na, 3))

    def gradient(self, dt, lam):
        """Q lambda + c of the optim slot's problem (current contacts and a-priori velocities)."""
        assert self.problem in (NOFRICTION, FRICTION)
        nl = int(self._L.oracle_n_contacts(self._h)) * (3 if self.problem == FRICTION else 1)
        lam = np.ascontiguousarray(lam, np.float64)
        assert lam.shape == (nl,), (lam.shape, nl)
        out = np.zeros(max(1, nl))
        assert self._L.oracle_gradient(self._h, float(dt), lam if nl else np.zeros(1), out) == 0
        return out[:nl]

    def update_contact_properties(self, dt):
        """The contacts' gamma from the last solve's multipliers (Viscous problems)."""
        assert self._L.oracle_update_contact_properties(self._h, float(dt)) == 0

    def set_forces(self, f):
        """Replace the external forces (n x 3); the last step's contacts and properties stay."""
        assert self._L.oracle_set_forces(self._h, np.ascontiguousarray(np.reshape(f, (self.n, 3)), np.float64)) == 0

    def contact_gamma(self):
        g = np.zeros(max(1, int(self._L.oracle_n_contacts(self._h))))
        assert self._L.oracle_get_contact_gamma(self._h, g) == 0
        return g[: int(self._L.oracle_n_contacts(self._h))]

    def set_contact_gamma(self, gamma):
        nc = int(self._L.oracle_n_contacts(self._h))
        g = np.ascontiguousarray(np.broadcast_to(np.asarray(gamma, np.float64), (nc,)))
        assert self._L.oracle_set_contact_gamma(self._h, g if nc else np.zeros(1)) == 0

    def stats(self):
        s = Stats()
        self._L.oracle_get_stats(self._h, C.byref(s))
        return {k: getattr(s, k) for k, _ in Stats._fields_}


def _v(x, n):
    o = np.zeros(n); x = np.atleast_1d(np.asarray(x, np.float64)); o[: x.size] = x
    return o


def closest_points(dim, shape_i, pos_i, q_i, size_i, shape_j, pos_j, q_j, size_j, sq_i=(1.0, 1.0),
                   sq_j=(1.0, 1.0)):
    L = lib()
    dij = np.zeros(1); nij = np.zeros(3); pi = np.zeros(3); pj = np.zeros(3)
    rc = L.oracle_closest_points(dim, shape_i, _v(pos_i, 3), np.ascontiguousarray(q_i, np.float64), _v(size_i, 3),
                                 _v(sq_i, 2), shape_j, _v(pos_j, 3), np.ascontiguousarray(q_j, np.float64),
                                 _v(size_j, 3), _v(sq_j, 2), dij, nij, pi, pj)
    if rc != 0:
        raise ValueError("no closest-point form for this pair")
    return dict(dij=float(dij[0]), nij=nij[:dim], pi=pi[:dim], pj=pj[:dim])


def se_closest_points(pos_i, q_i, size_i, sq_i, pos_j, q_j, size_j, sq_j):
    """closest_points(superellipsoid<3>, superellipsoid<3>) with hybrd1 diagnostics."""
    L = lib()
    dij = np.zeros(1); nij = np.zeros(3); pi = np.zeros(3); pj = np.zeros(3); u0 = np.zeros(4); u = np.zeros(4)
    info = np.zeros(1, np.int32); nfev = np.zeros(1, np.int32); njev = np.zeros(1, np.int32)
    L.oracle_se_closest_points(_v(pos_i, 3), _v(q_i, 4), _v(size_i, 3), _v(sq_i, 2), _v(pos_j, 3), _v(q_j, 4),
                               _v(size_j, 3), _v(sq_j, 2), dij, nij, pi, pj, u0, u, info, nfev, njev)
    return dict(dij=float(dij[0]), nij=nij, pi=pi, pj=pj, u0=u0, u=u, info=int(info[0]), nfev=int(nfev[0]),
                njev=int(njev[0]))


def se_residual(pos_i, q_i, size_i, sq_i, pos_j, q_j, size_j, sq_j, u, jacobian=False):
    L = lib()
    F = np.zeros(4); J = np.zeros(16)
    L.oracle_se_residual(_v(pos_i, 3), _v(q_i, 4), _v(size_i, 3), _v(sq_i, 2), _v(pos_j, 3), _v(q_j, 4),
                         _v(size_j, 3), _v(sq_j, 2), _v(u, 4), F, J.ctypes.data_as(C.c_void_p) if jacobian else None)
    return (F, J.reshape(4, 4)) if jacobian else F


def newton_closest_points(shape_i, pos_i, q_i, size_i, sq_i, shape_j, pos_j, q_j, size_j, sq_j):
    """closest_points of a sphere or plane against a superellipsoid (dispatcher order) + diagnostics."""
    L = lib()
    dij = np.zeros(1); nij = np.zeros(3); pi = np.zeros(3); pj = np.zeros(3); u0 = np.zeros(4); u = np.zeros(4)
    it = np.zeros(1, np.int32)
    rc = L.oracle_newton_closest_points(shape_i, _v(pos_i, 3), _v(q_i, 4), _v(size_i, 3), _v(sq_i, 2), shape_j,
                                        _v(pos_j, 3), _v(q_j, 4), _v(size_j, 3), _v(sq_j, 2), dij, nij, pi, pj, u0, u,
                                        it)
    if rc != 0:
        raise ValueError("not a sphere/plane - superellipsoid pair")
    return dict(dij=float(dij[0]), nij=nij, pi=pi, pj=pj, u0=u0, u=u, iterations=int(it[0]))


def newton_residual(kind, pos_i, q_i, size_i, sq_i, pos_j, q_j, size_j, u, jacobian=False):
    L = lib()
    F = np.zeros(4); J = np.zeros(16)
    L.oracle_newton_residual(kind, _v(pos_i, 3), _v(q_i, 4), _v(size_i, 3), _v(sq_i, 2), _v(pos_j, 3), _v(q_j, 4),
                             _v(size_j, 3), _v(u, 4), F, J.ctypes.data_as(C.c_void_p) if jacobian else None)
    return (F, J.reshape(4, 4)) if jacobian else F


def lu_solve4(A, b):
    x = np.zeros(4)
    lib().oracle_lu_solve4(np.ascontiguousarray(A, np.float64).ravel(), np.ascontiguousarray(b, np.float64), x)
    return x


def set_se_math(portable: bool):
    """Select the elementary functions of the superellipsoid solve (C library or portable)."""
    lib().oracle_set_se_math(1 if portable else 0)


def set_threads(n: int) -> int:
    """Threads of the reference's OpenMP loops in the restatement (search, vap, update and move)."""
    return int(lib().oracle_set_threads(int(n)))


MINPACK_ANL_ENORM, MINPACK_ANL_GIANT, MINPACK_ANL_DOGLEG, MINPACK_ANL_TERMINATION = 1, 2, 4, 8
MINPACK_ANL = 15


def set_minpack_variant(mask: int):
    """Test-only: 0 = the reference's src/minpack.cpp (default); each bit takes one ANL-80-74 form
    (MINPACK_ANL_* ; MINPACK_ANL = all of them)."""
    lib().oracle_set_minpack_variant(int(mask))


def set_jacobian_ad(ad: bool):
    """Test-only: forward-mode Jacobian (True) or the reference's hand-written one (False, default)."""
    lib().oracle_set_jacobian_ad(1 if ad else 0)


def minpack_dogleg(r_packed, diag, qtb, delta, wa4_last):
    """One dogleg step of the restated MINPACK; returns (x, took_gradient_branch)."""
    n = len(diag)
    x = np.zeros(n); g = np.zeros(1, np.int32)
    lib().oracle_minpack_dogleg(n, np.ascontiguousarray(r_packed, np.float64), np.ascontiguousarray(diag, np.float64),
                                np.ascontiguousarray(qtb, np.float64), float(delta), float(wa4_last), x, g)
    return x, bool(g[0])


def dogleg_fresh_wa4_count() -> int:
    """Scaled-gradient dogleg steps before hybrd's first trial point (reads 0 where the reference reads
    an uninitialised work element)."""
    return int(lib().oracle_dogleg_fresh_wa4_count())


MATH_FN = {"cos": 0, "sin": 1, "asin": 2, "log": 3, "exp": 4, "pow": 5}


def portable_math(fn, x, y=0.0):
    out = C.c_double()
    lib().oracle_math(MATH_FN[fn], float(x), float(y), C.byref(out))
    return out.value


def se_point_normal(pos, q, size, sq, a, b):
    L = lib()
    p = np.zeros(3); n = np.zeros(3)
    L.oracle_se_point_normal(_v(pos, 3), _v(q, 4), _v(size, 3), _v(sq, 2), float(a), float(b), p, n)
    return p, n


def gen_2d_case_spheres(n=3):
    L = lib()
    k = 2 * n * n
    x = np.zeros(k); y = np.zeros(k); r = np.zeros(k); vel = np.zeros(k)
    L.oracle_gen_2d_case_spheres(n, x, y, r, vel)
    return x, y, r, vel


def se2_residual(pos_i, q_i, size_i, e_i, pos_j, q_j, size_j, e_j, u, jacobian=False):
    """Residual (and Jacobian) of the SE2-SE2 system at u = (b1, b2)."""
    v = lambda x, n: np.ascontiguousarray(np.asarray(x, np.float64)[:n])
    F = np.zeros(2); J = np.zeros(4)
    lib().oracle_se2_residual(v(pos_i, 2), v(q_i, 4), v(size_i, 2), C.c_double(e_i), v(pos_j, 2), v(q_j, 4),
                              v(size_j, 2), C.c_double(e_j), v(u, 2), F.ctypes.data_as(C.c_void_p),
                              J.ctypes.data_as(C.c_void_p) if jacobian else None)
    return (F, J.reshape(2, 2)) if jacobian else F


def gen_segment_demo(n=50, max_radius=0.06):
    """The random draws (x, y, radius) of demos/segment.cpp's spheres, std::default_random_engine."""
    x = np.zeros(n); y = np.zeros(n); r = np.zeros(n)
    lib().oracle_gen_segment_demo(n, max_radius, x, y, r)
    return x, y, r


def gen_uniform(n, a, b):
    """n draws of uniform_real_distribution<double>(a, b) from a default std::default_random_engine."""
    out = np.zeros(n)
    lib().oracle_gen_uniform(n, a, b, out)
    return out


def gen_box_spheres_demo(n_side=10, r0=0.5):
    """The random draws of demos/box_spheres_3d.cpp (radius, mass, dx, dy, dz per sphere, in the demo's
    i, j, k loop order), std::default_random_engine."""
    n = n_side ** 3
    out = [np.zeros(n) for _ in range(5)]
    lib().oracle_gen_box_spheres_demo(n_side, r0, *out)
    return tuple(out)


def se2_closest_points(pos_i, q_i, size_i, e_i, pos_j, q_j, size_j, e_j):
    """closest_points(superellipsoid<2>, superellipsoid<2>) with hybrd1's root and info."""
    v = lambda x, n: np.ascontiguousarray(np.asarray(x, np.float64)[:n])
    dij = np.zeros(1); nij = np.zeros(2); pi = np.zeros(2); pj = np.zeros(2); u = np.zeros(2)
    info = np.zeros(1, np.int32)
    lib().oracle_se2_closest_points(v(pos_i, 2), v(q_i, 4), v(size_i, 2), float(e_i), v(pos_j, 2), v(q_j, 4),
                                    v(size_j, 2), float(e_j), dij, nij, pi, pj, u, info)
    return dict(dij=float(dij[0]), nij=nij, pi=pi, pj=pj, u=u, info=int(info[0]))
