"""ctypes bindings of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: it is the checker the product library is
compared against, never the thing measured or shipped.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

from tgh_desc import TghDesc, TghFit  # noqa: E402  (tests/ is on sys.path under pytest)

_lib = None
dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        D = C.c_double
        Lg = C.c_long
        desc = C.POINTER(TghDesc)
        sig = {
            "orc_derivative_weights": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), dp]),
            "orc_dissipation_weights": (C.c_int, [C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), dp]),
            "orc_metric_quantities": (None, [dp, dp]),
            "orc_metric_quantities_jvp": (None, [dp, dp, dp]),
            "orc_gh_node_source": (None, [dp, dp, dp, dp, D, D, dp]),
            "orc_gh_node_source_lean": (None, [dp, dp, dp, dp, dp, D, D, dp]),
            "orc_metric_divergences": (None, [dp, dp, dp]),
            "orc_metric_derivatives": (None, [dp, dp, dp]),
            "orc_gh_node_rhs_expanded": (None, [dp, dp, dp, dp, dp, dp, dp, D, D, dp, dp]),
            "orc_gh_node_rhs_flux": (None, [dp, dp, dp, dp, dp, D, D, dp, dp, dp]),
            "orc_matter_source": (None, [dp, D, dp, dp, dp]),
            "orc_ks_algebraic_source": (None, [D, dp, dp, dp, dp, dp, dp, dp]),
            "orc_dmetric": (None, [desc, D, dp, dp, dp]),
            "orc_background_state": (None, [desc, D, dp, dp, dp, dp]),
            "orc_gauge_source_grad": (None, [desc, D, dp, dp, dp]),
            "orc_case_state": (None, [desc, D, dp, dp]),
            "orc_point_draw": (D, [C.c_uint64, dp, C.c_int]),
            "orc_smoothstep": (D, [D]),
            "orc_problem_create": (P, [desc]),
            "orc_problem_create_forest": (P, [desc, C.POINTER(C.c_int), C.c_int, C.c_int]),
            "orc_block_level": (C.c_int, [P, Lg]),
            "orc_set_target": (None, [P, dp, D, C.c_int, D]),
            "orc_gauge_constraint": (None, [P, dp, D, dp, dp]),
            "orc_adm_constraint": (None, [P, dp, D, dp, dp, dp]),
            "orc_tau_blocks": (None, [P, dp, D, D, D, dp]),
            "orc_field_scale": (D, [P, dp, D]),
            "orc_transfer": (C.c_int, [P, P, dp, D, dp]),
            "orc_shape_series": (D, [dp, C.c_int, dp]),
            "orc_interior_geo": (None, [desc, D, dp, dp]),
            "orc_forest_classify": (C.c_int, [P, Lg, Lg, Lg, Lg, C.POINTER(Lg)]),
            "orc_prolong_weights": (None, [P, dp]),
            "orc_problem_destroy": (None, [P]),
            "orc_state_length": (Lg, [P]),
            "orc_work_length": (Lg, [P]),
            "orc_work": (C.POINTER(D), [P]),
            "orc_block_coords": (None, [P, C.POINTER(C.c_int)]),
            "orc_fill_exact": (None, [P, dp, D]),
            "orc_add_noise": (None, [P, dp, C.c_uint64, D]),
            "orc_prepare": (None, [P, dp, D]),
            "orc_rhs": (None, [P, dp, D, C.c_int]),
            "orc_gh_rhs": (None, [P, dp, dp, D, C.c_int]),
            "orc_host_rhs_at": (None, [P, Lg, Lg, Lg, Lg, D, dp]),
            "orc_rhs_scale": (D, [P, Lg, Lg, Lg, Lg]),
            "orc_rk4_step": (None, [P, dp, D, D, C.c_int]),
            "orc_settle": (None, [P, dp, D]),
            "orc_stage_limit": (None, [P, dp, D]),
            "orc_apply_bounds": (C.c_long, [P, dp, D]),
            "orc_bounds_project": (C.c_int, [dp, dp, dp, dp, dp]),
            "orc_sym_eigen3": (None, [dp, dp, dp]),
            "orc_max_speed": (D, [P, dp, D]),
            "orc_error_norms": (None, [P, dp, D, C.POINTER(D), C.POINTER(D)]),
            "orc_error_rows": (None, [P, dp, D, D, D, dp]),
            "orc_set_rho_max": (None, [P, D]),
            "orc_set_matter": (None, [P, P, P, P]),
            "orc_parity_odd_mask": (C.c_int, [C.c_int]),
            "orc_rotation": (C.c_int, [C.c_int]),
            "orc_is_primary": (C.c_int, [P, Lg, Lg, Lg, Lg]),
            "orc_count_rhs_flops": (D, [P, Lg, Lg, Lg, Lg, C.c_int]),
            "orc_max_threads": (C.c_int, []),
            "orc_fit_variables_at": (None, [C.POINTER(TghFit), dp, D, dp]),
            "orc_state_from_fit": (None, [dp, C.c_int, dp, dp]),
            "orc_fit_state": (None, [C.POINTER(TghFit), dp, D, dp, dp]),
            "orc_fill_target": (None, [P, dp, C.POINTER(TghFit), C.POINTER(TghFit), D, D, C.c_int, C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def weights(q, m):
    n = q + 1
    num = (C.c_longlong * n)()
    den = (C.c_longlong * n)()
    w = np.zeros(n)
    lib().orc_derivative_weights(q, m, num, den, w)
    return list(num), list(den), w


def diss_weights(r):
    n = 2 * r + 1
    num = (C.c_longlong * n)()
    den = (C.c_longlong * n)()
    w = np.zeros(n)
    lib().orc_dissipation_weights(r, num, den, w)
    return list(num), list(den), w


def metric_quantities(h):
    out = np.zeros(62)
    lib().orc_metric_quantities(np.ascontiguousarray(h, dtype=np.float64), out)
    return dict(alpha=out[0], beta=out[1:4], sqrtg=out[4], gamu=out[5:14].reshape(3, 3),
                gu4=out[14:30].reshape(4, 4), guo=out[30:46].reshape(4, 4), g4=out[46:62].reshape(4, 4))


class OracleProblem:
    """The oracle's uniform-forest problem for one case descriptor."""

    def __init__(self, desc, leaves=None, prolongation=None):
        """leaves: None (the uniform brick of desc.nblocks) or (level, i, j, k)
        quadruples of a frozen hierarchy, with the prolongation order."""
        self.desc = desc
        if leaves is None:
            self.p = lib().orc_problem_create(C.byref(desc))
        else:
            flat = [int(x) for lf in leaves for x in lf]
            arr = (C.c_int * len(flat))(*flat)
            self.p = lib().orc_problem_create_forest(C.byref(desc), arr, len(leaves), int(prolongation))
        self.n = lib().orc_state_length(self.p)

    def set_target(self, tw, t_f, rate=False, trail=0.0):
        """The :fitted target cache (40 variables per owned point, A then S)."""
        tw = np.ascontiguousarray(tw, dtype=np.float64)
        assert tw.size == 2 * self.n
        lib().orc_set_target(self.p, tw, float(t_f), 1 if rate else 0, float(trail))

    def fill_target(self, out, fa, fb, t, h, rate=False, clear=True):
        """fill_target! into the host cache out (40 variables per owned point)."""
        assert out.size == 2 * self.n and out.dtype == np.float64 and out.flags.c_contiguous
        lib().orc_fill_target(self.p, out, C.byref(fa), C.byref(fb) if fb is not None else None, float(t),
                              float(h), 1 if rate else 0, 1 if clear else 0)
        return out

    def gauge_constraint(self, u, t):
        """(L2[4], Linf[4]) of the gauge constraint C_a over the evolved points."""
        l2, li = np.zeros(4), np.zeros(4)
        lib().orc_gauge_constraint(self.p, np.ascontiguousarray(u), float(t), l2, li)
        return l2, li

    def adm_constraint(self, u, t):
        """((L2, Linf) of H, L2[3], Linf[3] of M_i) over the evolved points."""
        ham, ml2, mli = np.zeros(2), np.zeros(3), np.zeros(3)
        lib().orc_adm_constraint(self.p, np.ascontiguousarray(u), float(t), ham, ml2, mli)
        return ham, ml2, mli

    def tau_blocks(self, u, t, scale, eps=0.01):
        nb = self.n // ((self.desc.N + 1) ** 3 * 20)
        out = np.zeros(nb)
        lib().orc_tau_blocks(self.p, np.ascontiguousarray(u), float(t), float(scale), float(eps), out)
        return out

    def field_scale(self, u, t):
        return lib().orc_field_scale(self.p, np.ascontiguousarray(u), float(t))

    def transfer_from(self, other, u_other, t):
        """This forest's state carried over from other's (the regrid's transfer)."""
        out = self.zeros()
        if lib().orc_transfer(self.p, other.p, np.ascontiguousarray(u_other), float(t), out):
            raise RuntimeError("no transfer between these forests (a leaf two levels finer than the old forest)")
        return out

    def block_level(self, b):
        return lib().orc_block_level(self.p, b)

    def classify(self, b, i, j, k):
        """(kind, (block, l0, l1, l2, half mask, mirror | turns << 3)): kind 0
        primary, 1 copy, 2 Dirichlet, 3 prolongation."""
        src = (C.c_long * 6)()
        c = lib().orc_forest_classify(self.p, b, i, j, k, src)
        return c, tuple(src)

    def prolong_weights(self, p):
        w = np.zeros(p)
        lib().orc_prolong_weights(self.p, w)
        return w

    def __del__(self):
        if getattr(self, "p", None):
            lib().orc_problem_destroy(self.p)
            self.p = None

    def zeros(self):
        return np.zeros(self.n)

    def block_coords(self):
        nb = self.n // ((self.desc.N + 1) ** 3 * 20)
        out = (C.c_int * (3 * nb))()
        lib().orc_block_coords(self.p, out)
        return np.array(list(out)).reshape(nb, 3)

    def fill_exact(self, t):
        u = self.zeros()
        lib().orc_fill_exact(self.p, u, t)
        return u

    def add_noise(self, u, base, amp):
        lib().orc_add_noise(self.p, u, base, amp)

    def prepare(self, u, t):
        lib().orc_prepare(self.p, u, t)

    def work(self):
        n = lib().orc_work_length(self.p)
        ptr = lib().orc_work(self.p)
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    def gh_rhs(self, u, t, nthreads=1):
        du = self.zeros()
        lib().orc_gh_rhs(self.p, du, u, t, nthreads)
        return du

    def rk4_step(self, u, t, dt, nthreads=1):
        lib().orc_rk4_step(self.p, u, t, dt, nthreads)

    def settle(self, u, t):
        lib().orc_settle(self.p, u, t)

    def stage_limit(self, u, t):
        lib().orc_stage_limit(self.p, u, t)

    def apply_bounds(self, u, t=0.0):
        return lib().orc_apply_bounds(self.p, u, t)

    def max_speed(self, u, t=0.0):
        return lib().orc_max_speed(self.p, u, t)

    def set_matter(self, rho, S, Sij):
        """Matter fields (host arrays in the device layout) every right-hand
        side adds; None clears them."""
        if rho is None:
            lib().orc_set_matter(self.p, None, None, None)
            return
        self._mat = [np.ascontiguousarray(x, dtype=np.float64) for x in (rho, S, Sij)]
        lib().orc_set_matter(self.p, *[x.ctypes.data for x in self._mat])

    def set_rho_max(self, rho):
        lib().orc_set_rho_max(self.p, float(rho))

    def error_rows(self, u, t, shell=(0.0, -1.0)):
        """(err_l2, err_linf, residual, drift): gh_error! and error_norms."""
        out = np.zeros(4)
        lib().orc_error_rows(self.p, u, t, float(shell[0]), float(shell[1]), out)
        return tuple(float(x) for x in out)

    def error_norms(self, u, t):
        l2 = C.c_double()
        li = C.c_double()
        lib().orc_error_norms(self.p, u, t, C.byref(l2), C.byref(li))
        return l2.value, li.value

    def host_rhs_at(self, b, i, j, k, t):
        out = np.zeros(20)
        lib().orc_host_rhs_at(self.p, b, i, j, k, t, out)
        return out

    def rhs_scale(self, b, i, j, k):
        return lib().orc_rhs_scale(self.p, b, i, j, k)

    def count_flops(self, b, i, j, k, head_only=False):
        return lib().orc_count_rhs_flops(self.p, b, i, j, k, 1 if head_only else 0)


def bounds_project(h, P, bnd):
    """(h', P', hit, nonfinite) of the oracle's range projection at one point;
    bnd = (alpha_min, alpha_max, lambda_min, lambda_max, beta_max, K_max, r_gate)."""
    ho, Po = np.zeros(10), np.zeros(10)
    r = lib().orc_bounds_project(np.ascontiguousarray(h, dtype=np.float64), np.ascontiguousarray(P, dtype=np.float64),
                                 np.ascontiguousarray(bnd, dtype=np.float64), ho, Po)
    return ho, Po, bool(r & 1), bool(r & 2)


def sym_eigen3(A):
    lam, V = np.zeros(3), np.zeros(9)
    lib().orc_sym_eigen3(np.ascontiguousarray(A, dtype=np.float64).ravel(), lam, V)
    return lam, V.reshape(3, 3)


def shape_series(shape, lmax, n):
    sh = np.zeros(81)
    sh[:len(shape)] = shape
    return lib().orc_shape_series(sh, int(lmax), np.ascontiguousarray(n, dtype=np.float64))


def interior_geo(desc, t, x):
    out = np.zeros(3)
    lib().orc_interior_geo(C.byref(desc), float(t), np.ascontiguousarray(x, dtype=np.float64), out)
    return out


def evolve(orc, u, *, hmin, t_end, chunk, cfl=0.25, rho_max_factor=None, shell=(0.0, -1.0), nthreads=8):
    """evolve!'s chunk loop on the oracle (src/driver.jl:1732-1800; chunk_count
    and chunk_bounds, src/precision.jl:81-100): per chunk λ at its start, dt =
    cfl h_min / λ, steps = ⌈span / dt⌉ of span / steps, the layer's rate
    factor / dt when rho_max_factor is given (else the problem's own), and the
    record at the chunk's end: (t, steps, err_l2, err_linf, residual, drift,
    gauge_l2 = the largest component L2 of C_a)."""
    import math
    r = t_end / chunk
    m = round(r)
    nchunks = int(m) if m >= 1 and abs(r - m) <= 4 * (math.ulp(r) + math.ulp(t_end) / chunk) else math.ceil(r)
    recs = []
    for c in range(1, nchunks + 1):
        a = min((c - 1) * chunk, t_end)
        b = t_end if c == nchunks else min(c * chunk, t_end)
        lam = orc.max_speed(u, a)
        steps = max(1, math.ceil((b - a) / (cfl * hmin / lam)))
        dt = (b - a) / steps
        if rho_max_factor is not None:
            orc.set_rho_max(rho_max_factor / dt)
        for s in range(steps):
            orc.rk4_step(u, a + s * dt, dt, nthreads)
        rows = orc.error_rows(u, b, shell)
        gl2, _ = orc.gauge_constraint(u, b)
        recs.append({"t": b, "steps": steps, "err_l2": rows[0], "err_linf": rows[1], "residual": rows[2],
                     "drift": rows[3], "gauge_l2": float(max(gl2))})
    return recs


def _jvp_sigs():
    L = lib()
    P, D = C.c_void_p, C.c_double
    L.orc_gh_rhs_jvp.restype = C.c_int
    L.orc_gh_rhs_jvp.argtypes = [P, dp, dp, D, dp, dp, dp, C.c_int]
    L.orc_error_tangent.restype = C.c_double
    L.orc_error_tangent.argtypes = [P, dp, dp, D]
    L.orc_rk4_step_jvp.restype = C.c_int
    L.orc_rk4_step_jvp.argtypes = [P, dp, dp, D, D, dp, C.c_int]
    return L


def gh_rhs_jvp(orc, u, udot, t, pd, nthreads=8):
    """(F(u), its tangent along udot and the parameter tangents pd =
    d(gamma0_near, gamma0_far, gamma2, rho_max, rho_ramp)) by the dual pass."""
    L = _jvp_sigs()
    du, dd = np.zeros_like(u), np.zeros_like(u)
    rc = L.orc_gh_rhs_jvp(orc.p, np.ascontiguousarray(u), np.ascontiguousarray(udot), float(t),
                          np.ascontiguousarray(pd, dtype=np.float64), du, dd, nthreads)
    if rc:
        raise RuntimeError("the oracle's tangent pass does not cover this case")
    return du, dd


def rk4_step_jvp(orc, u, udot, t, dt, pd, nthreads=8):
    L = _jvp_sigs()
    rc = L.orc_rk4_step_jvp(orc.p, u, udot, float(t), float(dt), np.ascontiguousarray(pd, dtype=np.float64),
                            nthreads)
    if rc:
        raise RuntimeError("the oracle's tangent pass does not cover this case")


def error_tangent(orc, u, udot, t):
    """The directional derivative of the masked L2 error along udot."""
    return float(_jvp_sigs().orc_error_tangent(orc.p, np.ascontiguousarray(u), np.ascontiguousarray(udot), float(t)))


def fit_variables_at(f, x, t):
    out = np.zeros(20)
    lib().orc_fit_variables_at(C.byref(f), np.ascontiguousarray(x, dtype=np.float64), float(t), out)
    return out


def state_from_fit(v, tilde=False):
    h, P = np.zeros(10), np.zeros(10)
    lib().orc_state_from_fit(np.ascontiguousarray(v, dtype=np.float64), 1 if tilde else 0, h, P)
    return h, P


def fit_state(f, x, t):
    h, P = np.zeros(10), np.zeros(10)
    lib().orc_fit_state(C.byref(f), np.ascontiguousarray(x, dtype=np.float64), float(t), h, P)
    return h, P
