"""MI355X drop-in for the right-hand side verbs of TreeGeneralizedHarmonic.jl.

The host side above the C ABI of ``libtgh_mi355x.so`` (``include/tgh.h``), in
Python because the reference's own host language (Julia) is not installed
here. It mirrors the reference's interface for this path, with the same names,
argument meaning and errors (``!`` dropped from the verbs):

* cases:   ``minkowski_case``, ``gauge_wave_case``, ``kerr_schild_case``
           (src/initialdata.jl:869-1326), each float-type-first;
* forest:  ``gh_forest`` (src/initialdata.jl:1370-1400), a uniform brick;
* problem: ``GHProblem(forest, case; q, backend=MI355X())``
           (src/evolution.jl:1034-1149);
* verbs:   ``gh_rhs``, ``prepare``, ``rhs``, ``settle``, ``stage_limit``,
           ``step_limit``, ``max_speed``, ``initial_state``, ``evolved_sets``,
           ``gh_dt``, ``gh_solve`` (src/evolution.jl:1327-1560,
           src/stepping.jl:91-239).

State vectors are ``torch`` float64 CUDA tensors of ``tgh_state_length``
doubles in the reference's state layout; the library reads and writes them in
place through their device pointers. There is no CPU fallback: without the
built library or without a GPU every call raises.
"""
import ctypes as C
import math
import os

__all__ = [
    "ArgumentError", "TGHError", "MI355X", "Minkowski", "GaugeWave", "KerrSchild", "Harmonic",
    "ConstantDamping", "GaussianDamping", "Interior", "GHCase", "minkowski_case",
    "gauge_wave_case", "kerr_schild_case", "harmonic_kerr_case", "gh_forest", "GHProblem", "gh_rhs", "prepare",
    "rhs", "settle", "stage_limit", "step_limit", "max_speed", "initial_state",
    "evolved_sets", "gh_dt", "gh_solve", "add_noise", "error_norms", "library_path", "GHBatch",
    "gh_solve_batch", "Forest", "hole_forest", "Operators", "transfer", "regrid_flags", "regrid_forest", "regrid", "gh_solve_regrid",
    "field_scale", "refinement_indicator", "Refinement", "level_bounds", "refine_flags", "apply_bounds",
    "gauge_constraint_norms", "gauge_constraint_norms_noalloc", "adm_constraint_norms", "autotune",
    "set_matter", "set_rho_max", "vjp_rhs", "adjoint", "ADJOINT_PARAMS", "evolve", "chunk_count",
    "chunk_bounds", "FittedInterior", "StateBounds", "HorizonDissipation", "horizon_dissipation",
    "horizon_min_radius", "horizon_max_radius", "default_bounds", "default_gate", "with_interior",
    "record_norms", "minkowski_octant_case", "FitParams", "InteriorFit", "target_cache", "fill_target",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_NAME = "libtgh_mi355x.so"


def library_path():
    return os.path.join(_HERE, _LIB_NAME)


class ArgumentError(ValueError):
    """The reference's ArgumentError: an argument it refuses, with its reason."""


class TGHError(RuntimeError):
    """A device-side failure reported by the library."""


# --- the C ABI ---------------------------------------------------------------

TGH_OK, TGH_EINVAL, TGH_EDEVICE, TGH_ENOTFINITE, TGH_EUNSUPPORTED = 0, 1, 2, 3, 4
TGH_RK4_GRAPH = 1
TGH_RK4_FUSED = 2
TGH_RK4_BLOCKS = 4
TGH_RK4_SPLIT = 8
TGH_RK4_PER_BLOCK = 16


class TghDesc(C.Structure):
    """``tgh_desc`` of include/tgh.h."""
    _fields_ = [
        ("float_bits", C.c_int32), ("q", C.c_int32), ("N", C.c_int32),
        ("nblocks", C.c_int32 * 3), ("periodic", C.c_int32 * 3),
        ("background", C.c_int32), ("gauge", C.c_int32), ("damping", C.c_int32),
        ("interior", C.c_int32), ("dissipation", C.c_int32),
        ("lo", C.c_double * 3), ("hi", C.c_double * 3),
        ("M", C.c_double), ("a", C.c_double), ("center", C.c_double * 3),
        ("wave_amp", C.c_double), ("wave_len", C.c_double),
        ("gamma0", C.c_double), ("gamma0_near", C.c_double), ("gamma0_far", C.c_double),
        ("gamma0_width", C.c_double), ("gamma2", C.c_double), ("eps_ko", C.c_double),
        ("r0", C.c_double), ("r1", C.c_double), ("rho_max", C.c_double),
        ("w_ramp", C.c_double), ("rho_ramp", C.c_double),
        ("eps_in", C.c_double), ("eps_r1", C.c_double), ("eps_rh", C.c_double),
        ("alpha_min", C.c_double), ("alpha_max", C.c_double), ("lambda_min", C.c_double),
        ("lambda_max", C.c_double), ("beta_max", C.c_double), ("K_max", C.c_double), ("r_gate", C.c_double),
        ("velocity", C.c_double * 3),
        ("shaped", C.c_int32), ("shape_lmax", C.c_int32), ("shape", C.c_double * 81),
        ("r_in", C.c_double), ("r_out", C.c_double), ("offset", C.c_double), ("thickness", C.c_double),
        ("ownership", C.c_int32), ("reflecting", (C.c_int32 * 2) * 3), ("rotating", C.c_int32),
    ]


class TghMatter(C.Structure):
    _fields_ = [("rho", C.c_void_p), ("S", C.c_void_p), ("Sij", C.c_void_p)]


class TghFit(C.Structure):
    """include/tgh.h's tgh_fit: a fitted target as the cache's fill reads it."""
    _fields_ = [
        ("lmax", C.c_int32), ("cont", C.c_int32), ("tilde", C.c_int32), ("pad", C.c_int32),
        ("rbar", C.c_double), ("center", C.c_double * 3), ("velocity", C.c_double * 3),
        ("alpha_min", C.c_double), ("alpha_max", C.c_double), ("lambda_min", C.c_double),
        ("lambda_max", C.c_double), ("beta_max", C.c_double), ("K_max", C.c_double),
        ("t", C.c_double), ("coeffs", C.c_void_p),
    ]


_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise TGHError(f"{path} is not built: run __graft_entry__.build() (hipcc for gfx950) first; "
                       "this package has no CPU fallback")
    L = C.CDLL(path)
    P = C.c_void_p
    D = C.c_double
    sig = {
        "tgh_problem_create": (C.c_int, [C.POINTER(TghDesc), C.POINTER(P)]),
        "tgh_problem_create_forest": (C.c_int, [C.POINTER(TghDesc), C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                                                C.POINTER(P)]),
        "tgh_block_levels": (C.c_int, [P, C.POINTER(C.c_int32)]),
        "tgh_set_target": (C.c_int, [P, C.c_void_p, C.c_double, C.c_int32, C.c_double]),
        "tgh_fill_target": (C.c_int, [P, C.c_void_p, C.POINTER(TghFit), C.POINTER(TghFit), C.c_double,
                                      C.c_double, C.c_int32, C.c_int32]),
        "tgh_autotune": (C.c_int, [P, C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_double)]),
        "tgh_set_tile": (C.c_int, [P, C.c_int32]),
        "tgh_gauge_constraint": (C.c_int, [P, C.c_void_p, C.c_double, C.POINTER(C.c_double),
                                           C.POINTER(C.c_double)]),
        "tgh_gauge_constraint_noalloc": (C.c_int, [P, C.c_void_p, C.c_double, C.POINTER(C.c_double),
                                                   C.POINTER(C.c_double)]),
        "tgh_adm_constraint": (C.c_int, [P, C.c_void_p, C.c_double, C.POINTER(C.c_double),
                                         C.POINTER(C.c_double)]),
        "tgh_refinement_indicator": (C.c_int, [P, C.c_void_p, C.c_double, C.c_double, C.c_double,
                                               C.POINTER(C.c_double)]),
        "tgh_field_scale": (C.c_int, [P, C.c_void_p, C.c_double, C.POINTER(C.c_double)]),
        "tgh_transfer": (C.c_int, [P, C.c_void_p, C.c_double, P, C.c_void_p]),
        "tgh_problem_destroy": (C.c_int, [P]),
        "tgh_state_length": (C.c_int64, [P]),
        "tgh_work_length": (C.c_int64, [P]),
        "tgh_block_coords": (C.c_int, [P, C.POINTER(C.c_int32)]),
        "tgh_set_stream": (C.c_int, [P, P]),
        "tgh_prepare": (C.c_int, [P, P, D]),
        "tgh_rhs": (C.c_int, [P, P, D, C.POINTER(TghMatter)]),
        "tgh_gh_rhs": (C.c_int, [P, P, P, D, C.POINTER(TghMatter)]),
        "tgh_settle": (C.c_int, [P, P, D]),
        "tgh_stage_limit": (C.c_int, [P, P, D]),
        "tgh_step_limit": (C.c_int, [P, P, D]),
        "tgh_max_speed": (C.c_int, [P, P, D, C.POINTER(D)]),
        "tgh_initial_state": (C.c_int, [P, P, D]),
        "tgh_add_noise": (C.c_int, [P, P, C.c_uint64, D]),
        "tgh_rk4": (C.c_int, [P, P, D, D, C.c_int32, C.c_int32]),
        "tgh_error_norms": (C.c_int, [P, P, D, C.POINTER(D), C.POINTER(D)]),
        "tgh_error_rows": (C.c_int, [P, P, D, D, D, C.POINTER(D)]),
        "tgh_set_rho_max": (C.c_int, [P, D]),
        "tgh_set_matter": (C.c_int, [P, C.POINTER(TghMatter)]),
        "tgh_vjp_rhs": (C.c_int, [P, P, D, P, P, C.POINTER(D)]),
        "tgh_adjoint": (C.c_int, [P, P, D, D, C.c_int32, C.c_int32, C.POINTER(D), C.POINTER(D), C.POINTER(D)]),
        "tgh_copy_work": (C.c_int, [P, P]),
        "tgh_profile": (C.c_int, [P, C.c_int32]),
        "tgh_profile_read": (C.c_int, [P, C.POINTER(D)]),
        "tgh_last_error": (C.c_char_p, []),
        "tgh_apply_bounds": (C.c_int, [P, P, D, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(D)]),
        "tgh_batch_create": (C.c_int, [C.POINTER(P), C.c_int32, C.POINTER(P)]),
        "tgh_batch_rk4": (C.c_int, [P, C.POINTER(P), D, D, C.c_int32, C.c_int32]),
        "tgh_batch_destroy": (C.c_int, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc):
    if rc == TGH_OK:
        return
    msg = _load().tgh_last_error().decode()
    if rc in (TGH_EINVAL, TGH_ENOTFINITE):
        raise ArgumentError(msg)
    if rc == TGH_EUNSUPPORTED:
        raise NotImplementedError(msg)
    raise TGHError(msg)


# --- float types and backends ------------------------------------------------

def _check_float(T):
    import torch
    if T in (None, "Float64", float, torch.float64):
        return
    raise ArgumentError(f"this backend computes in Float64 only (the horizon interior and the "
                        f"constraint damping need it), got {T}")


class MI355X:
    """The backend object passed as ``backend=`` (the reference's
    KernelAbstractions backend keyword, src/driver.jl:467-479)."""

    def __init__(self, device=0):
        import torch
        if not torch.cuda.is_available():
            raise TGHError("MI355X backend: no GPU is visible to this process")
        self.device = torch.device("cuda", device)

    def __repr__(self):
        return f"MI355X({self.device.index})"


# --- backgrounds, profiles, interior, case -----------------------------------

class Minkowski:
    kind = 0


class GaugeWave:
    """The AwA gauge wave H = 1 - A sin(2 pi (x - t)/d) (src/initialdata.jl:897-901)."""
    kind = 1

    def __init__(self, A, d):
        self.A, self.d = float(A), float(d)


class KerrSchild:
    """KerrSchild(M, a): spin a along +z (SpacetimeMetrics, restated)."""
    kind = 2

    def __init__(self, M, a):
        self.M, self.a = float(M), float(a)


class Harmonic:
    """Harmonic(M, a): the Kerr hole in harmonic coordinates, spin a along +z
    (SpacetimeMetrics, restated from Cook & Scheel 1997). H_a = 0 exactly, so a
    case on it takes no gauge source (src/gauge.jl:64)."""
    kind = 4

    def __init__(self, M, a):
        self.M, self.a = float(M), float(a)
        if not (self.M > 0 and abs(self.a) < self.M):
            raise ArgumentError(f"the harmonic Kerr chart needs M > 0 and |a| < M, got M = {M}, "
                                f"a = {a}")


class ConstantDamping:
    def __init__(self, g0):
        self.g0 = float(g0)


class GaussianDamping:
    """far + (near - far) exp(-r^2 / 2 width^2) (src/gauge.jl:1122-1147)."""

    def __init__(self, *, near, far, width, center=(0.0, 0.0, 0.0)):
        if not float(width) > 0:
            raise ArgumentError(f"GaussianDamping: width = {width}; a Gaussian needs a positive width")
        if float(near) < 0 or float(far) < 0:
            raise ArgumentError(f"GaussianDamping: near = {near}, far = {far}; negative damping "
                                f"rates would drive the constraints instead of damping them")
        self.near, self.far, self.width = float(near), float(far), float(width)
        self.center = tuple(float(c) for c in center)


class HorizonDissipation:
    """HorizonDissipation(T; ε_out, ε_in, r_1, r_h, center) (src/gauge.jl:1192-1252): the
    Kreiss–Oliger amplitude ε_out at and outside r_h, ε_in at and inside r_1, a quintic
    smoothstep of r = |x − c| between."""

    def __init__(self, T=None, *, eps_out, eps_in, r_1, r_h, center=(0.0, 0.0, 0.0)):
        _check_float(T)
        eo, ei, r1, rh = float(eps_out), float(eps_in), float(r_1), float(r_h)
        if not (eo >= 0 and ei >= 0):
            raise ArgumentError(f"HorizonDissipation: ε_out = {eo}, ε_in = {ei}; a negative amplitude "
                                f"turns the dissipation into growth of the highest-frequency modes "
                                f"(the sign is already in the operator's weights)")
        if not (eo <= 4 and ei <= 4):
            raise ArgumentError(f"HorizonDissipation: ε_out = {eo}, ε_in = {ei}; amplitudes above 4 "
                                f"put the corner mode's decay rate 3ε/h outside RK4's real-axis "
                                f"stability interval (|z| ≤ 2.79) at the stepper's CFL numbers")
        if not 0 < r1 < rh:
            raise ArgumentError(f"HorizonDissipation: r_1 = {r1}, r_h = {rh}; the ramp runs inward "
                                f"from r_h to r_1, which needs 0 < r_1 < r_h")
        self.eps_out, self.eps_in, self.r_1, self.r_h = eo, ei, r1, rh
        self.center = tuple(float(c) for c in center)


def horizon_min_radius(background):
    """The horizon's smallest coordinate radius (src/interior.jl:760-764): r₊ = M + √(M² − a²)
    on the Kerr–Schild axis, R₊ = √(M² − a²) in harmonic coordinates."""
    if isinstance(background, KerrSchild):
        return background.M + (background.M ** 2 - background.a ** 2) ** 0.5
    if isinstance(background, Harmonic):
        return (background.M ** 2 - background.a ** 2) ** 0.5
    raise ArgumentError(f"{type(background).__name__} has no horizon")


def horizon_max_radius(background):
    """The horizon's largest coordinate radius (src/interior.jl:761-765): √(r₊² + a²)
    with r₊ the smallest (the equator's)."""
    return (horizon_min_radius(background) ** 2 + background.a ** 2) ** 0.5


def horizon_dissipation(case, *, eps_in, eps_out=None):
    """horizon_dissipation(case; ε_in, ε_out = case.ε_KO) (src/initialdata.jl:838-860)."""
    if case.interior is None:
        raise ArgumentError("horizon_dissipation reads r_1 from the case's interior, and this case "
                            "has none")
    eo = case.eps_ko if eps_out is None else eps_out
    if isinstance(eo, HorizonDissipation):
        raise ArgumentError("this case's ε_KO is already a HorizonDissipation; give eps_out "
                            "as a number")
    return HorizonDissipation(eps_out=eo, eps_in=eps_in, r_1=case.interior.r_1,
                              r_h=horizon_min_radius(case.background), center=case.interior.center)


class Interior:
    """Interior(T; center, r_0, r_1, ρ_max, variant, w_ramp, ρ_ramp, margin)
    (src/interior.jl:178-238). `margin` is m, the number of spacings the
    horizon keeps outside r_1 (r_1 ≤ r_h,min − m h), which the refinement's
    level floor is derived from."""
    VARIANTS = {"damped": 1, "frozen": 2, "pasted": 3, "fitted": 4}

    def __init__(self, T=None, *, center=(0.0, 0.0, 0.0), r_0, r_1, rho_max=0.0,
                 variant="damped", w_ramp=0.5, rho_ramp=0.5, margin=8):
        _check_float(T)
        if variant not in self.VARIANTS:
            raise ArgumentError(f"unknown interior variant :{variant} (known: "
                                f"{', '.join(self.VARIANTS)})")
        if not r_0 > 0:
            raise ArgumentError(f"Interior: r_0 = {r_0}; the core radius has to be positive")
        if not r_1 > r_0:
            raise ArgumentError(f"Interior: r_0 = {r_0}, r_1 = {r_1}; the layer is the shell r_0 ≤ r < r_1, so r_1 > r_0")
        if not (0 < w_ramp <= 1 and 0 < rho_ramp <= 1):
            raise ArgumentError(f"Interior: w_ramp = {w_ramp}, ρ_ramp = {rho_ramp}; the ramps are "
                                f"fractions of the layer, in (0, 1]")
        if rho_max < 0:
            raise ArgumentError(f"ρ_max = {rho_max} is negative")
        if int(margin) != margin or margin < 1:
            raise ArgumentError(f"the margin counts spacings between r_1 and the horizon and must be a "
                                f"positive integer, got {margin}")
        self.center = tuple(float(c) for c in center)
        self.r_0, self.r_1, self.rho_max = float(r_0), float(r_1), float(rho_max)
        self.variant, self.w_ramp, self.rho_ramp = variant, float(w_ramp), float(rho_ramp)
        self.margin = int(margin)

    def with_rho_max(self, rho_max):
        import copy
        if not rho_max >= 0:
            raise ArgumentError(f"ρ_max = {rho_max} is negative")
        c = copy.copy(self)
        c.rho_max = float(rho_max)
        return c


def shape_series(shape, lmax, n):
    """shape_series (src/interior.jl:1294-1329): Σ a_lm ỹ_lm(n̂) for real
    coefficients in real_harmonic_index order (l² + l + m), by the fully
    normalised Legendre recurrences (Condon–Shortley phase); numpy, host side."""
    import numpy as np
    n = np.asarray(n, dtype=float)
    x, y, z = n[..., 0], n[..., 1], n[..., 2]
    s2 = math.sqrt(2.0)
    acc = np.zeros_like(x)
    qmm = np.full_like(x, 1.0 / math.sqrt(4.0 * math.pi))
    C, S = np.ones_like(x), np.zeros_like(x)

    def term(l, m, q):
        if m == 0:
            return shape[l * l + l] * q
        return s2 * q * (shape[l * l + l + m] * C - shape[l * l + l - m] * S)

    for m in range(lmax + 1):
        if m > 0:
            qmm = -math.sqrt((2 * m + 1) / (2 * m)) * qmm
            C, S = C * x - S * y, C * y + S * x
        acc = acc + term(m, m, qmm)
        if m + 1 <= lmax:
            qa, qb = qmm, math.sqrt(2 * m + 3) * z * qmm
            acc = acc + term(m + 1, m, qb)
            for l in range(m + 2, lmax + 1):
                al = math.sqrt((4 * l * l - 1) / (l * l - m * m))
                be = math.sqrt(((l - 1) ** 2 - m * m) / (4 * (l - 1) ** 2 - 1))
                qa, qb = qb, al * (z * qb - be * qa)
                acc = acc + term(l, m, qb)
    return acc


def shape_bounds(shape, lmax):
    """shape_bounds (src/interior.jl:1331-1375): the extremes of the series over
    both poles and n_θ = 4 lmax + 3 colatitudes by 2 n_θ longitudes."""
    import numpy as np
    nth = 4 * int(lmax) + 3
    nph = 2 * nth
    th = np.pi * np.arange(1, nth + 1) / (nth + 1)
    ph = 2 * np.pi * np.arange(nph) / nph
    T, P = np.meshgrid(th, ph, indexing="ij")
    dirs = np.concatenate([np.array([[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]]),
                           np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)])
    v = shape_series(np.asarray(shape, dtype=float), lmax, dirs)
    return float(v.min()), float(v.max())


class FittedInterior:
    """FittedInterior(T; center, shape, lmax, offset, thickness, ρ_max, variant,
    w_ramp, ρ_ramp, r_in, r_out) (src/interior.jl:1752-1816): the tracked
    horizon r_h(n̂) as a real spherical-harmonic series about center (a point,
    or (c0, v) for a moving one), clamped into [r_in, r_out] (its sampled
    extremes by default); the layer is r_1 = r_h − offset, r_0 = r_1 − thickness
    along each ray."""
    VARIANTS = {"damped": 1, "frozen": 2, "fitted": 4}

    def __init__(self, T=None, *, center, shape, lmax=None, offset, thickness, rho_max=0.0,
                 variant="damped", w_ramp=0.5, rho_ramp=0.5, r_in=None, r_out=None, margin=8, h=0.0):
        _check_float(T)
        if variant not in self.VARIANTS:
            raise ArgumentError(f"unknown variant :{variant} for a tracked geometry (built here: "
                                f"{', '.join(self.VARIANTS)})")
        shape = [float(a) for a in shape]
        L = int(math.isqrt(len(shape)) - 1) if lmax is None else int(lmax)
        if (L + 1) ** 2 != len(shape):
            raise ArgumentError(f"FittedInterior: {len(shape)} coefficients do not make a series "
                                f"of degree {L}, which has {(L + 1) ** 2}")
        if L > 8:
            raise ArgumentError(f"FittedInterior: lmax = {L}; this library carries series up to degree 8")
        if r_in is None or r_out is None:
            r_in, r_out = shape_bounds(shape, L)
        if not 0 < r_in <= r_out:
            raise ArgumentError(f"FittedInterior: r_in = {r_in}, r_out = {r_out}; the clamp "
                                f"interval needs 0 < r_in ≤ r_out")
        if not (offset > 0 and thickness > 0):
            raise ArgumentError(f"FittedInterior: offset = {offset}, thickness = {thickness}; both "
                                f"are distances and have to be positive")
        if not (r_in - offset) - thickness > 0:
            raise ArgumentError(f"FittedInterior: r_in − offset − thickness = "
                                f"{(r_in - offset) - thickness} ≤ 0, so the core would swallow the center")
        if not (0 < w_ramp <= 1 and 0 < rho_ramp <= 1):
            raise ArgumentError(f"FittedInterior: w_ramp = {w_ramp}, ρ_ramp = {rho_ramp}; the "
                                f"ramps are fractions of the layer, in (0, 1]")
        if rho_max < 0:
            raise ArgumentError(f"ρ_max = {rho_max} is negative")
        self.center = tuple(float(c) for c in center)
        self.shape, self.lmax = shape, L
        self.r_in, self.r_out = float(r_in), float(r_out)
        self.offset, self.thickness = float(offset), float(thickness)
        self.rho_max, self.variant = float(rho_max), variant
        self.w_ramp, self.rho_ramp = float(w_ramp), float(rho_ramp)
        self.margin = int(margin)
        if not (math.isfinite(h) and h >= 0):
            raise ArgumentError(f"FittedInterior: the spacing h = {h} the geometry was built at must be ≥ 0")
        self.h = float(h)  # read by fill_target only (its fill radius and translation step)
        # the bounding radii of the layer, for the checks that read r_0 and r_1
        self.r_0 = (self.r_in - self.offset) - self.thickness
        self.r_1 = self.r_out - self.offset

    def with_rho_max(self, rho_max):
        import copy
        if not rho_max >= 0:
            raise ArgumentError(f"ρ_max = {rho_max} is negative")
        c = copy.copy(self)
        c.rho_max = float(rho_max)
        return c


class StateBounds:
    """StateBounds(T; α_min, α_max, λ_min, λ_max, β_max, K_max, r_gate)
    (src/bounds.jl:51-136): the range projection's ranges and its gate radius."""

    def __init__(self, T=None, *, alpha_min, alpha_max, lambda_min, lambda_max, beta_max, K_max, r_gate):
        _check_float(T)
        vals = [float(x) for x in (alpha_min, alpha_max, lambda_min, lambda_max, beta_max, K_max, r_gate)]
        if not all(math.isfinite(x) for x in vals):
            raise ArgumentError(f"StateBounds: {vals}; all seven numbers have to be finite (a case "
                                f"without bounds passes bounds=None)")
        a0, a1, l0, l1, b, K, rg = vals
        if not 0 < a0 < 1 < a1:
            raise ArgumentError(f"StateBounds: α_min = {a0}, α_max = {a1}; the lapse interval has "
                                f"to contain 1 and stay positive")
        if not 0 < l0 < 1 < l1:
            raise ArgumentError(f"StateBounds: λ_min = {l0}, λ_max = {l1}; the eigenvalue "
                                f"interval of γ_ij has to contain 1 and stay positive")
        if not b > 0:
            raise ArgumentError(f"StateBounds: β_max = {b} is not positive")
        if not K > 0:
            raise ArgumentError(f"StateBounds: K_max = {K} is not positive")
        if not rg > 0:
            raise ArgumentError(f"StateBounds: r_gate = {rg} is not positive")
        (self.alpha_min, self.alpha_max, self.lambda_min, self.lambda_max, self.beta_max, self.K_max,
         self.r_gate) = vals


def default_bounds(T=None, *, M, r_gate):
    """default_bounds(T; M, r_gate) (src/bounds.jl:138-160)."""
    return StateBounds(T, alpha_min=1 / 50, alpha_max=50.0, lambda_min=1 / 100, lambda_max=1000.0,
                       beta_max=10.0, K_max=100.0 / M, r_gate=r_gate)


def default_gate(interior, forest, q):
    """default_gate(interior, forest, q) = r_1 − 2 G h (src/bounds.jl:162-184), h the
    uniform forest's spacing."""
    box = forest.case.box if forest.case is not None else None
    if box is None:
        raise ArgumentError("default_gate needs the forest's case for its spacing")
    h = (box[0][1] - box[0][0]) / (forest.roots[0] * forest.N)
    return interior.r_1 - 2 * (q // 2 + 1) * h


class GHCase:
    """GHCase(T, background; box, periodic, ε_KO, γ0, γ2, interior, gauge_source)
    (src/initialdata.jl:158-311). ``gauge_source`` is None (harmonic
    backgrounds) or "algebraic" (the closed-form Kerr-Schild source)."""

    def __init__(self, T, background, *, box, periodic, eps_ko, gamma0, gamma2,
                 interior=None, gauge_source=None, bounds=None, velocity=(0.0, 0.0, 0.0), reflecting=None,
                 rotating=None):
        _check_float(T)
        refl = ((False, False),) * 3 if reflecting is None else \
            tuple((bool(r[0]), bool(r[1])) for r in reflecting)
        if len(refl) != 3:
            raise ArgumentError(f"reflecting is one (low, high) pair per axis, got {reflecting!r}")
        for dd in range(3):
            if periodic[dd] and any(refl[dd]):
                raise ArgumentError(f"axis {dd} cannot wrap and mirror at once: periodic[{dd}] is set and "
                                    f"reflecting[{dd}] = {refl[dd]}")
        # a symmetry plane (src/boundaries.jl:84-142): the ghosts mirror the
        # interior with each variable's parity; the hook fills only the faces
        # that neither wrap nor reflect
        self.reflecting = refl
        # the rotating seam (src/boundaries.jl:145-205): the low faces of d1 and
        # d2 glued by a quarter turn; (0, 0) for none
        rot = (0, 0) if rotating is None else tuple(int(x) for x in rotating)
        if rot != (0, 0):
            if len(rot) != 2 or rot[0] == rot[1] or not all(1 <= x <= 3 for x in rot):
                raise ArgumentError(f"rotating is a pair of two different axes (1-based), got {rotating!r}")
            if rot != (1, 2):
                raise NotImplementedError(f"the rotating seam is built for the axes (1, 2), a quarter turn "
                                          f"about z; got {rot}")
            for dd in (0, 1):
                if periodic[dd]:
                    raise ArgumentError(f"the seam's axis {dd + 1} is periodic: a seam ends the box there")
                if any(refl[dd]):
                    raise ArgumentError(f"the seam's axis {dd + 1} is reflecting: a face is glued or "
                                        f"mirrored, not both")
            w = [b[1] - b[0] for b in box]
            if abs(w[0] - w[1]) > 1e-12 * abs(w[0]) or box[0][0] != box[1][0]:
                raise ArgumentError(f"the seam turns x into y about the axis through their low faces, so "
                                    f"the two need one width and one low end; got {box[0]} and {box[1]}")
        self.rotating = rot
        self.velocity = tuple(float(x) for x in velocity)
        if any(self.velocity) and not isinstance(background, (KerrSchild, Harmonic)):
            raise ArgumentError("only a Kerr-Schild or harmonic hole moves")
        if bounds is not None and not isinstance(bounds, StateBounds):
            raise ArgumentError("bounds is None or a StateBounds")
        if bounds is not None and interior is None:
            raise ArgumentError("StateBounds act inside r_gate about the hole's center, and this "
                                "case has no hole interior to measure r from")
        self.bounds = bounds
        if not gamma2 > -1:
            raise ArgumentError(f"γ2 must satisfy γ2 > −1, got {gamma2}")
        if isinstance(eps_ko, HorizonDissipation):
            c0 = interior.center if interior is not None else (0.0, 0.0, 0.0)
            if tuple(eps_ko.center) != tuple(c0):
                raise ArgumentError("the dissipation profile is centred on the case's hole")
        elif eps_ko < 0:
            raise ArgumentError(f"ε_KO = {eps_ko} is negative")
        self.background = background
        self.box = tuple((float(a), float(b)) for a, b in box)
        self.periodic = tuple(bool(p) for p in periodic)
        self.eps_ko = eps_ko if isinstance(eps_ko, HorizonDissipation) else float(eps_ko)
        self.gamma0 = gamma0 if isinstance(gamma0, (ConstantDamping, GaussianDamping)) \
            else ConstantDamping(gamma0)
        self.gamma2 = float(gamma2)
        self.interior = interior
        if gauge_source not in (None, "algebraic"):
            raise ArgumentError(f"gauge_source is nothing (sampled where the background is not "
                                f"harmonic) or :algebraic, got {gauge_source!r}")
        if gauge_source == "algebraic" and not isinstance(background, KerrSchild):
            raise ArgumentError("the algebraic gauge source is the Kerr-Schild hole's own")
        self.gauge_source = gauge_source


def with_interior(case, interior):
    """with_interior (src/evolution.jl:1150-1190 spelling for a case): the same
    case with another interior (an Interior or a FittedInterior)."""
    import copy
    c = copy.copy(case)
    c.interior = interior
    return c


def minkowski_case(T=None, *, L, eps_ko, gamma0, gamma2):
    """minkowski_case (src/initialdata.jl:869-871): Minkowski space, periodic box [0, L]^3."""
    return GHCase(T, Minkowski(), box=[(0.0, L)] * 3, periodic=(True,) * 3, eps_ko=eps_ko,
                  gamma0=gamma0, gamma2=gamma2)


def minkowski_octant_case(T=None, *, L, eps_ko, gamma0, gamma2, rotating=False):
    """minkowski_octant_case (src/initialdata.jl:873-893): flat space on the
    octant [0, L]^3 of the box [-L, L]^3, a symmetry plane at each of the three
    faces through the origin and Minkowski's h = Pi = 0 at the three outer
    ones -- the robust-stability case on the mesh of a hole at the origin.
    `rotating` glues x = 0 to y = 0 by a quarter turn about z instead, and
    only z = 0 reflects."""
    return GHCase(T, Minkowski(), box=[(0.0, L)] * 3, periodic=(False,) * 3, eps_ko=eps_ko,
                  gamma0=gamma0, gamma2=gamma2,
                  reflecting=tuple((not rotating or d == 2, False) for d in range(3)),
                  rotating=(1, 2) if rotating else None)


def gauge_wave_case(T=None, *, A=1 / 20, d=1.0, eps_ko, gamma0, gamma2):
    """The AwA gauge wave on one periodic wavelength cubed (src/initialdata.jl:908-911)."""
    return GHCase(T, GaugeWave(A, d), box=[(0.0, d)] * 3, periodic=(True,) * 3, eps_ko=eps_ko,
                  gamma0=gamma0, gamma2=gamma2)


def _moving_hole(v, x0, axis):
    """moving_hole_case's placement (src/initialdata.jl:1283-1298): the hole at
    x0 + v t along axis (1, 2 or 3), the background translate(boost(bg, −v ê), x0 ê)."""
    if axis not in (1, 2, 3):
        raise ArgumentError(f"axis = {axis}: the hole moves along coordinate 1 (x), 2 (y) or 3 (z)")
    v, x0 = float(v), float(x0)
    if not abs(v) < 1:
        raise ArgumentError(f"|v| = {abs(v)} ≥ 1: a hole moves slower than light")
    on_axis = lambda s: tuple(s if d == axis - 1 else 0.0 for d in range(3))  # noqa: E731
    return on_axis(x0), on_axis(v)


def _octant_box(octant, center, vel, a, L):
    """The box, the reflecting faces and the rotating seam of hole_case's
    `octant` (src/initialdata.jl:993-1035, 1075-1170): None or False the whole
    box [-L, L]^3; True or "reflecting" the octant [0, L]^3 mirrored at the
    three planes through the origin; "rotating" the same octant with x = 0
    glued to y = 0 by a quarter turn about z and a mirror at z = 0;
    "quadrant" (-L, L) x (0, L)^2 mirrored at y = 0 and z = 0;
    "rotating_quadrant" (0, L)^2 x (-L, L) with the seam and no mirror;
    "bitant" (-L, L)^2 x (0, L) mirrored at z = 0. Each asks for a hole that
    is its own image under its maps at all times."""
    kind = "none" if octant in (None, False) else "reflecting" if octant is True else octant
    if kind not in ("none", "reflecting", "rotating", "quadrant", "rotating_quadrant", "bitant"):
        raise ArgumentError(f"octant is None, True, \"reflecting\", \"rotating\", \"quadrant\", "
                            f"\"rotating_quadrant\" or \"bitant\", got {octant!r}")
    c, v = center, vel
    if kind == "reflecting" and not (not any(c) and not any(v) and a == 0.0):
        raise ArgumentError(f"the octant mirrors the solution in the three coordinate planes, which holds "
                            f"only for a hole at the origin, at rest and without spin (a spinning one "
                            f"keeps only the z mirror); got center {c}, velocity {v}, a = {a}")
    if kind == "quadrant" and not (c[1] == c[2] == 0.0 and v[1] == v[2] == 0.0 and a == 0.0):
        raise ArgumentError(f"the quadrant mirrors the solution in y = 0 and z = 0 at every time, which holds "
                            f"only for a hole without spin on the x axis moving along it; got center {c}, "
                            f"velocity {v}, a = {a}")
    if kind == "bitant" and not (c[2] == 0.0 and v[2] == 0.0):
        raise ArgumentError(f"the bitant mirrors the solution in z = 0 at every time, which holds only for a "
                            f"hole in that plane moving in it, its spin along z; got center {c}, velocity {v}")
    if kind == "rotating" and not (not any(c) and not any(v)):
        raise ArgumentError(f"the rotating octant turns the solution by quarter turns about z and mirrors it "
                            f"in z = 0, which holds only for a hole at the origin and at rest, spinning about "
                            f"z; got center {c}, velocity {v}")
    if kind == "rotating_quadrant" and not (c[0] == c[1] == 0.0 and v[0] == v[1] == 0.0):
        raise ArgumentError(f"the rotating quadrant turns the solution by quarter turns about z at every time, "
                            f"which holds only for a hole on the z axis moving along it; got center {c}, "
                            f"velocity {v}")
    half = {"none": (), "reflecting": (0, 1, 2), "rotating": (0, 1, 2), "quadrant": (1, 2),
            "rotating_quadrant": (0, 1), "bitant": (2,)}[kind]
    mirror = {"reflecting": (0, 1, 2), "rotating": (2,), "quadrant": (1, 2), "bitant": (2,)}.get(kind, ())
    box = [((0.0 if d in half else -L), L) for d in range(3)]
    return box, tuple((d in mirror, False) for d in range(3)), (1, 2) if kind.startswith("rotating") else None


def kerr_schild_case(T=None, *, M=1.0, a=0.0, v=0.0, x0=0.0, axis=1, halfwidth, r_0, r_1, rho_max=None,
                     eps_ko=0.5, gamma0=None, gamma2=0.0, interior="damped", w_ramp=0.5, rho_ramp=0.5,
                     gauge_source=None, bounds=None, margin=8, octant=None):
    """kerr_schild_case / hole_case (src/initialdata.jl:1044-1174,1230-1298) for a
    hole in a Dirichlet box, at x0 + v t along axis. ρ_max defaults to the
    physical rate 4/M the driver sets per chunk. gauge_source=None samples
    H_a = −Γ_a of the background once (the reference's default, a hole at rest
    only); "algebraic" is the closed-form KerrSchildSource evaluated in the
    kernel, which moves with the hole. `octant` puts the hole on a symmetric
    part of the box with reflecting faces (_octant_box)."""
    center, vel = _moving_hole(v, x0, axis)
    if any(vel) and gauge_source is None:
        raise ArgumentError("a moving Kerr-Schild hole has a time-dependent H_a, which a source sampled once at "
                            "creation cannot follow: pass gauge_source=\"algebraic\" (evaluated in the kernel at "
                            "each point and time) or move a harmonic_kerr_case hole instead")
    damping = gamma0 if gamma0 is not None else GaussianDamping(near=1 / M, far=1 / (10 * M),
                                                                width=3 * M, center=center)
    inter = Interior(T, center=center, r_0=r_0, r_1=r_1, rho_max=(4.0 / M if rho_max is None else rho_max),
                     variant=interior, w_ramp=w_ramp, rho_ramp=rho_ramp, margin=margin)
    L = float(halfwidth)
    box, refl, rot = _octant_box(octant, center, vel, float(a), L)
    return GHCase(T, KerrSchild(M, a), box=box, periodic=(False,) * 3, eps_ko=eps_ko,
                  gamma0=damping, gamma2=gamma2, interior=inter, gauge_source=gauge_source,
                  bounds=bounds, velocity=vel, reflecting=refl, rotating=rot)


def harmonic_kerr_case(T=None, *, M=1.0, a=0.0, v=0.0, x0=0.0, axis=1, halfwidth, r_0, r_1, rho_max=None,
                       eps_ko=0.5, gamma0=None, gamma2=0.0, interior="damped", w_ramp=0.5, rho_ramp=0.5,
                       bounds=None, margin=8, octant=None):
    """harmonic_kerr_case / hole_case (src/initialdata.jl:1300-1326, 1044-1174):
    Harmonic(M, a) in a Dirichlet box, at x0 + v t along axis, with no gauge
    source (a boost of the harmonic chart is harmonic). Defaults as
    kerr_schild_case's."""
    center, vel = _moving_hole(v, x0, axis)
    damping = gamma0 if gamma0 is not None else GaussianDamping(near=1 / M, far=1 / (10 * M),
                                                                width=3 * M, center=center)
    inter = Interior(T, center=center, r_0=r_0, r_1=r_1, rho_max=(4.0 / M if rho_max is None else rho_max),
                     variant=interior, w_ramp=w_ramp, rho_ramp=rho_ramp, margin=margin)
    L = float(halfwidth)
    box, refl, rot = _octant_box(octant, center, vel, float(a), L)
    return GHCase(T, Harmonic(M, a), box=box, periodic=(False,) * 3, eps_ko=eps_ko,
                  gamma0=damping, gamma2=gamma2, interior=inter, bounds=bounds, velocity=vel,
                  reflecting=refl, rotating=rot)


def _morton(i, j, k):
    key = 0
    for bit in range(21):
        key |= ((i >> bit) & 1) << (3 * bit) | ((j >> bit) & 1) << (3 * bit + 1) | ((k >> bit) & 1) << (3 * bit + 2)
    return key


class Forest:
    """gh_forest's forest (src/initialdata.jl:1370-1395): a brick of root blocks
    of N cells (roots), uniform, or a frozen hierarchy whose leaves are
    (level, i, j, k), block (i, j, k) of the brick of level l (roots_d 2^l
    blocks of spacing h / 2^l). `leaves` is None on a uniform forest.

    `ownership` names which block owns a point on a block face: "closed"
    (TreeAMR 0.3, the reference's mesh: a block owns both face planes, a
    shared point's primary is the finest leaf holding it, both outer planes
    of a Dirichlet box are evolved) or "half_open" (TreeAMR 0.1: a block owns
    [0, N) per axis and the upper outer plane is a Dirichlet ghost), the mesh
    the reference's published frozen-hierarchy tables were measured on."""

    OWNERSHIP = {"closed": 0, "half_open": 1}

    def __init__(self, case, N, roots, leaves=None, ownership="closed"):
        roots = (roots,) * 3 if isinstance(roots, int) else tuple(roots)
        if len(roots) != 3 or any(r < 1 for r in roots):
            raise ArgumentError(f"roots = {roots!r}: give one positive count, or three (one per "
                                f"axis)")
        if ownership not in self.OWNERSHIP:
            raise ArgumentError(f"ownership is \"closed\" or \"half_open\", got {ownership!r}")
        self.ownership = ownership
        self.N, self.roots, self.case = int(N), roots, case
        self._leaves = None if leaves is None else set(tuple(int(x) for x in lf) for lf in leaves)
        if case is not None:
            widths = [b[1] - b[0] for b in case.box]
            cell = [widths[d] / roots[d] for d in range(3)]
            if any(abs(c - cell[0]) > 1e-12 * abs(cell[0]) for c in cell):
                raise ArgumentError(f"the box widths {widths} over roots {roots} give root blocks of "
                                    f"widths {cell}; blocks have one spacing on all three axes, so "
                                    f"the widths must be in the roots' proportions")

    @property
    def refined(self):
        return self._leaves is not None

    @property
    def leaves(self):
        """The leaves in Morton order of their lower corners at the finest level."""
        if self._leaves is None:
            return [(0, i, j, k) for k in range(self.roots[2]) for j in range(self.roots[1])
                    for i in range(self.roots[0])]
        L = self.maxlevel()
        return sorted(self._leaves, key=lambda lf: (_morton(*(c << (L - lf[0]) for c in lf[1:])), lf[0]))

    def nleaves(self):
        return (self.roots[0] * self.roots[1] * self.roots[2]) if self._leaves is None else len(self._leaves)

    def maxlevel(self):
        return 0 if self._leaves is None else max(lf[0] for lf in self._leaves)

    def root_spacing(self):
        lo, hi = self.case.box[0]
        return (hi - lo) / (self.roots[0] * self.N)

    def minimum_spacing(self):
        return self.root_spacing() / 2 ** self.maxlevel()

    def block_extent(self, leaf):
        """((lo, hi) per axis) of leaf (level, i, j, k)."""
        l = leaf[0]
        h = self.root_spacing() / 2 ** l
        return tuple((self.case.box[d][0] + leaf[1 + d] * self.N * h,
                      self.case.box[d][0] + (leaf[1 + d] + 1) * self.N * h) for d in range(3))

    def _containing(self, leafset, l, c):
        """The leaf holding level-l block c: its ancestor of the lowest depth
        that is a leaf, or None where the region is refined below l."""
        for lv in range(l, -1, -1):
            key = (lv,) + tuple(x >> (l - lv) for x in c)
            if key in leafset:
                return key
        return None

    def _neighbours(self, lf):
        l = lf[0]
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    if dx == dy == dz == 0:
                        continue
                    c = [lf[1] + dx, lf[2] + dy, lf[3] + dz]
                    ok = True
                    for d in range(3):
                        n = self.roots[d] << l
                        if not 0 <= c[d] < n:
                            if not self.case.periodic[d]:
                                ok = False
                            c[d] %= n
                    if ok:
                        yield tuple(c)

    def refine(self, targets):
        """refine!(forest, targets): each target leaf replaced by its 8 children."""
        if self._leaves is None:
            self._leaves = set(self.leaves)
        for t in list(targets):
            t = tuple(t)
            if t not in self._leaves:
                raise ArgumentError(f"{t} is not a leaf of the forest")
            self._leaves.remove(t)
            l, i, j, k = t
            for c in range(8):
                self._leaves.add((l + 1, 2 * i + (c & 1), 2 * j + (c >> 1 & 1), 2 * k + (c >> 2 & 1)))

    def coarsen(self, parents):
        """coarsen!(forest, parents): each parent (level, i, j, k) whose eight
        children are all leaves becomes a leaf again."""
        if self._leaves is None:
            raise ArgumentError("a uniform forest has nothing to coarsen")
        for t in list(parents):
            l, i, j, k = (int(x) for x in t)
            kids = [(l + 1, 2 * i + (c & 1), 2 * j + (c >> 1 & 1), 2 * k + (c >> 2 & 1)) for c in range(8)]
            if l < 0 or any(c not in self._leaves for c in kids):
                raise ArgumentError(f"{tuple(t)} is not the parent of eight leaves")
            self._leaves.difference_update(kids)
            self._leaves.add((l, i, j, k))

    def _unbalanced(self):
        leafset = self._leaves or set()
        bad = set()
        seen = set()  # (level, block): siblings share most of their neighbours
        for lf in leafset:
            if lf[0] < 2:
                continue
            for c in self._neighbours(lf):
                key = (lf[0],) + c
                if key in seen:
                    continue
                seen.add(key)
                o = self._containing(leafset, lf[0], c)
                if o is not None and o[0] < lf[0] - 1:
                    bad.add(o)
        return bad

    def balance(self):
        """balance!(forest): refine until leaves that touch (faces, edges,
        corners, across periodic seams) differ by at most one level."""
        while True:
            bad = self._unbalanced()
            if not bad:
                return self
            self.refine(bad)

    def isbalanced(self):
        return not self._unbalanced()

    def leaf_array(self):
        import numpy as np
        return np.array(self.leaves, dtype=np.int32).reshape(-1, 4)


def gh_forest(T=None, case=None, *, N, roots, refined=False, ownership="closed"):
    """gh_forest(T, case; N, roots, refined = false) (src/initialdata.jl:1370-1395):
    the uniform brick, or with `refined` the two-level mesh whose root block
    holding lo + 3/8 of the box is refined once (every face of it a coarse-fine
    face)."""
    _check_float(T)
    f = Forest(case, N, roots, ownership=ownership)
    if not refined:
        return f
    xref = [case.box[d][0] + 3 * (case.box[d][1] - case.box[d][0]) / 8 for d in range(3)]
    targets = [lf for lf in f.leaves
               if all(f.block_extent(lf)[d][0] <= xref[d] < f.block_extent(lf)[d][1] for d in range(3))]
    if len(targets) != 1:
        raise ArgumentError(f"gh_forest(refined=True): {len(targets)} root blocks contain the "
                            f"point {xref} that selects the refined block; expected exactly one")
    f.refine(targets)
    return f.balance()


def _segment_box_dist2(lo, hi, c, dc):
    """The squared distance from the segment c + s dc, s in [0, 1], to each box
    [lo, hi] (arrays n x 3): the per-axis excess max(lo - x, 0, x - hi) is
    piecewise linear in s, so its square summed over axes is a convex
    piecewise quadratic; its minimum is at a breakpoint (where x_d(s) crosses
    lo_d or hi_d), at an end, or at the vertex of the quadratic on a piece."""
    import numpy as np
    n = lo.shape[0]
    cand = [np.zeros(n), np.ones(n)]
    for d in range(3):
        if dc[d] != 0.0:
            for b in (lo[:, d], hi[:, d]):
                cand.append(np.clip((b - c[d]) / dc[d], 0.0, 1.0))
    S = np.sort(np.stack(cand, 1), 1)  # breakpoints per box
    def f(s):
        x = c[None, None, :] + s[..., None] * dc[None, None, :]
        e = np.maximum(np.maximum(lo[:, None, :] - x, 0.0), x - hi[:, None, :])
        return (e * e).sum(-1)
    best = f(S).min(1)
    # inside each piece [S_k, S_k+1] the active sides are fixed: the quadratic's
    # vertex from the excess's slope and offset at the piece's midpoint
    mid = 0.5 * (S[:, 1:] + S[:, :-1])
    x = c[None, None, :] + mid[..., None] * dc[None, None, :]
    below = x < lo[:, None, :]
    above = x > hi[:, None, :]
    # excess = a + b s on the active axes: a = lo - c (below) or c - hi (above), b = -dc or dc
    a = np.where(below, lo[:, None, :] - c, np.where(above, c - hi[:, None, :], 0.0))
    b = np.where(below, -dc, np.where(above, dc, 0.0))
    bb = (b * b).sum(-1)
    ab = (a * b).sum(-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        sv = np.where(bb > 0, -ab / bb, mid)
    sv = np.clip(sv, S[:, :-1], S[:, 1:])
    best = np.minimum(best, f(sv).min(1))
    return best

def _boxes_meet_shell(lo, hi, c, R, shape, dc):
    """Which boxes [lo, hi] (n x 3 arrays) meet the shell of radius R about c:
    the closed ball, the open cube, or either swept along [c, c + dc]."""
    import numpy as np
    c = np.asarray(c, float)
    if dc is None:
        if shape == "ball":
            e = np.maximum(np.maximum(lo - c, 0.0), c - hi)
            return (e * e).sum(1) <= R * R
        return np.all((lo < c + R) & (hi > c - R), axis=1)
    dc = np.asarray(dc, float)
    if shape == "ball":
        return _segment_box_dist2(lo, hi, c, dc) <= R * R
    # the union of open cubes about c + s dc: per axis the s where the box's
    # [lo, hi] overlaps (c + s dc - R, c + s dc + R), an open interval in s
    s_lo = np.zeros(lo.shape[0])
    s_hi = np.ones(lo.shape[0])
    ok = np.ones(lo.shape[0], bool)
    for d in range(3):
        a, b = lo[:, d] - R - c[d], hi[:, d] + R - c[d]  # need a < s dc < b
        if dc[d] == 0.0:
            ok &= (a < 0.0) & (0.0 < b)
        else:
            t0, t1 = np.minimum(a / dc[d], b / dc[d]), np.maximum(a / dc[d], b / dc[d])
            s_lo = np.maximum(s_lo, t0)
            s_hi = np.minimum(s_hi, t1)
    return ok & (s_lo < s_hi)


def hole_forest(T=None, case=None, *, N, roots, center=None, radii, levels=None, shape="ball", path=None,
                ownership="closed"):
    """hole_forest(T, case; N, roots, center, radii, levels, shape, path)
    (src/initialdata.jl:1417-1510): a fixed nested hierarchy about `center`
    (the case's hole by default). Starting from the uniform forest, level l
    refines every level-(l − 1) leaf whose box meets shell l of radius
    radii[l] -- the closed ball |x − c| ≤ R (shape "ball") or the open cube
    |x − c|_∞ < R ("cube"), each swept along the segment [c, c + path] when a
    path is given -- and 2:1 balance follows each level. The layout depends on
    roots and radii only, so raising N halves every spacing and moves no
    block."""
    import numpy as np
    _check_float(T)
    if shape not in ("ball", "cube"):
        raise ArgumentError(f"hole_forest: shape = :{shape}; the shells are \"ball\" or \"cube\"")
    if path is not None and len(path) != 3:
        raise ArgumentError(f"hole_forest: path = {path!r}; give None or a displacement of "
                            f"three numbers")
    radii = [float(r) for r in radii]
    levels = len(radii) if levels is None else int(levels)
    if len(radii) != levels:
        raise ArgumentError(f"hole_forest: {len(radii)} radii for {levels} levels; one "
                            f"radius refines one level")
    if min(radii) <= 0:
        raise ArgumentError(f"hole_forest: radii = {radii} include a non-positive one")
    if any(b > a for a, b in zip(radii, radii[1:])):
        raise ArgumentError(f"hole_forest: radii = {radii} grow somewhere; a finer level's "
                            f"shell lies inside the coarser one's")
    if center is None:
        center = case.interior.center if case.interior is not None else (0.0, 0.0, 0.0)
    c = np.asarray(center, dtype=float)
    dc = None if path is None else np.asarray(path, dtype=float)
    f = Forest(case, N, roots, ownership=ownership)
    for l, R in enumerate(radii, start=1):
        cand = [lf for lf in f.leaves if lf[0] == l - 1]
        ext = np.array([f.block_extent(lf) for lf in cand])  # leaves x axis x (lo, hi)
        hit = _boxes_meet_shell(ext[:, :, 0], ext[:, :, 1], c, R, shape, dc) if cand else []
        targets = [lf for lf, h in zip(cand, hit) if h]
        if not targets:
            raise ArgumentError(f"hole_forest: shell {l} (radius {R} about {tuple(c)}) meets no "
                                f"level-{l - 1} leaf, so it would add no level")
        f.refine(targets)
        f.balance()
    return f


class Operators:
    """TreeAMR's Operators(prolongation = p, restriction = r): the ghost
    schedule's interpolation orders. Restriction on a vertex-centred mesh is
    injection (a coincident fine point copied), whatever r says."""

    def __init__(self, *, prolongation, restriction=None):
        self.prolongation = int(prolongation)
        self.restriction = None if restriction is None else int(restriction)


def _desc(forest, case, q):
    d = TghDesc()
    d.float_bits = 64
    d.q, d.N = int(q), forest.N
    d.nblocks[:] = list(forest.roots)
    d.periodic[:] = [1 if p else 0 for p in case.periodic]
    for dd in range(3):
        r = getattr(case, "reflecting", ((False, False),) * 3)[dd]
        d.reflecting[dd][0], d.reflecting[dd][1] = int(r[0]), int(r[1])
    d.rotating = 1 if getattr(case, "rotating", (0, 0)) == (1, 2) else 0
    bg = case.background
    d.background = bg.kind
    d.lo[:] = [b[0] for b in case.box]
    d.hi[:] = [b[1] for b in case.box]
    if isinstance(bg, GaugeWave):
        d.wave_amp, d.wave_len = bg.A, bg.d
    if isinstance(bg, (KerrSchild, Harmonic)):
        d.M, d.a = bg.M, bg.a
    harmonic = isinstance(bg, (Minkowski, GaugeWave, Harmonic))
    d.gauge = 2 if case.gauge_source == "algebraic" else (0 if harmonic else 1)
    g0 = case.gamma0
    if isinstance(g0, GaussianDamping):
        d.damping = 1
        d.gamma0_near, d.gamma0_far, d.gamma0_width = g0.near, g0.far, g0.width
    else:
        d.damping, d.gamma0 = 0, g0.g0
    if isinstance(case.eps_ko, HorizonDissipation):
        hd = case.eps_ko
        d.dissipation = 1
        d.eps_ko, d.eps_in, d.eps_r1, d.eps_rh = hd.eps_out, hd.eps_in, hd.r_1, hd.r_h
    else:
        d.dissipation, d.eps_ko = 0, case.eps_ko
    d.gamma2 = case.gamma2
    bd = getattr(case, "bounds", None)
    if bd is not None:
        (d.alpha_min, d.alpha_max, d.lambda_min, d.lambda_max, d.beta_max, d.K_max,
         d.r_gate) = (bd.alpha_min, bd.alpha_max, bd.lambda_min, bd.lambda_max, bd.beta_max, bd.K_max,
                      bd.r_gate)
    d.velocity[:] = list(getattr(case, "velocity", (0.0, 0.0, 0.0)))
    d.ownership = Forest.OWNERSHIP[getattr(forest, "ownership", "closed")]
    it = case.interior
    if it is not None:
        d.interior = Interior.VARIANTS[it.variant]
        d.center[:] = list(it.center)
        d.r0, d.r1, d.rho_max, d.w_ramp, d.rho_ramp = it.r_0, it.r_1, it.rho_max, it.w_ramp, it.rho_ramp
        if isinstance(it, FittedInterior):
            d.shaped, d.shape_lmax = 1, it.lmax
            d.shape[:len(it.shape)] = it.shape
            d.r_in, d.r_out, d.offset, d.thickness = it.r_in, it.r_out, it.offset, it.thickness
    return d


class GHProblem:
    """GHProblem(forest, case; q, backend=MI355X()) (src/evolution.jl:1034-1149):
    the device work arrays, the block table and the case, built once."""

    def __init__(self, forest, case, *, q, backend=None, ops=None, target=None, t_target=0.0,
                 target_rate=False, trail=0.0):
        if not (q >= 2 and q % 2 == 0):
            raise ArgumentError(f"GHProblem: q = {q}; the centred stencils are of even order "
                                f"2, 4, 6 or 8")
        self.backend = backend if backend is not None else MI355X()
        self.forest, self.case, self.q, self.ops = forest, case, int(q), ops
        self.desc = _desc(forest, case, q)
        self._ptr = C.c_void_p()
        import torch
        planned = forest.refined or forest.ownership != "closed"
        if planned and ops is None:
            raise ArgumentError(f"a forest with levels (or half-open blocks) takes its ghost operators "
                                f"from the caller, ops = Operators(prolongation = p): the interface rule "
                                f"wants p = q + 2 = {q + 2}, and this library does not pick it")
        with torch.cuda.device(self.backend.device):
            if planned:
                leaves = forest.leaf_array()
                arr = leaves.ravel()
                _check(_load().tgh_problem_create_forest(
                    C.byref(self.desc), arr.ctypes.data_as(C.POINTER(C.c_int32)), len(leaves),
                    ops.prolongation, C.byref(self._ptr)))
            else:
                _check(_load().tgh_problem_create(C.byref(self.desc), C.byref(self._ptr)))
            # order the library's work after torch's on the same stream
            _check(_load().tgh_set_stream(self._ptr, C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        self.state_length = int(_load().tgh_state_length(self._ptr))
        self.work_length = int(_load().tgh_work_length(self._ptr))
        self.nblocks = forest.nleaves()
        self.target = None
        if target is not None or (case.interior is not None and case.interior.variant == "fitted"):
            if target is None:
                raise ArgumentError("the :fitted variant reads its target from a cache the caller supplies: "
                                    "pass GHProblem(..., target=...) with a device array of 40 variables "
                                    "per owned point (A then S)")
            self.set_target(target, t_target, rate=target_rate, trail=trail)

    def set_target(self, target, t_target, *, rate=False, trail=0.0):
        """The :fitted target cache (src/evolution.jl:1007-1021): a float64 CUDA
        tensor of 40 (N+1)^3 nblocks values (A_v then S_v per block), read as
        A + (t − t_target) S; the problem keeps a reference to it."""
        import torch
        if (not isinstance(target, torch.Tensor) or target.dtype != torch.float64 or not target.is_cuda
                or not target.is_contiguous() or target.numel() != 2 * self.state_length):
            raise ArgumentError(f"the target cache must be a contiguous float64 CUDA tensor of "
                                f"{2 * self.state_length} values (40 per owned point)")
        self.target = target
        _check(_load().tgh_set_target(self._ptr, C.c_void_p(target.data_ptr()), float(t_target),
                                      1 if rate else 0, float(trail)))

    def __del__(self):
        if getattr(self, "_ptr", None) and self._ptr.value and _lib is not None:
            _lib.tgh_problem_destroy(self._ptr)
            self._ptr = C.c_void_p()

    @property
    def ptr(self):
        return self._ptr

    def statevector(self):
        import torch
        return torch.zeros(self.state_length, dtype=torch.float64, device=self.backend.device)

    def block_coords(self):
        out = (C.c_int32 * (3 * self.nblocks))()
        _check(_load().tgh_block_coords(self._ptr, out))
        return [tuple(out[3 * b:3 * b + 3]) for b in range(self.nblocks)]

    def leaves(self):
        """The leaves (level, i, j, k) in the state's block order."""
        if self.forest.refined:
            return self.forest.leaves
        return [(0,) + bc for bc in self.block_coords()]

    def work(self):
        import torch
        w = torch.empty(self.work_length, dtype=torch.float64, device=self.backend.device)
        _check(_load().tgh_copy_work(self._ptr, C.c_void_p(w.data_ptr())))
        return w


def _dev(x, p, name):
    import torch
    if not isinstance(x, torch.Tensor) or x.dtype != torch.float64 or not x.is_cuda:
        raise ArgumentError(f"{name} must be a float64 CUDA tensor, got {type(x).__name__}")
    if not x.is_contiguous() or x.numel() < p.state_length:
        raise ArgumentError(f"{name} must be contiguous with at least {p.state_length} entries")
    return C.c_void_p(x.data_ptr())


def _matter(matter):
    if matter is None:
        return None
    rho, S, Sij = matter
    m = TghMatter(C.c_void_p(rho.data_ptr()), C.c_void_p(S.data_ptr()), C.c_void_p(Sij.data_ptr()))
    return C.byref(m)


def prepare(p, u, t):
    """prepare!(p, u, t) -> p (src/evolution.jl:1349-1357)."""
    _check(_load().tgh_prepare(p.ptr, _dev(u, p, "u"), float(t)))
    return p


def rhs(du, p, t, matter=None):
    """rhs!(du, p, t; matter) (src/evolution.jl:1386-1390)."""
    _check(_load().tgh_rhs(p.ptr, _dev(du, p, "du"), float(t), _matter(matter)))


def gh_rhs(du, u, p, t, matter=None):
    """gh_rhs!(du, u, p, t; matter) (src/evolution.jl:1327-1331)."""
    _check(_load().tgh_gh_rhs(p.ptr, _dev(du, p, "du"), _dev(u, p, "u"), float(t), _matter(matter)))


def settle(p, u, t):
    """settle!(p, u, t) -> u (src/evolution.jl:1370-1374)."""
    _check(_load().tgh_settle(p.ptr, _dev(u, p, "u"), float(t)))
    return u


def stage_limit(u, p, t):
    """stage_limit!(u, p, t) (src/stepping.jl:91-95)."""
    _check(_load().tgh_stage_limit(p.ptr, _dev(u, p, "u"), float(t)))


def step_limit(u, p, t):
    """step_limit!(u, p, t) (src/stepping.jl:97)."""
    _check(_load().tgh_step_limit(p.ptr, _dev(u, p, "u"), float(t)))


def apply_bounds(p, u, t):
    """apply_bounds!(p, u, t) -> points moved (src/bounds.jl:733-770): the range
    projection inside the gate, written back only where it fired."""
    n, nf, rm = C.c_int64(), C.c_int64(), C.c_double()
    _check(_load().tgh_apply_bounds(p.ptr, _dev(u, p, "u"), float(t), C.byref(n), C.byref(nf), C.byref(rm)))
    p.bounds_last = (n.value, nf.value, rm.value)
    return n.value


def max_speed(p, u, t):
    """max_speed(p, u, t) -> λ, checked (src/evolution.jl:1539-1560)."""
    out = C.c_double()
    _check(_load().tgh_max_speed(p.ptr, _dev(u, p, "u"), float(t), C.byref(out)))
    return out.value


def initial_state(u, p, t):
    """initial_state!(u, p, t) -> u (src/stepping.jl:110-115)."""
    _check(_load().tgh_initial_state(p.ptr, _dev(u, p, "u"), float(t)))
    return u


def add_noise(u, p, base, amplitude):
    """add_noise! with the SplitMix64 point draw (src/initialdata.jl:1961-2013)."""
    _check(_load().tgh_add_noise(p.ptr, _dev(u, p, "u"), C.c_uint64(base), float(amplitude)))
    return u


def evolved_sets(p):
    """evolved_sets(p) (src/stepping.jl:125): the one evolved field set."""
    return (p,)


def error_norms(p, u, t, *, shell=None):
    """gh_error!(p, u, t; shell) + error_norms(p) (src/constraints.jl:311-396,
    629-671). Without `shell`: (err_l2, err_linf), the masked volume-weighted
    L2 and the L-inf of e = |u − u_exact(t)| (the 2-norm over the 20
    variables) over evolved primaries. With `shell = (lo, hi)`: the record's
    four rows as a dict, err_l2, err_linf, residual (the layer's L-inf of e)
    and drift (the L-inf of |h_tt − h_tt,exact| over evolved primaries with
    lo ≤ r ≤ hi)."""
    out = (C.c_double * 4)()
    lo, hi = (0.0, -1.0) if shell is None else (float(shell[0]), float(shell[1]))
    _check(_load().tgh_error_rows(p.ptr, _dev(u, p, "u"), float(t), lo, hi, out))
    if shell is None:
        return out[0], out[1]
    return {"err_l2": out[0], "err_linf": out[1], "residual": out[2], "drift": out[3]}


TGH_PI_TILES = 6


def autotune(p, u, dt, t=0.0):
    """The brick stepper's autotuner (no reference counterpart): times each
    tile shape of the Π kernel on a scratch copy of u and keeps the fastest
    for p's later steps; returns the average ms per Π launch of each shape."""
    w = u.clone()
    ms = (C.c_double * TGH_PI_TILES)()
    _check(_load().tgh_autotune(p.ptr, _dev(w, p, "u"), float(t), float(dt), ms))
    return list(ms)


def gauge_constraint_norms(p, u, t):
    """gh_constraint!(p, u, t) + constraint_norms(p)'s gauge rows
    (src/constraints.jl:211-309,559-566,848-876): the masked (L2, L∞) of the
    four components of C_a = g_ab (Γ^b + g^{bc} H_c)."""
    l2, li = (C.c_double * 4)(), (C.c_double * 4)()
    _check(_load().tgh_gauge_constraint(p.ptr, _dev(u, p, "u"), float(t), l2, li))
    return list(l2), list(li)


def gauge_constraint_norms_noalloc(p, u, t):
    """gauge_constraint_norms through the entry point that allocates nothing
    (no reference counterpart): the partials live in the library's static
    device array, so it serves callers that cannot allocate; the same norms
    to the bit, for problems of at most 116 506 workgroup partials (5825
    blocks of 16^3 cells)."""
    l2, li = (C.c_double * 4)(), (C.c_double * 4)()
    _check(_load().tgh_gauge_constraint_noalloc(p.ptr, _dev(u, p, "u"), float(t), l2, li))
    return list(l2), list(li)


def adm_constraint_norms(p, u, t):
    """adm_constraint!(p, u, t) + constraint_norms(p)'s ham and mom rows
    (src/constraints.jl:432-534,584-605,848-876): (L2, L∞) of (ℋ, ℳ_x, ℳ_y, ℳ_z)."""
    l2, li = (C.c_double * 4)(), (C.c_double * 4)()
    _check(_load().tgh_adm_constraint(p.ptr, _dev(u, p, "u"), float(t), l2, li))
    return list(l2), list(li)


class FitParams:
    """FitParams{T}(lmax, cont, rbar, center, bounds[, tilde]) (src/fit.jl:492-502):
    what an evaluation of a fit needs besides its coefficients -- the degree,
    the radial order, the radius the ansatz is scaled by, the center (a point,
    or (c0, v) for c(t) = c0 + v t), the StateBounds its states are projected
    into (the gate is not used) and whether the momentum slots hold
    (α/√γ) Π."""

    def __init__(self, T=None, *, lmax, cont, rbar, center, bounds, tilde=False):
        _check_float(T)
        if int(lmax) != lmax or not 0 <= lmax <= 32:
            raise ArgumentError(f"FitParams: lmax = {lmax}; this library evaluates degrees 0 to 32")
        if cont not in (0, 1, 2):
            raise ArgumentError(f"FitParams: cont = {cont}; the radial order is 0, 1 or 2")
        if not (math.isfinite(rbar) and rbar > 0):
            raise ArgumentError(f"FitParams: r̄ = {rbar} has to be a positive length")
        if not isinstance(bounds, StateBounds):
            raise ArgumentError("FitParams: bounds must be a StateBounds (fit_state always projects)")
        c = list(center)
        if len(c) == 2 and all(hasattr(x, "__len__") for x in c):
            c0, v = (tuple(float(x) for x in c[0]), tuple(float(x) for x in c[1]))
        else:
            c0, v = tuple(float(x) for x in c), (0.0, 0.0, 0.0)
        if len(c0) != 3 or len(v) != 3:
            raise ArgumentError("FitParams: the center is a point or (c0, v), three coordinates each")
        self.lmax, self.cont, self.rbar, self.tilde = int(lmax), int(cont), float(rbar), bool(tilde)
        self.c0, self.v, self.bounds = c0, v, bounds

    @property
    def nslots(self):
        return (self.lmax + 1) ** 2


class InteriorFit:
    """The part of InteriorFit (src/fit.jl:504-539) the cache's fill reads:
    the params, the coefficients C[slot, k, v] as a float64 CUDA tensor of
    shape (20, cont + 1, (L+1)^2) (the reference's ((L+1)^2, cont+1, 20)
    column-major array, so the same memory) and the time t it was built at.
    Building a fit (the least squares of build_fit) is the caller's."""

    def __init__(self, params, coeffs, t):
        import torch
        if not isinstance(params, FitParams):
            raise ArgumentError("InteriorFit: params must be a FitParams")
        shape = (20, params.cont + 1, params.nslots)
        if (not isinstance(coeffs, torch.Tensor) or coeffs.dtype != torch.float64 or not coeffs.is_cuda
                or tuple(coeffs.shape) != shape or not coeffs.is_contiguous()):
            raise ArgumentError(f"InteriorFit: the coefficients must be a contiguous float64 CUDA tensor of "
                                f"shape {shape}")
        if not math.isfinite(t):
            raise ArgumentError(f"InteriorFit: t = {t} is not finite")
        self.params, self.coeffs, self.t = params, coeffs, float(t)

    def _struct(self):
        p, b = self.params, self.params.bounds
        f = TghFit()
        f.lmax, f.cont, f.tilde, f.rbar = p.lmax, p.cont, 1 if p.tilde else 0, p.rbar
        f.center[:], f.velocity[:] = list(p.c0), list(p.v)
        f.alpha_min, f.alpha_max, f.lambda_min, f.lambda_max = b.alpha_min, b.alpha_max, b.lambda_min, b.lambda_max
        f.beta_max, f.K_max, f.t = b.beta_max, b.K_max, self.t
        f.coeffs = self.coeffs.data_ptr()
        return f


def target_cache(p):
    """target_cache(U) (src/fit.jl:1118-1127): a zeroed cache of 40 variables per
    owned point (A_h, A_Π, S_h, S_Π) on p's forest, as a float64 CUDA tensor."""
    import torch
    return torch.zeros(2 * p.state_length, dtype=torch.float64, device=p.backend.device)


def fill_target(target, p, fits, t, *, rate=False, clear=True):
    """fill_target!(target, origins, spacings, interior, fits, t; rate, clear)
    (src/fit.jl:1129-1151, fit_target_kernel! :1057-1116) on p's mesh and its
    tracked geometry (a FittedInterior): fits = (latest, previous), previous
    None before there are two. At every owned point with |x − c(t)| < r_fill =
    (r_out − offset) + h the cache takes A = F_a + (t − t_a) S and S = (F_a −
    F_b)/(t_a − t_b) (0 without an older previous fit), F the fits' states at
    t; with rate, S also carries the latest fit's translation at the point.
    With clear the other points are zeroed. Returns target."""
    it = p.case.interior
    if not isinstance(it, FittedInterior):
        raise ArgumentError("fill_target fills the tracked geometry's cache: the problem's interior must be "
                            "a FittedInterior")
    fa, fb = fits
    if not isinstance(fa, InteriorFit) or (fb is not None and not isinstance(fb, InteriorFit)):
        raise ArgumentError("fill_target: fits = (latest, previous) of InteriorFit, previous None before "
                            "there are two")
    _dev(target, p, "the target cache")
    if target.numel() != 2 * p.state_length:
        raise ArgumentError(f"the target cache has {target.numel()} values, not {2 * p.state_length} (40 per "
                            f"owned point)")
    sa = fa._struct()
    sb = fb._struct() if fb is not None else None
    _check(_load().tgh_fill_target(p.ptr, C.c_void_p(target.data_ptr()), C.byref(sa),
                                   C.byref(sb) if sb is not None else None, float(t), float(it.h),
                                   1 if rate else 0, 1 if clear else 0))
    return target


def record_norms(p, u, t, *, adm=True, shell=None):
    """record_norms(p) (src/constraints.jl:868-881): constraint_norms' rows and
    error_norms' rows together, what the driver's analysis row reads. The
    reference folds the three monitors' slots of its diagnostics field set in
    one sweep after running them; here each monitor folds its own partials,
    so the call runs them on (u, t): the gauge constraint always, the ADM
    constraints when `adm` (the driver runs them on their own cadence; without
    them the ham and mom fields are None), and the error rows with the drift
    over `shell` (lo, hi) when given. Returns a dict: gauge_l2, gauge_linf
    (4 each), ham_l2, ham_linf, mom_l2, mom_linf (3 each), err_l2, err_linf,
    residual, drift."""
    g2, gi = gauge_constraint_norms(p, u, t)
    rec = {"gauge_l2": g2, "gauge_linf": gi, "ham_l2": None, "ham_linf": None, "mom_l2": None,
           "mom_linf": None}
    if adm:
        a2, ai = adm_constraint_norms(p, u, t)
        rec.update(ham_l2=a2[0], ham_linf=ai[0], mom_l2=a2[1:], mom_linf=ai[1:])
    rec.update(error_norms(p, u, t, shell=shell if shell is not None else (0.0, -1.0)))
    return rec


def field_scale(p, u, t):
    """field_scale (src/refinement.jl:221-300): the largest |h_v| over the evolved points."""
    out = C.c_double()
    _check(_load().tgh_field_scale(p.ptr, _dev(u, p, "u"), float(t), C.byref(out)))
    return out.value


def refinement_indicator(p, u, t, *, scale, eps=0.01):
    """gh_tau! + each block's maximum (src/refinement.jl:152-219,312-378): the
    Löhner indicator's largest value per block, in the state's block order."""
    out = (C.c_double * p.nblocks)()
    _check(_load().tgh_refinement_indicator(p.ptr, _dev(u, p, "u"), float(t), float(scale), float(eps), out))
    return list(out)


def transfer(p_from, u_from, t, p_to, u_to):
    """The regrid's state transfer (src/refinement.jl:973-1021, refine!/coarsen!
    and the state carried over): u_to on p_to's leaves takes at every owned
    point what p_from's ghost fill at t gives there -- a copy of the owner
    (injection from a finer leaf) or the prolongation from the leaf one
    level coarser."""
    _check(_load().tgh_transfer(p_from.ptr, _dev(u_from, p_from, "u_from"), float(t), p_to.ptr,
                                _dev(u_to, p_to, "u_to")))
    return u_to


class Refinement:
    """Refinement(T; refine_tol, coarsen_tol, maxlevel_cap, floor_margin,
    ceiling_cells = 1, ceiling_level = 0, ε = 1/100) (src/refinement.jl:
    93-134): the regrid's thresholds on the Löhner indicator τ ∈ [0, 1], the
    level cap, how far past the horizon's largest radius the level floor
    reaches, and the ceiling: blocks within ceiling_cells root spacings of a
    non-periodic face stay at ceiling_level."""

    def __init__(self, T=None, *, refine_tol, coarsen_tol, maxlevel_cap, floor_margin, ceiling_cells=1,
                 ceiling_level=0, eps=0.01):
        _check_float(T)
        rtol, ctol = float(refine_tol), float(coarsen_tol)
        if not 0 < ctol < rtol:
            raise ArgumentError(f"the two thresholds need 0 < coarsen_tol < refine_tol (the band between "
                                f"them keeps a block at the threshold from refining and coarsening in "
                                f"turn), got coarsen_tol = {ctol}, refine_tol = {rtol}")
        if not rtol < 1:
            raise ArgumentError(f"τ never exceeds 1, so refine_tol = {rtol} would never refine anything; "
                                f"pick a threshold below 1")
        if int(maxlevel_cap) != maxlevel_cap or not 0 <= maxlevel_cap <= 10:
            raise ArgumentError(f"maxlevel_cap is a level in 0 .. 10, got {maxlevel_cap}")
        if not float(floor_margin) >= 0:
            raise ArgumentError(f"the floor's shell ends floor_margin beyond the horizon's largest radius, "
                                f"which needs floor_margin ≥ 0, got {floor_margin}")
        if int(ceiling_cells) != ceiling_cells or ceiling_cells < 0:
            raise ArgumentError(f"ceiling_cells counts root spacings from the outer faces and must be a "
                                f"non-negative integer, got {ceiling_cells}")
        if int(ceiling_level) != ceiling_level or not 0 <= ceiling_level <= maxlevel_cap:
            raise ArgumentError(f"ceiling_level must lie in 0 .. maxlevel_cap = {maxlevel_cap}, got "
                                f"{ceiling_level}")
        if not float(eps) > 0:
            raise ArgumentError(f"the indicator's noise floor ε must be positive, got {eps}")
        self.refine_tol, self.coarsen_tol = rtol, ctol
        self.maxlevel_cap, self.floor_margin = int(maxlevel_cap), float(floor_margin)
        self.ceiling_cells, self.ceiling_level, self.eps = int(ceiling_cells), int(ceiling_level), float(eps)


def _box_radii(ext, c):
    """The nearest and farthest distance from c to the box ext."""
    near = sum(max(ext[d][0] - c[d], 0.0, c[d] - ext[d][1]) ** 2 for d in range(3)) ** 0.5
    far = sum(max(abs(ext[d][0] - c[d]), abs(ext[d][1] - c[d])) ** 2 for d in range(3)) ** 0.5
    return near, far


def level_bounds(forest, case, q, refinement, t, leaves=None):
    """block_level_bounds over `leaves` (forest.leaves by default;
    src/refinement.jl:430-668): per
    leaf (floor, cap). The floor is the coarsest level whose spacing meets
    both interior requirements, h ≤ (r_h,min − r_1)/m and h ≤ (r_1 −
    r_0)/(2(G+1)) (for a tracked geometry h ≤ offset/m and thickness/(2(G+1))),
    on the leaves meeting the shell from r_1 (the tracked core surface) to the
    horizon's largest radius plus floor_margin about the hole's center at t;
    the cap is ceiling_level within ceiling_cells root spacings of a
    non-periodic face and maxlevel_cap elsewhere. A floor the cap forbids is
    an error."""
    ref, f = refinement, forest
    it = case.interior
    G = q // 2 + 1
    h0 = f.root_spacing()
    lo, hi, L = 0.0, -1.0, 0
    c = (0.0, 0.0, 0.0)
    if it is not None:
        vel = getattr(case, "velocity", (0.0, 0.0, 0.0))
        c = tuple(it.center[d] + vel[d] * t for d in range(3))
        if isinstance(it, FittedInterior):
            need = min(it.offset / it.margin, it.thickness / (2 * (G + 1))) * (1 + 8 * 2.0 ** -52)
            lo, hi = it.r_0, it.r_out + ref.floor_margin
        else:
            rmin, rmax = horizon_min_radius(case.background), horizon_max_radius(case.background)
            if not rmin > it.r_1:
                raise ArgumentError(f"r_1 = {it.r_1} is not inside the horizon's smallest radius {rmin}: "
                                    f"no level can keep m = {it.margin} spacings between them")
            need = min((rmin - it.r_1) / it.margin, (it.r_1 - it.r_0) / (2 * (G + 1)))
            lo, hi = it.r_1, rmax + ref.floor_margin
        h = h0
        while h > need and L < 10:
            L, h = L + 1, h / 2
        if L > ref.maxlevel_cap:
            raise ArgumentError(f"the interior needs level {L} around the hole (h ≤ {need:.6g} at root "
                                f"spacing {h0:.6g}), above maxlevel_cap = {ref.maxlevel_cap}: raise the "
                                f"cap, refine the root brick or move the layer outward")
    margin = ref.ceiling_cells * h0
    out = []
    for lf in (f.leaves if leaves is None else leaves):
        ext = f.block_extent(lf)
        near, far = _box_radii(ext, c)
        flo = L if (L > 0 and hi >= lo and near <= hi and far >= lo) else 0
        capped = any(not case.periodic[d] and (ext[d][0] - case.box[d][0] <= margin or
                                               case.box[d][1] - ext[d][1] <= margin) for d in range(3))
        cap = ref.ceiling_level if capped else ref.maxlevel_cap
        if flo > cap:
            raise ArgumentError(f"leaf {lf} meets the horizon shell [{lo:.6g}, {hi:.6g}], which wants level "
                                f"{flo}, and lies within {margin:.6g} of the outer boundary, which caps it at "
                                f"{cap}: widen the box, narrow floor_margin or ceiling_cells")
        out.append((flo, cap))
    return out


def refine_flags(leaves, tau, bounds, refinement):
    """refine_flags (src/refinement.jl:670-730) on per-leaf indicator maxima:
    each leaf asks for one level finer where τ passes refine_tol below its
    cap, its own level where τ passes coarsen_tol, one coarser elsewhere; the
    request is clamped into the leaf's (floor, cap) and returned as +1
    (refine), 0 (keep) or -1 (coarsen)."""
    ref = refinement
    if not len(leaves) == len(tau) == len(bounds):
        raise ArgumentError(f"one τ and one (floor, cap) per leaf: {len(leaves)} leaves, {len(tau)} τ, "
                            f"{len(bounds)} bounds")
    flags = []
    for lf, tb, (flo, cap) in zip(leaves, tau, bounds):
        l = lf[0]
        if tb > ref.refine_tol and l < cap:
            want = l + 1
        elif tb > ref.coarsen_tol:
            want = l
        else:
            want = max(l - 1, 0)
        want = min(max(want, flo), cap)
        flags.append((want > l) - (want < l))
    return flags


def regrid_flags(p, u, t, *, refinement, scale=None):
    """The regrid's marking on the device state: the Löhner indicator's block
    maxima (scaled by the field scale of u unless `scale` is given) through
    refine_flags with the problem's level_bounds at t. Returns (flags, τ per
    leaf of p.leaves())."""
    leaves = p.leaves()
    bounds = level_bounds(p.forest, p.case, p.q, refinement, t, leaves)
    if scale is None:
        scale = field_scale(p, u, t)
    tau = refinement_indicator(p, u, t, scale=scale, eps=refinement.eps)
    return refine_flags(leaves, tau, bounds, refinement), tau


def regrid_forest(forest, leaves, flags):
    """The new mesh of regrid! (src/refinement.jl:973-1021 with TreeAMR's
    refine!/coarsen!/balance!) from per-leaf flags (+1 refine, 0 keep, -1
    coarsen) over `leaves`: the marked leaves refine once, balance! closes
    2:1, and each sibling group whose eight leaves are all marked to coarsen
    merges where no neighbour region of the parent's level is refined below
    the children's level (so the parent stays balanced)."""
    if len(flags) != len(leaves):
        raise ArgumentError(f"one flag per leaf: {len(flags)} flags for {len(leaves)} leaves")
    f = Forest(forest.case, forest.N, forest.roots, leaves, ownership=forest.ownership)
    f.refine([lf for lf, fl in zip(leaves, flags) if fl > 0])
    f.balance()
    want = {lf for lf, fl in zip(leaves, flags) if fl < 0}
    parents = {(lf[0] - 1, lf[1] >> 1, lf[2] >> 1, lf[3] >> 1) for lf in want if lf[0] > 0}
    for par in sorted(parents):
        l = par[0] + 1
        kids = [(l, 2 * par[1] + (c & 1), 2 * par[2] + (c >> 1 & 1), 2 * par[3] + (c >> 2 & 1)) for c in range(8)]
        if not all(c in want and c in f._leaves for c in kids):
            continue
        if any(f._containing(f._leaves, l, (2 * n[0] + (c & 1), 2 * n[1] + (c >> 1 & 1), 2 * n[2] + (c >> 2 & 1)))
               is None for n in f._neighbours(par) for c in range(8)):
            continue
        f.coarsen([par])
    return f


def regrid(p, u, t, *, refinement, scale=None):
    """regrid!(forest, U; flags) with the flags of regrid_flags: flagged leaves
    refine once, sibling groups all flagged to coarsen merge where the parent
    stays 2:1 balanced, balance! closes the rest, the state moves to the new
    leaves (transfer), and the stage limiter projects it as the driver does
    after a regrid (src/driver.jl:1692-1700: prolongation into a new leaf can
    leave a point outside the state bounds). Returns (the new problem, its
    state, the flags); a regrid that leaves the mesh as it was returns p and
    u themselves."""
    if p.target is not None:
        raise ArgumentError("a :fitted problem relaxes toward a target cache that the fit computes on its "
                            "own leaves; after a regrid the fit must fill the new mesh's cache, and the "
                            "fit is not built here: regrid a problem without a target")
    flags, _ = regrid_flags(p, u, t, refinement=refinement, scale=scale)
    leaves = p.leaves()
    f = regrid_forest(p.forest, leaves, flags)
    if set(f.leaves) == set(leaves):  # the mesh as it was: keep the problem
        return p, u, flags
    new = GHProblem(f, p.case, q=p.q, backend=p.backend,
                    ops=p.ops if p.ops is not None else Operators(prolongation=p.q + 2))
    u_new = new.statevector()
    transfer(p, u, t, new, u_new)
    stage_limit(u_new, new, t)
    return new, u_new, flags


def gh_dt(p, u, *, cfl, t=0.0):
    """gh_dt(p, u; cfl) = cfl · h_min / λ_max (src/evolution.jl:1511-1521)."""
    lam = max_speed(p, u, t)
    return float(cfl) * p.forest.minimum_spacing() / lam


def gh_solve(p, u, tspan, *, dt, graph=False):
    """gh_solve(p, u, (t0, t1); dt) (src/stepping.jl:235-239): fixed-step RK4 with
    nsteps = ⌈(t1 − t0)/dt⌉ and Δt = (t1 − t0)/nsteps, the limiters on every
    stage, in place. Returns u (settled)."""
    t0, t1 = float(tspan[0]), float(tspan[1])
    nsteps = max(1, math.ceil((t1 - t0) / dt * (1 - 4 * 2.0 ** -52)))
    step = (t1 - t0) / nsteps
    _check(_load().tgh_rk4(p.ptr, _dev(u, p, "u"), t0, step, nsteps, TGH_RK4_GRAPH if graph else 0))
    return u


def set_matter(p, matter):
    """gh_integrator(...; matter) (src/stepping.jl:150-154): matter fields
    (rho, S, Sij as float64 CUDA tensors in the tgh_matter layout, or None for
    vacuum) that every right-hand side of gh_solve adds, held fixed over the
    steps. The problem keeps a reference to them."""
    _check(_load().tgh_set_matter(p.ptr, _matter(matter)))
    p.matter = matter
    return p


def set_rho_max(p, rho_max):
    """with_interior(p, with_ρ_max(p.interior, ρ)): the layer's relaxation rate
    of the problem, nothing else changed. Returns p."""
    _check(_load().tgh_set_rho_max(p.ptr, float(rho_max)))
    if p.case.interior is not None:
        p.case = with_interior(p.case, p.case.interior.with_rho_max(float(rho_max)))
    return p


ADJOINT_PARAMS = ("gamma0_near", "gamma0_far", "gamma2", "rho_max", "rho_ramp")


def vjp_rhs(p, u, t, du_bar):
    """The vector-Jacobian product of gh_rhs! (src/evolution.jl:1327) at (u, t):
    (u_bar, theta_bar) = ((dF/du)^T du_bar, (dF/dθ)^T du_bar) with θ =
    ADJOINT_PARAMS (the damping's Gaussian rates, γ2, the layer's ρ_max and ρ
    ramp). u_bar is zero at the face copies the fill overwrites."""
    import torch
    u_bar = torch.empty_like(u)
    th = (C.c_double * 5)()
    _check(_load().tgh_vjp_rhs(p.ptr, _dev(u, p, "u"), float(t), _dev(du_bar, p, "du_bar"), _dev(u_bar, p, "u_bar"), th))
    return u_bar, list(th)


def adjoint(p, u, t0, dt, nsteps, *, every):
    """J and dJ/dθ for J = the masked L2 error after nsteps RK4 steps of dt from
    u (in place: u ends at T = t0 + nsteps dt), θ = ADJOINT_PARAMS: the forward
    sweep keeps a checkpoint every `every` steps, the reverse sweep recomputes
    each segment's stages from its checkpoint and applies the transposed
    right-hand side and RK4 combine. Returns (J, grad, (forward_ms, reverse_ms))."""
    J = C.c_double()
    g = (C.c_double * 5)()
    ms = (C.c_double * 2)()
    _check(_load().tgh_adjoint(p.ptr, _dev(u, p, "u"), float(t0), float(dt), int(nsteps), int(every), C.byref(J), g, ms))
    return J.value, list(g), (ms[0], ms[1])


def chunk_count(t_end, chunk):
    """chunk_count (src/precision.jl:81-87): ⌈t_end / chunk⌉, except that a
    quotient within a few ulp of an integer m gives m."""
    r = t_end / chunk
    m = round(r)
    tol = 4 * (math.ulp(r) + math.ulp(t_end) / chunk)
    return int(m) if m >= 1 and abs(r - m) <= tol else math.ceil(r)


def chunk_bounds(c, nchunks, t_end, chunk):
    """chunk_bounds (src/precision.jl:96-100): chunk c (1-based) spans
    (c − 1)·chunk to c·chunk, both capped at t_end, the last ending at t_end."""
    return min((c - 1) * chunk, t_end), (t_end if c == nchunks else min(c * chunk, t_end))


def evolve(p, u, *, t_end, chunk, cfl=0.25, rho_max_factor=None, rho_max_fixed=None, graph=True):
    """evolve!'s chunk loop for an analytic case on a frozen mesh
    (src/driver.jl:1732-1960, the record row :1540-1560): for each chunk
    (chunk_count / chunk_bounds) the largest speed λ over the evolved points
    at the chunk's start, dt = cfl · h_min / λ, steps = ⌈span / dt⌉ and the
    step span / steps; the layer's rate for the chunk (chunk_interior,
    :2391-2411): ρ_max = rho_max_factor / dt, else rho_max_fixed, else the
    case's own (hole_case's default 4/M), a fixed rate refused above 1/dt;
    then the steps, and at the chunk's end the record: t, dt, steps, λ, the
    error rows (gh_error! with the horizon shell) and the gauge constraint's
    largest component norms. The state starts as u holds it at t = 0 (the
    caller's initial_state). Holes at rest only: a moving hole's step takes a
    growth factor this loop does not model. Returns the records."""
    if rho_max_factor is not None and rho_max_fixed is not None:
        raise ArgumentError(f"evolve was given both rho_max_factor = {rho_max_factor} and rho_max_fixed = "
                            f"{rho_max_fixed}: the layer relaxes at the grid rate factor/dt or at a fixed "
                            f"rate, one or the other")
    case = p.case
    if any(getattr(case, "velocity", (0.0, 0.0, 0.0))):
        raise ArgumentError("evolve's chunk loop is built for holes at rest; a moving hole's time step "
                            "carries a growth factor of its own")
    it = case.interior
    shell = (0.0, -1.0)
    if it is not None and isinstance(case.background, (KerrSchild, Harmonic)) and not isinstance(it, FittedInterior):
        shell = (horizon_min_radius(case.background), horizon_max_radius(case.background) + (it.r_1 - it.r_0))
    hmin = p.forest.minimum_spacing()
    nchunks = chunk_count(float(t_end), float(chunk))
    records = []
    for c in range(1, nchunks + 1):
        tstart, stop = chunk_bounds(c, nchunks, float(t_end), float(chunk))
        lam = max_speed(p, u, tstart)
        dt = cfl * hmin / lam
        steps = max(1, math.ceil((stop - tstart) / dt))
        dt_used = (stop - tstart) / steps
        if it is not None:
            if rho_max_factor is not None:
                set_rho_max(p, rho_max_factor / dt_used)
            else:
                fixed = rho_max_fixed if rho_max_fixed is not None else it.rho_max
                if not fixed * dt_used <= 1:
                    raise ArgumentError(f"ρ_max = {fixed} exceeds this chunk's "
                                        f"grid rate 1/dt = {1 / dt_used}: relax at most once per step "
                                        f"(refine the mesh, lower the rate, or give rho_max_factor)")
                if fixed != p.case.interior.rho_max:
                    set_rho_max(p, fixed)
        _check(_load().tgh_rk4(p.ptr, _dev(u, p, "u"), tstart, dt_used, steps, TGH_RK4_GRAPH if graph else 0))
        rows = error_norms(p, u, stop, shell=shell)
        gl2, gli = gauge_constraint_norms(p, u, stop)
        import torch
        records.append({"t": stop, "dt": dt_used, "steps": steps, "lambda": lam, **rows,
                        "gauge_l2": max(gl2), "gauge_linf": max(gli),
                        "rho_max": p.case.interior.rho_max if it is not None else None,
                        "finite": bool(torch.isfinite(u).all())})
    return records


def gh_solve_regrid(p, u, tspan, *, cfl, every, refinement, graph=False):
    """The driver's regrid cadence (src/driver.jl:1640-1700, 1786-1800): chunks
    of `every` RK4 steps, each chunk's step sized on its own mesh as the
    driver sizes it, dt = cfl · h_min / λ (λ the largest speed over the
    evolved points at the chunk's start), and every chunk but the last
    followed by regrid at its end time. The last chunk takes ⌈rest / dt⌉
    steps of rest / steps to end at tspan[1]. Returns (the final problem, its
    state, [(t, nblocks, dt)] per regrid)."""
    if every < 1:
        raise ArgumentError(f"the regrid cadence is a positive number of steps, got {every}")
    t, t1 = float(tspan[0]), float(tspan[1])
    log = []
    while True:
        dt = gh_dt(p, u, cfl=cfl, t=t)
        rest = t1 - t
        n = max(1, math.ceil(rest / dt * (1 - 4 * 2.0 ** -52)))
        last = n <= every
        step = rest / n if last else dt
        k = n if last else every
        _check(_load().tgh_rk4(p.ptr, _dev(u, p, "u"), t, step, k, TGH_RK4_GRAPH if graph else 0))
        if last:
            return p, u, log
        t = t + k * step
        p, u, _ = regrid(p, u, t, refinement=refinement)
        log.append((t, p.nblocks, step))


class GHBatch:
    """Several independent problems (holes of uneven block counts) stepped
    together: each RK4 stage is one head and one Pi launch over all of them,
    driven by a device-side table of the problems' ranges and arguments. The
    problems share q, the gauge-source kind and whether they have an interior;
    each computes the bits gh_solve computes for it alone. No reference
    counterpart: the reference steps each problem by its own gh_solve
    (src/stepping.jl:208-239)."""

    def __init__(self, problems):
        self.problems = list(problems)
        n = len(self.problems)
        arr = (C.c_void_p * n)(*[p.ptr.value for p in self.problems])
        self._ptr = C.c_void_p()
        _check(_load().tgh_batch_create(arr, n, C.byref(self._ptr)))

    def __del__(self):
        if getattr(self, "_ptr", None) and self._ptr.value and _lib is not None:
            _lib.tgh_batch_destroy(self._ptr)
            self._ptr = C.c_void_p()

    @property
    def ptr(self):
        return self._ptr


def gh_solve_batch(batch, us, tspan, *, dt, graph=False):
    """gh_solve for every problem of the batch over the same span (nsteps and
    Δt as gh_solve's); us[i] is problem i's state, stepped in place."""
    if len(us) != len(batch.problems):
        raise ArgumentError(f"the batch holds {len(batch.problems)} problems, got {len(us)} states")
    t0, t1 = float(tspan[0]), float(tspan[1])
    nsteps = max(1, math.ceil((t1 - t0) / dt * (1 - 4 * 2.0 ** -52)))
    step = (t1 - t0) / nsteps
    ptrs = (C.c_void_p * len(us))(*[_dev(u, p, f"us[{i}]").value
                                    for i, (u, p) in enumerate(zip(us, batch.problems))])
    import torch
    torch.cuda.current_stream().synchronize()
    _check(_load().tgh_batch_rk4(batch.ptr, ptrs, t0, step, nsteps, TGH_RK4_GRAPH if graph else 0))
    return us
