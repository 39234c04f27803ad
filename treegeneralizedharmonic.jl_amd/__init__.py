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
    "field_scale", "refinement_indicator",
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
    ]


class TghMatter(C.Structure):
    _fields_ = [("rho", C.c_void_p), ("S", C.c_void_p), ("Sij", C.c_void_p)]


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
        "tgh_autotune": (C.c_int, [P, C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_double)]),
        "tgh_set_tile": (C.c_int, [P, C.c_int32]),
        "tgh_gauge_constraint": (C.c_int, [P, C.c_void_p, C.c_double, C.POINTER(C.c_double),
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
            raise ArgumentError(f"the damping profile's width must be positive, got {width}")
        if float(near) < 0 or float(far) < 0:
            raise ArgumentError(f"both constraint-damping rates must satisfy γ0 ≥ 0, got near = "
                                f"{near} and far = {far}")
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
            raise ArgumentError(f"both Kreiss–Oliger amplitudes must be non-negative, got ε_out = "
                                f"{eo} and ε_in = {ei}: the operator carries the damping sign in "
                                f"its weights, so a negative amplitude amplifies the grid-scale "
                                f"modes it is there to remove.")
        if not (eo <= 4 and ei <= 4):
            raise ArgumentError(f"the Kreiss–Oliger amplitude is capped at 4, got ε_out = {eo} and "
                                f"ε_in = {ei}: the operator damps the lattice's corner mode at "
                                f"3ε/h, and RK4 is stable on the negative real axis only to 2.79.")
        if not 0 < r1 < rh:
            raise ArgumentError(f"the profile rises from the horizon inward, so it needs "
                                f"0 < r_1 < r_h, got r_1 = {r1} and r_h = {rh}")
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


def horizon_dissipation(case, *, eps_in, eps_out=None):
    """horizon_dissipation(case; ε_in, ε_out = case.ε_KO) (src/initialdata.jl:838-860)."""
    if case.interior is None:
        raise ArgumentError("the dissipation profile rises from the horizon to the layer's outer "
                            "radius r_1, and this case has no interior, so no r_1.")
    eo = case.eps_ko if eps_out is None else eps_out
    if isinstance(eo, HorizonDissipation):
        raise ArgumentError("the profile's exterior amplitude is a number: pass eps_out "
                            "explicitly for a case whose ε_KO is already a profile.")
    return HorizonDissipation(eps_out=eo, eps_in=eps_in, r_1=case.interior.r_1,
                              r_h=horizon_min_radius(case.background), center=case.interior.center)


class Interior:
    """Interior(T; center, r_0, r_1, ρ_max, variant, w_ramp, ρ_ramp)
    (src/interior.jl:178-238)."""
    VARIANTS = {"damped": 1, "frozen": 2, "pasted": 3, "fitted": 4}

    def __init__(self, T=None, *, center=(0.0, 0.0, 0.0), r_0, r_1, rho_max=0.0,
                 variant="damped", w_ramp=0.5, rho_ramp=0.5):
        _check_float(T)
        if variant not in self.VARIANTS:
            raise ArgumentError(f"the interior variant must be one of {tuple(self.VARIANTS)}, "
                                f"got :{variant}")
        if not r_0 > 0:
            raise ArgumentError(f"the frozen core's radius must be positive, got r_0 = {r_0}")
        if not r_1 > r_0:
            raise ArgumentError(f"the layer needs r_1 > r_0, got r_0 = {r_0} and r_1 = {r_1}")
        if not (0 < w_ramp <= 1 and 0 < rho_ramp <= 1):
            raise ArgumentError(f"the ramp fractions must lie in (0, 1], but w_ramp = {w_ramp} "
                                f"and ρ_ramp = {rho_ramp}")
        if rho_max < 0:
            raise ArgumentError(f"the relaxation rate must satisfy ρ_max ≥ 0, got {rho_max}")
        self.center = tuple(float(c) for c in center)
        self.r_0, self.r_1, self.rho_max = float(r_0), float(r_1), float(rho_max)
        self.variant, self.w_ramp, self.rho_ramp = variant, float(w_ramp), float(rho_ramp)

    def with_rho_max(self, rho_max):
        return Interior(center=self.center, r_0=self.r_0, r_1=self.r_1, rho_max=rho_max,
                        variant=self.variant, w_ramp=self.w_ramp, rho_ramp=self.rho_ramp)


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
                 variant="damped", w_ramp=0.5, rho_ramp=0.5, r_in=None, r_out=None):
        _check_float(T)
        if variant not in self.VARIANTS:
            raise ArgumentError(f"the interior variant of a tracked geometry must be one of "
                                f"{tuple(self.VARIANTS)} here, got :{variant}")
        shape = [float(a) for a in shape]
        L = int(math.isqrt(len(shape)) - 1) if lmax is None else int(lmax)
        if (L + 1) ** 2 != len(shape):
            raise ArgumentError(f"a shape of degree lmax = {L} has (lmax + 1)² = {(L + 1) ** 2} real "
                                f"coefficients, got {len(shape)}")
        if L > 8:
            raise ArgumentError(f"the shape's degree is at most 8 here, got lmax = {L}")
        if r_in is None or r_out is None:
            r_in, r_out = shape_bounds(shape, L)
        if not 0 < r_in <= r_out:
            raise ArgumentError(f"the horizon's bounding radii must satisfy 0 < r_in ≤ r_out, got "
                                f"r_in = {r_in}, r_out = {r_out}")
        if not (offset > 0 and thickness > 0):
            raise ArgumentError(f"the offset and the thickness are lengths and must be positive, got "
                                f"offset = {offset} and thickness = {thickness}")
        if not (r_in - offset) - thickness > 0:
            raise ArgumentError(f"the core surface r_h(n̂) − offset − thickness reaches the center: "
                                f"r_in = {r_in}, offset = {offset}, thickness = {thickness}")
        if not (0 < w_ramp <= 1 and 0 < rho_ramp <= 1):
            raise ArgumentError(f"the ramp fractions must lie in (0, 1], got w_ramp = {w_ramp} and "
                                f"ρ_ramp = {rho_ramp}")
        if rho_max < 0:
            raise ArgumentError(f"the relaxation rate must satisfy ρ_max ≥ 0, got {rho_max}")
        self.center = tuple(float(c) for c in center)
        self.shape, self.lmax = shape, L
        self.r_in, self.r_out = float(r_in), float(r_out)
        self.offset, self.thickness = float(offset), float(thickness)
        self.rho_max, self.variant = float(rho_max), variant
        self.w_ramp, self.rho_ramp = float(w_ramp), float(rho_ramp)
        # the bounding radii of the layer, for the checks that read r_0 and r_1
        self.r_0 = (self.r_in - self.offset) - self.thickness
        self.r_1 = self.r_out - self.offset


class StateBounds:
    """StateBounds(T; α_min, α_max, λ_min, λ_max, β_max, K_max, r_gate)
    (src/bounds.jl:51-136): the range projection's ranges and its gate radius."""

    def __init__(self, T=None, *, alpha_min, alpha_max, lambda_min, lambda_max, beta_max, K_max, r_gate):
        _check_float(T)
        vals = [float(x) for x in (alpha_min, alpha_max, lambda_min, lambda_max, beta_max, K_max, r_gate)]
        if not all(math.isfinite(x) for x in vals):
            raise ArgumentError(f"every bound must be finite, got {vals}: a bound at infinity is no "
                                f"bound, and it is spelled by leaving the case's bounds at None instead.")
        a0, a1, l0, l1, b, K, rg = vals
        if not 0 < a0 < 1 < a1:
            raise ArgumentError(f"the lapse range needs 0 < α_min < 1 < α_max, got α_min = {a0} and "
                                f"α_max = {a1}")
        if not 0 < l0 < 1 < l1:
            raise ArgumentError(f"the spatial metric's eigenvalue range needs 0 < λ_min < 1 < λ_max, "
                                f"got λ_min = {l0} and λ_max = {l1}")
        if not b > 0:
            raise ArgumentError(f"the shift's norm bound must be positive, got β_max = {b}")
        if not K > 0:
            raise ArgumentError(f"the momentum's scale bound must be positive, got K_max = {K}")
        if not rg > 0:
            raise ArgumentError(f"the gate radius must be positive, got r_gate = {rg}")
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
                 interior=None, gauge_source=None, bounds=None, velocity=(0.0, 0.0, 0.0)):
        _check_float(T)
        self.velocity = tuple(float(x) for x in velocity)
        if any(self.velocity) and not isinstance(background, (KerrSchild, Harmonic)):
            raise ArgumentError("only a Kerr-Schild or harmonic hole moves")
        if bounds is not None and not isinstance(bounds, StateBounds):
            raise ArgumentError("bounds is None or a StateBounds")
        if bounds is not None and interior is None:
            raise ArgumentError("this case carries range bounds but no interior: the projection is "
                                "gated on r < r_gate from the hole's center")
        self.bounds = bounds
        if not gamma2 > -1:
            raise ArgumentError(f"γ2 must satisfy γ2 > −1, got {gamma2}")
        if isinstance(eps_ko, HorizonDissipation):
            c0 = interior.center if interior is not None else (0.0, 0.0, 0.0)
            if tuple(eps_ko.center) != tuple(c0):
                raise ArgumentError("the dissipation profile is centred on the case's hole")
        elif eps_ko < 0:
            raise ArgumentError(f"the Kreiss–Oliger amplitude must be ≥ 0, got {eps_ko}")
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
    """Flat space on a periodic box of side L (src/initialdata.jl:869-871)."""
    return GHCase(T, Minkowski(), box=[(0.0, L)] * 3, periodic=(True,) * 3, eps_ko=eps_ko,
                  gamma0=gamma0, gamma2=gamma2)


def gauge_wave_case(T=None, *, A=1 / 20, d=1.0, eps_ko, gamma0, gamma2):
    """The AwA gauge wave on one periodic wavelength cubed (src/initialdata.jl:908-911)."""
    return GHCase(T, GaugeWave(A, d), box=[(0.0, d)] * 3, periodic=(True,) * 3, eps_ko=eps_ko,
                  gamma0=gamma0, gamma2=gamma2)


def _moving_hole(v, x0, axis):
    """moving_hole_case's placement (src/initialdata.jl:1283-1298): the hole at
    x0 + v t along axis (1, 2 or 3), the background translate(boost(bg, −v ê), x0 ê)."""
    if axis not in (1, 2, 3):
        raise ArgumentError(f"axis is the coordinate the hole sits and moves on, 1, 2 or 3, got {axis}.")
    v, x0 = float(v), float(x0)
    if not abs(v) < 1:
        raise ArgumentError(f"the hole's velocity must be below the speed of light, got |v| = {abs(v)}")
    on_axis = lambda s: tuple(s if d == axis - 1 else 0.0 for d in range(3))  # noqa: E731
    return on_axis(x0), on_axis(v)


def kerr_schild_case(T=None, *, M=1.0, a=0.0, v=0.0, x0=0.0, axis=1, halfwidth, r_0, r_1, rho_max=None,
                     eps_ko=0.5, gamma0=None, gamma2=0.0, interior="damped", w_ramp=0.5, rho_ramp=0.5,
                     gauge_source=None, bounds=None):
    """kerr_schild_case / hole_case (src/initialdata.jl:1044-1174,1230-1298) for a
    hole in a Dirichlet box, at x0 + v t along axis. ρ_max defaults to the
    physical rate 4/M the driver sets per chunk. gauge_source=None samples
    H_a = −Γ_a of the background once (the reference's default, a hole at rest
    only); "algebraic" is the closed-form KerrSchildSource evaluated in the
    kernel, which moves with the hole."""
    center, vel = _moving_hole(v, x0, axis)
    if any(vel) and gauge_source is None:
        raise ArgumentError("this background is neither harmonic nor static, so its gauge source depends on "
                            "time and the sampled source cannot represent it: give a moving Kerr-Schild hole "
                            "gauge_source=\"algebraic\" (the KerrSchildSource that moves with it), or use "
                            "harmonic_kerr_case.")
    damping = gamma0 if gamma0 is not None else GaussianDamping(near=1 / M, far=1 / (10 * M),
                                                                width=3 * M, center=center)
    inter = Interior(T, center=center, r_0=r_0, r_1=r_1, rho_max=(4.0 / M if rho_max is None else rho_max),
                     variant=interior, w_ramp=w_ramp, rho_ramp=rho_ramp)
    L = float(halfwidth)
    return GHCase(T, KerrSchild(M, a), box=[(-L, L)] * 3, periodic=(False,) * 3, eps_ko=eps_ko,
                  gamma0=damping, gamma2=gamma2, interior=inter, gauge_source=gauge_source,
                  bounds=bounds, velocity=vel)


def harmonic_kerr_case(T=None, *, M=1.0, a=0.0, v=0.0, x0=0.0, axis=1, halfwidth, r_0, r_1, rho_max=None,
                       eps_ko=0.5, gamma0=None, gamma2=0.0, interior="damped", w_ramp=0.5, rho_ramp=0.5,
                       bounds=None):
    """harmonic_kerr_case / hole_case (src/initialdata.jl:1300-1326, 1044-1174):
    Harmonic(M, a) in a Dirichlet box, at x0 + v t along axis, with no gauge
    source (a boost of the harmonic chart is harmonic). Defaults as
    kerr_schild_case's."""
    center, vel = _moving_hole(v, x0, axis)
    damping = gamma0 if gamma0 is not None else GaussianDamping(near=1 / M, far=1 / (10 * M),
                                                                width=3 * M, center=center)
    inter = Interior(T, center=center, r_0=r_0, r_1=r_1, rho_max=(4.0 / M if rho_max is None else rho_max),
                     variant=interior, w_ramp=w_ramp, rho_ramp=rho_ramp)
    L = float(halfwidth)
    return GHCase(T, Harmonic(M, a), box=[(-L, L)] * 3, periodic=(False,) * 3, eps_ko=eps_ko,
                  gamma0=damping, gamma2=gamma2, interior=inter, bounds=bounds, velocity=vel)


def _morton(i, j, k):
    key = 0
    for bit in range(21):
        key |= ((i >> bit) & 1) << (3 * bit) | ((j >> bit) & 1) << (3 * bit + 1) | ((k >> bit) & 1) << (3 * bit + 2)
    return key


class Forest:
    """gh_forest's forest (src/initialdata.jl:1370-1395): a brick of root blocks
    of N cells (roots), uniform, or a frozen hierarchy whose leaves are
    (level, i, j, k), block (i, j, k) of the brick of level l (roots_d 2^l
    blocks of spacing h / 2^l). `leaves` is None on a uniform forest."""

    def __init__(self, case, N, roots, leaves=None):
        roots = (roots,) * 3 if isinstance(roots, int) else tuple(roots)
        if len(roots) != 3 or any(r < 1 for r in roots):
            raise ArgumentError(f"roots is one count of root blocks for every dimension or a "
                                f"brick of three counts, got {roots!r}")
        self.N, self.roots, self.case = int(N), roots, case
        self._leaves = None if leaves is None else set(tuple(int(x) for x in lf) for lf in leaves)
        if case is not None:
            widths = [b[1] - b[0] for b in case.box]
            cell = [widths[d] / roots[d] for d in range(3)]
            if any(abs(c - cell[0]) > 1e-12 * abs(cell[0]) for c in cell):
                raise ArgumentError(f"TreeAMR's blocks are cubes with a single spacing per level, so the "
                                    f"root blocks must be cubes: the box's widths {widths} over the brick "
                                    f"of roots {roots} give root blocks of widths {cell}")

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
        for lf in leafset:
            if lf[0] < 2:
                continue
            for c in self._neighbours(lf):
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


def gh_forest(T=None, case=None, *, N, roots, refined=False):
    """gh_forest(T, case; N, roots, refined = false) (src/initialdata.jl:1370-1395):
    the uniform brick, or with `refined` the two-level mesh whose root block
    holding lo + 3/8 of the box is refined once (every face of it a coarse-fine
    face)."""
    _check_float(T)
    f = Forest(case, N, roots)
    if not refined:
        return f
    xref = [case.box[d][0] + 3 * (case.box[d][1] - case.box[d][0]) / 8 for d in range(3)]
    targets = [lf for lf in f.leaves
               if all(f.block_extent(lf)[d][0] <= xref[d] < f.block_extent(lf)[d][1] for d in range(3))]
    if len(targets) != 1:
        raise ArgumentError(f"the two-level mesh refines exactly the one root block holding {xref}, but "
                            f"{len(targets)} blocks claim it")
    f.refine(targets)
    return f.balance()


def _box_meets_ball(ext, c, R):
    d2 = sum(max(ext[d][0] - c[d], 0.0, c[d] - ext[d][1]) ** 2 for d in range(3))
    return d2 <= R * R


def _box_meets_cube(ext, c, R):
    return all(ext[d][0] < c[d] + R and ext[d][1] > c[d] - R for d in range(3))


def _box_meets_capsule(ext, c, dc, R):
    # the distance from c + s dc to the box is convex in s: ternary search
    def dist2(s):
        return sum(max(ext[d][0] - (c[d] + s * dc[d]), 0.0, (c[d] + s * dc[d]) - ext[d][1]) ** 2
                   for d in range(3))
    a, b = 0.0, 1.0
    for _ in range(200):
        m1, m2 = a + (b - a) / 3, b - (b - a) / 3
        if dist2(m1) <= dist2(m2):
            b = m2
        else:
            a = m1
    return min(dist2(0.0), dist2(1.0), dist2((a + b) / 2)) <= R * R


def _box_meets_swept_cube(ext, c, dc, R):
    # some s in [0, 1] with lo_d < c_d + s dc_d + R and hi_d > c_d + s dc_d - R
    # on every axis: an intersection of open intervals in s
    lo, hi = 0.0, 1.0
    lo_open = hi_open = False
    for d in range(3):
        a, b = ext[d][0] - c[d] - R, ext[d][1] - c[d] + R  # need a < s dc < b
        if dc[d] == 0.0:
            if not (a < 0.0 < b):
                return False
            continue
        s0, s1 = sorted((a / dc[d], b / dc[d]))
        if s0 >= lo:
            lo, lo_open = s0, True
        if s1 <= hi:
            hi, hi_open = s1, True
    return lo < hi or (lo == hi and not (lo_open or hi_open))


def hole_forest(T=None, case=None, *, N, roots, center=None, radii, levels=None, shape="ball", path=None):
    """hole_forest(T, case; N, roots, center, radii, levels, shape, path)
    (src/initialdata.jl:1469-1515): the uniform forest, then for each level
    every leaf of the level below that meets shell radii[l] about center (a
    closed ball, an open cube, or either swept along [center, center + path])
    refined, and the forest balanced after each level."""
    _check_float(T)
    if shape not in ("ball", "cube"):
        raise ArgumentError(f"a shell is a :ball (|x − c| ≤ R) or a :cube (|x − c|_∞ < R), got :{shape}")
    if path is not None and len(path) != 3:
        raise ArgumentError(f"path is None or the displacement Δc of the segment [c, c + Δc], three numbers, "
                            f"got {path!r}")
    radii = [float(r) for r in radii]
    levels = len(radii) if levels is None else int(levels)
    if len(radii) != levels:
        raise ArgumentError(f"hole_forest takes one radius per refinement level, but levels = {levels} and "
                            f"radii has {len(radii)} entries")
    if not all(r > 0 for r in radii):
        raise ArgumentError(f"every shell radius must be positive, got radii = {radii}")
    if not all(radii[l] <= radii[l - 1] for l in range(1, levels)):
        raise ArgumentError(f"the shells must nest, so the radii must not increase with the level, but "
                            f"radii = {radii}")
    if center is None:
        center = case.interior.center if case.interior is not None else (0.0, 0.0, 0.0)
    c = [float(x) for x in center]
    f = Forest(case, N, roots)
    for l in range(1, levels + 1):
        R = radii[l - 1]
        targets = []
        for lf in f.leaves:
            if lf[0] != l - 1:
                continue
            ext = f.block_extent(lf)
            if path is None:
                hit = _box_meets_ball(ext, c, R) if shape == "ball" else _box_meets_cube(ext, c, R)
            else:
                dc = [float(x) for x in path]
                hit = _box_meets_capsule(ext, c, dc, R) if shape == "ball" else _box_meets_swept_cube(ext, c, dc, R)
            if hit:
                targets.append(lf)
        if not targets:
            raise ArgumentError(f"no leaf of level {l - 1} comes within radii[{l}] = {R} of {c}, so shell {l} "
                                f"would refine nothing")
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
            raise ArgumentError(f"the finite-difference order must be even and at least 2, so "
                                f"that the centered stencils have an integer half-width q/2, but q={q}")
        self.backend = backend if backend is not None else MI355X()
        self.forest, self.case, self.q, self.ops = forest, case, int(q), ops
        self.desc = _desc(forest, case, q)
        self._ptr = C.c_void_p()
        import torch
        if forest.refined and ops is None:
            raise ArgumentError(f"a forest with levels needs its ghost operators spelled out, "
                                f"ops = Operators(prolongation = p): the interface-order rule asks for "
                                f"p = q + 2 = {q + 2}, and it is never defaulted")
        with torch.cuda.device(self.backend.device):
            if forest.refined:
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
                raise ArgumentError("a :fitted interior relaxes toward a target the kernel reads from a "
                                    "cache: give GHProblem target = a device array of 40 variables per "
                                    "owned point (A then S)")
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


def error_norms(p, u, t):
    """(L2, L∞) of the masked error u − u_exact(t) over evolved points
    (src/constraints.jl:311-396,655-671)."""
    l2, li = C.c_double(), C.c_double()
    _check(_load().tgh_error_norms(p.ptr, _dev(u, p, "u"), float(t), C.byref(l2), C.byref(li)))
    return l2.value, li.value


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


def adm_constraint_norms(p, u, t):
    """adm_constraint!(p, u, t) + constraint_norms(p)'s ham and mom rows
    (src/constraints.jl:432-534,584-605,848-876): (L2, L∞) of (ℋ, ℳ_x, ℳ_y, ℳ_z)."""
    l2, li = (C.c_double * 4)(), (C.c_double * 4)()
    _check(_load().tgh_adm_constraint(p.ptr, _dev(u, p, "u"), float(t), l2, li))
    return list(l2), list(li)


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


def regrid_flags(p, u, t, *, refine_tol, coarsen_tol, maxlevel_cap, eps=0.01, scale=None):
    """refine_flags (src/refinement.jl:738-795) on the Löhner indicator's block
    maxima: per leaf of p.leaves() +1 (refine: τ above refine_tol below the
    cap), 0 (keep: τ above coarsen_tol) or -1 (coarsen, above level 0)."""
    if not (0 <= coarsen_tol < refine_tol):
        raise ArgumentError(f"the tolerances need 0 <= coarsen_tol < refine_tol, got "
                            f"coarsen_tol={coarsen_tol}, refine_tol={refine_tol}")
    if not 0 <= maxlevel_cap <= 10:
        raise ArgumentError(f"maxlevel_cap must lie in 0 .. 10, got {maxlevel_cap}")
    if scale is None:
        scale = field_scale(p, u, t)
    tau = refinement_indicator(p, u, t, scale=scale, eps=eps)
    flags = []
    for lf, tb in zip(p.leaves(), tau):
        if tb > refine_tol and lf[0] < maxlevel_cap:
            flags.append(1)
        elif tb > coarsen_tol or lf[0] == 0:
            flags.append(0)
        else:
            flags.append(-1)
    return flags, tau


def regrid_forest(forest, leaves, flags):
    """The new mesh of regrid! (src/refinement.jl:973-1021 with TreeAMR's
    refine!/coarsen!/balance!) from per-leaf flags (+1 refine, 0 keep, -1
    coarsen) over `leaves`: the marked leaves refine once, balance! closes
    2:1, and each sibling group whose eight leaves are all marked to coarsen
    merges where no neighbour region of the parent's level is refined below
    the children's level (so the parent stays balanced)."""
    if len(flags) != len(leaves):
        raise ArgumentError(f"one flag per leaf: {len(flags)} flags for {len(leaves)} leaves")
    f = Forest(forest.case, forest.N, forest.roots, leaves)
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


def regrid(p, u, t, *, refine_tol, coarsen_tol, maxlevel_cap, eps=0.01, scale=None):
    """regrid!(forest, U; flags) with the flags of regrid_flags: flagged leaves
    refine once, sibling groups all flagged to coarsen merge where the parent
    stays 2:1 balanced, balance! closes the rest, and the state moves to the
    new leaves (transfer). Returns (the new problem, its state, the flags); a
    regrid that leaves the mesh as it was returns p and u themselves."""
    if p.target is not None:
        raise ArgumentError("a :fitted problem relaxes toward a target cache that the fit computes on its "
                            "own leaves; after a regrid the fit must fill the new mesh's cache, and the "
                            "fit is not built here: regrid a problem without a target")
    flags, _ = regrid_flags(p, u, t, refine_tol=refine_tol, coarsen_tol=coarsen_tol,
                            maxlevel_cap=maxlevel_cap, eps=eps, scale=scale)
    leaves = p.leaves()
    f = regrid_forest(p.forest, leaves, flags)
    if set(f.leaves) == set(leaves):  # the mesh as it was: keep the problem
        return p, u, flags
    new = GHProblem(f, p.case, q=p.q, backend=p.backend,
                    ops=p.ops if p.ops is not None else Operators(prolongation=p.q + 2))
    u_new = new.statevector()
    transfer(p, u, t, new, u_new)
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
    case = getattr(p, "case", None)
    if case is not None and case.interior is not None and case.interior.variant == "pasted":
        # The fused stepper does not run the paste; use the verbs.
        import torch
        k = [torch.empty_like(u) for _ in range(4)]
        for n in range(nsteps):
            t = t0 + n * step
            y = u.clone()
            gh_rhs(k[0], u, p, t)
            torch.add(u, k[0], alpha=step / 2, out=y); stage_limit(y, p, t + step / 2)
            gh_rhs(k[1], y, p, t + step / 2)
            torch.add(u, k[1], alpha=step / 2, out=y); stage_limit(y, p, t + step / 2)
            gh_rhs(k[2], y, p, t + step / 2)
            torch.add(u, k[2], alpha=step, out=y); stage_limit(y, p, t + step)
            gh_rhs(k[3], y, p, t + step)
            u.add_((k[0] + 2 * k[1] + 2 * k[2] + k[3]) * (step / 6)); step_limit(u, p, t + step)
        return u
    _check(_load().tgh_rk4(p.ptr, _dev(u, p, "u"), t0, step, nsteps, TGH_RK4_GRAPH if graph else 0))
    return u


def gh_solve_regrid(p, u, tspan, *, dt, every, refine_tol, coarsen_tol, maxlevel_cap, eps=0.01,
                    graph=False):
    """The driver's regrid cadence (src/driver.jl:1640-1700): gh_solve over
    chunks of `every` steps of the one step size gh_solve would take over
    tspan, each chunk but the last followed by regrid at its end time (the
    hook's t). Returns (the final problem, its state, [(t, nblocks)] per
    regrid)."""
    if every < 1:
        raise ArgumentError(f"the regrid cadence is a positive number of steps, got {every}")
    t0, t1 = float(tspan[0]), float(tspan[1])
    nsteps = max(1, math.ceil((t1 - t0) / dt * (1 - 4 * 2.0 ** -52)))
    step = (t1 - t0) / nsteps
    done, log = 0, []
    while done < nsteps:
        k = min(every, nsteps - done)
        t = t0 + done * step
        _check(_load().tgh_rk4(p.ptr, _dev(u, p, "u"), t, step, k, TGH_RK4_GRAPH if graph else 0))
        done += k
        if done < nsteps:
            t = t0 + done * step
            p, u, _ = regrid(p, u, t, refine_tol=refine_tol, coarsen_tol=coarsen_tol,
                             maxlevel_cap=maxlevel_cap, eps=eps)
            log.append((t, p.nblocks))
    return p, u, log


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
