"""The case descriptor of include/tgh.h as a ctypes structure, and the cases
the tests and the benchmark build (restating the reference's case builders,
src/initialdata.jl:869-1326)."""
import ctypes as C

TGH_BG_MINKOWSKI, TGH_BG_GAUGE_WAVE, TGH_BG_KERR_SCHILD, TGH_BG_SHIFTED_MINKOWSKI, TGH_BG_HARMONIC = 0, 1, 2, 3, 4
TGH_GAUGE_NONE, TGH_GAUGE_SAMPLED, TGH_GAUGE_KERR_SCHILD = 0, 1, 2
TGH_DAMPING_CONSTANT, TGH_DAMPING_GAUSSIAN = 0, 1
TGH_INTERIOR_NONE, TGH_INTERIOR_DAMPED, TGH_INTERIOR_FROZEN, TGH_INTERIOR_PASTED, TGH_INTERIOR_FITTED = 0, 1, 2, 3, 4
TGH_DISSIPATION_CONSTANT, TGH_DISSIPATION_HORIZON = 0, 1


class TghDesc(C.Structure):
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


class TghFit(C.Structure):
    """include/tgh.h's tgh_fit (coeffs: a device pointer for the library, a
    host pointer for the oracle)."""
    _fields_ = [
        ("lmax", C.c_int32), ("cont", C.c_int32), ("tilde", C.c_int32), ("pad", C.c_int32),
        ("rbar", C.c_double), ("center", C.c_double * 3), ("velocity", C.c_double * 3),
        ("alpha_min", C.c_double), ("alpha_max", C.c_double), ("lambda_min", C.c_double),
        ("lambda_max", C.c_double), ("beta_max", C.c_double), ("K_max", C.c_double),
        ("t", C.c_double), ("coeffs", C.c_void_p),
    ]


WIDE_BOUNDS = (1e-3, 1e3, 1e-4, 1e4, 1e3, 1e6)  # ranges no test state reaches


def make_fit(coeffs, *, lmax, cont, rbar=1.0, center=(0.0, 0.0, 0.0), velocity=(0.0, 0.0, 0.0), t=0.0,
             tilde=False, bounds=WIDE_BOUNDS):
    """A tgh_fit over a host coefficient array of shape (20, cont+1, (lmax+1)^2)
    (the reference's ((L+1)^2, cont+1, 20) column-major); the struct keeps
    the array alive as ._host."""
    import numpy as np
    arr = np.ascontiguousarray(coeffs, dtype=np.float64)
    assert arr.shape == (20, cont + 1, (lmax + 1) ** 2)
    f = TghFit()
    f.lmax, f.cont, f.tilde, f.rbar = lmax, cont, 1 if tilde else 0, rbar
    f.center[:], f.velocity[:] = list(center), list(velocity)
    (f.alpha_min, f.alpha_max, f.lambda_min, f.lambda_max, f.beta_max, f.K_max) = bounds
    f.t = t
    f.coeffs = arr.ctypes.data
    f._host = arr
    return f


def make_desc(*, q=4, N=16, nblocks=(1, 1, 1), periodic=(1, 1, 1), lo=(0.0, 0.0, 0.0),
              hi=(1.0, 1.0, 1.0), background=TGH_BG_MINKOWSKI, gauge=TGH_GAUGE_NONE,
              damping=TGH_DAMPING_CONSTANT, gamma0=0.0, gamma0_near=1.0, gamma0_far=0.1,
              gamma0_width=3.0, gamma2=0.0, eps_ko=0.0, interior=TGH_INTERIOR_NONE,
              M=1.0, a=0.0, center=(0.0, 0.0, 0.0), wave_amp=0.05, wave_len=1.0,
              r0=0.0, r1=0.0, rho_max=0.0, w_ramp=0.5, rho_ramp=0.5, dissipation=0, eps_in=0.0,
              eps_r1=0.0, eps_rh=0.0, velocity=(0.0, 0.0, 0.0), reflecting=((0, 0), (0, 0), (0, 0)),
              rotating=0):
    d = TghDesc()
    d.float_bits = 64
    d.q, d.N = q, N
    d.nblocks[:] = list(nblocks)
    d.periodic[:] = list(periodic)
    d.background, d.gauge, d.damping, d.interior = background, gauge, damping, interior
    d.lo[:] = list(lo)
    d.hi[:] = list(hi)
    d.M, d.a = M, a
    d.center[:] = list(center)
    d.wave_amp, d.wave_len = wave_amp, wave_len
    d.gamma0, d.gamma0_near, d.gamma0_far, d.gamma0_width = gamma0, gamma0_near, gamma0_far, gamma0_width
    d.gamma2, d.eps_ko = gamma2, eps_ko
    d.r0, d.r1, d.rho_max, d.w_ramp, d.rho_ramp = r0, r1, rho_max, w_ramp, rho_ramp
    d.dissipation, d.eps_in, d.eps_r1, d.eps_rh = dissipation, eps_in, eps_r1, eps_rh
    d.velocity[:] = list(velocity)
    for dd in range(3):
        d.reflecting[dd][0], d.reflecting[dd][1] = int(reflecting[dd][0]), int(reflecting[dd][1])
    d.rotating = int(rotating)
    return d


def minkowski_noise_case(N=16, q=4):
    """Flat space, periodic unit box, one block: the CPU-runnable workload
    (robust-stability parameters of test/noise_tests.jl:29-30)."""
    return make_desc(q=q, N=N, nblocks=(1, 1, 1), periodic=(1, 1, 1), lo=(0, 0, 0), hi=(1, 1, 1),
                     background=TGH_BG_MINKOWSKI, gamma0=1.0, gamma2=0.0, eps_ko=0.5)


def gauge_wave_case(N=8, roots=1, q=4, eps_ko=0.0, gamma0=1.0, gamma2=0.0, A=0.05, d=1.0):
    """gauge_wave_case (src/initialdata.jl:908-911): one wavelength cubed, periodic."""
    return make_desc(q=q, N=N, nblocks=(roots,) * 3, periodic=(1, 1, 1), lo=(0, 0, 0), hi=(d, d, d),
                     background=TGH_BG_GAUGE_WAVE, gamma0=gamma0, gamma2=gamma2, eps_ko=eps_ko,
                     wave_amp=A, wave_len=d)


def kerr_schild_hole_case(*, N=16, nblocks=16, halfwidth=8.0, a=0.7, M=1.0, q=4,
                          gauge=TGH_GAUGE_KERR_SCHILD, r0=0.71, r1=1.21, rho_max=4.0):
    """hole_case(KerrSchild(M, a)) (src/initialdata.jl:1044-1174): Dirichlet box,
    :damped layer, Gaussian damping near 1/M far 1/(10M) width 3M, eps_KO = 1/2."""
    L = halfwidth
    return make_desc(q=q, N=N, nblocks=(nblocks,) * 3, periodic=(0, 0, 0), lo=(-L, -L, -L), hi=(L, L, L),
                     background=TGH_BG_KERR_SCHILD, gauge=gauge, damping=TGH_DAMPING_GAUSSIAN,
                     gamma0_near=1.0 / M, gamma0_far=1.0 / (10 * M), gamma0_width=3.0 * M, gamma2=0.0,
                     eps_ko=0.5, interior=TGH_INTERIOR_DAMPED, M=M, a=a, r0=r0, r1=r1, rho_max=rho_max)


def harmonic_hole_case(*, N=16, nblocks=4, halfwidth=2.0, a=0.3, M=1.0, q=4, r0=0.35, r1=0.9,
                       rho_max=4.0):
    """harmonic_kerr_case (src/initialdata.jl:1300-1326): hole_case(Harmonic(M, a)) with
    no gauge source, the rest as kerr_schild_hole_case."""
    L = halfwidth
    return make_desc(q=q, N=N, nblocks=(nblocks,) * 3, periodic=(0, 0, 0), lo=(-L, -L, -L), hi=(L, L, L),
                     background=TGH_BG_HARMONIC, gauge=TGH_GAUGE_NONE, damping=TGH_DAMPING_GAUSSIAN,
                     gamma0_near=1.0 / M, gamma0_far=1.0 / (10 * M), gamma0_width=3.0 * M, gamma2=0.0,
                     eps_ko=0.5, interior=TGH_INTERIOR_DAMPED, M=M, a=a, r0=r0, r1=r1, rho_max=rho_max)


def moving_hole_case(*, background=TGH_BG_HARMONIC, a=0.3, v=0.3, x0=0.25, axis=1, N=16, nblocks=4, halfwidth=2.0,
                     M=1.0, q=4, r0=None, r1=None, rho_max=4.0, interior=TGH_INTERIOR_DAMPED):
    """moving_hole_case (src/initialdata.jl:1283-1298): the hole at x0 + v t along axis,
    translate(boost(bg, -v), x0); the Kerr-Schild one with its moving closed-form source."""
    L = halfwidth
    harm = background == TGH_BG_HARMONIC
    r0 = (0.35 if harm else 0.55) if r0 is None else r0
    r1 = (0.9 if harm else 1.05) if r1 is None else r1
    c = [0.0, 0.0, 0.0]
    vel = [0.0, 0.0, 0.0]
    c[axis - 1] = x0
    vel[axis - 1] = v
    return make_desc(q=q, N=N, nblocks=(nblocks,) * 3, periodic=(0, 0, 0), lo=(-L, -L, -L), hi=(L, L, L),
                     background=background, gauge=TGH_GAUGE_NONE if harm else TGH_GAUGE_KERR_SCHILD,
                     damping=TGH_DAMPING_GAUSSIAN, gamma0_near=1.0 / M, gamma0_far=1.0 / (10 * M),
                     gamma0_width=3.0 * M, gamma2=0.0, eps_ko=0.5, interior=interior, M=M, a=a, r0=r0, r1=r1,
                     rho_max=rho_max, center=tuple(c), velocity=tuple(vel))
