"""Pin the oracle (oracle/fdtd_oracle.cpp) against the reference.

CPU only.  Three kinds of check:
  * golden fixtures written by tests/golden/make_golden.py from the reference's
    own compiled functions (FDTD_CPU.cxx / FDTD_common.cxx): bit-exact;
  * the compiled reference itself (oracle/_ref) on further inputs, when built;
  * special cases and closed forms for what only exists in FDTD_GPU.cu.
"""
import math
import os

import numpy as np
import pytest

import tomography_amd as t
from cases import (CYL, CYL_PROBES, CYL_REF, DISC, DISC_PROBES, DISC_TAU, FAT, PHYS_PARAMS, cylinder_far_pattern,
                   cylinder_field, debye_eps, disc_incident_spectrum, fat_medium, fat_with_cylinder,
                   line_source_ratio, phys_small, plane_wave_far_field, plane_wave_scattered, ref_cpu_params,
                   ref_debye_disc, ref_params, ref_small, rel_l2, skin_cylinder, spectrum, tapered)
from oracle.binding import Reference, unpack_ref_cj

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PI = 3.141592653589793238
EPS0 = 8.85418e-12
MU0 = PI * 4e-7
C0 = 1 / math.sqrt(MU0 * EPS0)


def ref_cpu_full():
    return t.preset(t.PRESET_REF_CPU)


def test_ref_presets_are_the_reference_macros():
    # FDTD_common.hxx:3-65, evaluated independently here
    c = t.preset(t.PRESET_REF_GPU)
    assert c.nx == c.ny == math.ceil(12.7 * 0.0315 / 0.001) == 401
    assert c.ntff_offset == math.floor(3.2 * 0.0315 / 0.001) == 100
    assert c.steps == 2000 and c.n_exc == 4 and c.ncpml == 10 and c.n_freq == 10 and c.n_angles == 100
    assert c.dt == 0.001 / C0 / 2
    assert c.sigma_max == 3 / (150 * PI * 0.001)
    assert c.cpml_order == 2 and c.kappa_max == 10 and c.alpha_max == 0.01
    assert c.pulse_tau == 3.3445267e-11 and c.pulse_t0 == math.sqrt(20.0) * 3.3445267e-11
    assert c.debye_tau == 7e-12 and c.numpatches == 3 and c.n_params == 27 and c.ff_norm == 33.329
    assert c.contour_len() == 800 and c.farfield_len() == 4000
    cc = t.preset(t.PRESET_REF_CPU)
    assert cc.numpatches == 9 and cc.n_params == 162 and cc.flags == t.FLAG_CPU_REF_VARIANT


def test_contour_matches_reference(oracle):
    g = np.load(os.path.join(GOLD, "ref_cpu_run.npz"))
    xs, ys, legs = oracle.contour(ref_cpu_full())
    assert np.array_equal(xs, g["contour_x"]) and np.array_equal(ys, g["contour_y"])
    # legs: yn, xp, yp, xn with 199, 201, 199, 201 points (size_cjzy / size_cjzx per frequency)
    assert np.bincount(legs).tolist() == [199, 201, 199, 201]


def test_cpml_tables_bit_exact(oracle):
    g = np.load(os.path.join(GOLD, "ref_cpu_cpml.npz"))["tables"]
    s = oracle.sim(ref_cpu_full(), ref_cpu_params(), exc=0)
    ours = s.cpml()  # kex, kmx, aex, bex, amx, bmx
    assert np.array_equal(ours.astype(np.float32), g)


@pytest.mark.parametrize("name", ["ref_cpu_run", "ref_cpu_dispersionless"])
def test_materials_bit_exact(oracle, name):
    g = np.load(os.path.join(GOLD, name + ".npz"))
    s = oracle.sim(ref_cpu_full(), g["params"], exc=int(g["exc"]))
    mat, coef = s.materials()
    assert np.array_equal(mat, g["patch_map"])
    # coefficient planes: Ceze, Cezhy, ..., Cezeic, Cezeip (oracle order 0,1,3,4)
    for m in np.unique(mat):
        sel = mat == m
        got = np.array([coef[0][sel][0], coef[1][sel][0], coef[3][sel][0], coef[4][sel][0]], np.float32)
        assert np.array_equal(got, g["coef"][m]), m


@pytest.mark.parametrize("name", ["ref_cpu_run", "ref_cpu_dispersionless"])
def test_fields_and_dft_bit_exact_vs_reference(oracle, name):
    """The FP32 oracle in CPU-path mode reproduces FDTD_CPU.cxx bit for bit."""
    g = np.load(os.path.join(GOLD, name + ".npz"))
    s = oracle.sim(ref_cpu_full(), g["params"], exc=int(g["exc"]))
    n = int(g["nsteps"])
    trace = []
    for _ in range(n):
        s.step(1)
        trace.append(s.probe(200, 200))
    trace = np.array(trace, np.float32)
    assert np.array_equal(trace, g["trace"])
    ez, hx, hy, _ = s.fields()
    assert np.array_equal(ez.astype(np.float32)[::4, ::4], g["ez_sub"])
    assert np.sqrt((ez.astype(np.float32).astype(np.float64) ** 2).sum()) == g["ez_l2"]
    assert np.sqrt((hx.astype(np.float32).astype(np.float64) ** 2).sum()) == g["hx_l2"]
    assert np.sqrt((hy.astype(np.float32).astype(np.float64) ** 2).sum()) == g["hy_l2"]
    J, M = s.dft()
    assert np.array_equal(J.astype(np.complex64), g["J"])
    assert np.array_equal(M.astype(np.complex64), g["M"])
    D, _ = s.farfield()
    # N2FPostProcess itself runs in float with double libm; the oracle projects in double
    assert rel_l2(D, g["D"]) < 2e-6


def test_gpu_semantics_reduce_to_cpu_path_without_dispersion(oracle):
    """Special case: delta_eps = 0 makes Beta_p = 0 and Cezjp*Jp = 0, so the
    FDTD_GPU.cu E update with the CPU fwf origin is the FDTD_CPU.cxx update:
    the GPU-semantics oracle must reproduce the reference CPU fixture exactly."""
    g = np.load(os.path.join(GOLD, "ref_cpu_dispersionless.npz"))
    p = g["params"]
    gp = np.concatenate([p[:81], np.zeros(81), p[81:]])  # eps_inf, delta_eps = 0, sigma
    cfg = t.preset(t.PRESET_REF_GPU).copy(numpatches=9, n_params=243,
                                           flags=t.FLAG_FWF_ABS_ORIGIN | t.FLAG_NO_NAN_CLAMP)
    s = oracle.sim(cfg, gp, exc=int(g["exc"]))
    s.step(int(g["nsteps"]))
    ez = s.fields()[0].astype(np.float32)
    assert np.array_equal(ez[::4, ::4], g["ez_sub"])
    J, M = s.dft()
    assert np.array_equal(J.astype(np.complex64), g["J"]) and np.array_equal(M.astype(np.complex64), g["M"])


def test_vacuum_phantom_scatters_nothing(oracle):
    """Special case: eps_inf = 1, delta_eps = sigma = 0 gives Cezeic = Cezeip = 0,
    so the scattered field stays identically zero and D == 0."""
    cfg = ref_small(steps=200)
    p = np.concatenate([np.ones(9), np.zeros(9), np.zeros(9)])
    fit, D, P = oracle.eval(cfg.copy(n_exc=2), p, target=np.zeros(2 * 10 * 100, np.float32))
    assert np.all(D == 0) and fit == 0.0


def test_float_oracle_tracks_fp64_oracle(oracle):
    cfg = ref_small(steps=250)
    p = ref_params(3)
    a, b = oracle.sim(cfg, p, exc=0), oracle.sim(cfg, p, exc=0, fp64=True)
    a.step(250)
    b.step(250)
    for x, y in zip(a.fields(), b.fields()):
        assert rel_l2(x, y) < 1e-5
    assert rel_l2(a.farfield()[0], b.farfield()[0]) < 1e-4


def test_incident_power_is_gaussian_spectrum(oracle):
    """calc_incident_power (FDTD_common.cxx:48-51) is (0.5/eta0)|G(w)|^2 with G the
    Fourier transform of exp(-t^2/tau^2): check against a numerical transform."""
    cfg = t.preset(t.PRESET_REF_GPU)
    tau = cfg.pulse_tau
    eta0 = math.sqrt(MU0 / EPS0)
    tt = np.linspace(-12 * tau, 12 * tau, 200001)
    for f in (1e9, 4e9, 1e10):
        G = np.trapezoid(np.exp(-(tt / tau) ** 2) * np.exp(-2j * PI * f * tt), tt)
        assert abs(oracle.lib.orc_incident_power(ctypes_cfg(cfg), f) / (0.5 / eta0 * abs(G) ** 2) - 1) < 1e-9


def ctypes_cfg(cfg):
    import ctypes
    return ctypes.byref(cfg)


def test_physical_point_source_in_vacuum_is_isotropic(oracle):
    """Closed form: a point source at the centre of an empty domain radiates the
    same far field in every direction.  This pins the running DFT, the closed
    Huygens box, the projection phases and the relative sign of L and eta0*N
    (a wrong sign or a half-cell shift makes the pattern directional)."""
    cfg = phys_small(steps=1000, n_exc=1)
    p = PHYS_PARAMS.copy()
    p[2] = 0.0  # no tumour; the phantom shrinks to the single centre node
    cfgv = cfg.copy(breast_radius=0.0002, skin_thickness=0.0001, src_radius=0.0)
    s = oracle.sim(cfgv, p, exc=0, fp64=True)
    s.step(1000)
    D, _ = s.farfield()
    spread = (D.max(axis=1) - D.min(axis=1)) / D.mean(axis=1)
    assert np.all(spread < 1e-3), spread


def test_physical_dft_phase_convention(oracle):
    """The running DFT is sum_n s*h_n*dt*exp(-j w (n + t_off) dt): check one
    accumulator against a direct transform of the oracle's own samples."""
    cfg = phys_small(steps=300, n_exc=1)
    s = oracle.sim(cfg, PHYS_PARAMS, exc=0, fp64=True)
    xs, ys, legs = oracle.contour(cfg)
    p_idx = 5  # a point on the yn leg: J = +Hx, M = +Ez average
    hx_hist, ez_hist = [], []
    for n in range(300):
        s.step(1)
        ez, hx, _, _ = s.fields()
        x, y = xs[p_idx], ys[p_idx]
        hx_hist.append(hx[y, x])
        ez_hist.append(0.5 * (ez[y, x] + ez[y + 1, x]))
    J, M = s.dft()
    n = np.arange(300)
    for f in range(cfg.n_freq):
        w = 2 * PI * (cfg.f_start + f * (cfg.f_stop - cfg.f_start) / (cfg.n_freq - 1))
        Jd = np.sum(np.array(hx_hist) * cfg.dt * np.exp(-1j * w * (n + 0.5) * cfg.dt))
        Md = np.sum(np.array(ez_hist) * cfg.dt * np.exp(-1j * w * (n + 1.0) * cfg.dt))
        assert abs(J[f, p_idx] - Jd) <= 1e-12 * max(1e-30, abs(Jd)) + 1e-30
        assert abs(M[f, p_idx] - Md) <= 1e-9 * abs(Md)


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built (no reference tree)")
def test_more_vectors_against_compiled_reference(oracle):
    ref = Reference()
    p = ref_cpu_params(99)
    out = ref.run(p, 3, 620)
    s = oracle.sim(ref_cpu_full(), p, exc=3)
    s.step(620)
    ez, hx, hy, _ = s.fields()
    assert np.array_equal(ez.astype(np.float32), out["ez"])
    assert np.array_equal(hx.astype(np.float32), out["hx"]) and np.array_equal(hy.astype(np.float32), out["hy"])
    J, M = s.dft()
    Jr, Mr = unpack_ref_cj(out["cj"])
    assert np.array_equal(J.astype(np.complex64), Jr.astype(np.complex64))


def test_planted_mistake_is_caught(oracle):
    """A wrong CPML order (a plausible restatement slip) must break the
    bit-exact comparison with the reference fixture."""
    g = np.load(os.path.join(GOLD, "ref_cpu_cpml.npz"))["tables"]
    s = oracle.sim(ref_cpu_full().copy(cpml_order=3), ref_cpu_params(), exc=0)
    assert not np.array_equal(s.cpml().astype(np.float32), g)


def _debye_coefficients(einf, deps, tau, sigma, dt, dx):
    """First principles: the trapezoidal-in-time discretisation at one node of
        eps0 einf dE/dt + sigma E + Jp = curl H,   tau dJp/dt + Jp = eps0 deps dE/dt,
    solved as a 2x2 linear system for (E^{n+1}, Jp^{n+1}) from (E^n, Jp^n,
    curl).  Returns the E coefficients on E^n, on (dx curl), on Jp^n (negated)
    and the Jp recursion Jp^{n+1} = Kp Jp^n + g (E^{n+1} - E^n)."""
    a = EPS0 * einf / dt
    if tau > 0:
        q = tau / dt
        kp, g = (q - 0.5) / (q + 0.5), EPS0 * deps / dt / (q + 0.5)
    else:
        kp, g = 0.0, 0.0
    lhs = np.array([[a + sigma / 2, 0.5], [-g, 1.0]])
    col = {}
    for name, e0, j0, cu in (("E", 1.0, 0.0, 0.0), ("J", 0.0, 1.0, 0.0), ("C", 0.0, 0.0, 1.0)):
        rhs = np.array([(a - sigma / 2) * e0 - 0.5 * j0 + cu, kp * j0 - g * e0])
        col[name] = np.linalg.solve(lhs, rhs)
    # curl = (dHy - dHx) / dx, so the coefficient of (dHy - dHx) is col["C"][0] / dx
    return col["E"][0], col["C"][0], -col["J"][0], kp, g


def test_physical_debye_coefficients_follow_from_the_discretised_pdes(oracle):
    """The physical-mode E / Debye-current coefficients of every tissue equal
    the solution of the trapezoidal discretisation of Ampere's law with a
    single-pole Debye current (an independent 2x2 linear solve in FP64, not the
    closed-form expressions of the restatement), to FP32 rounding."""
    cfg = phys_small()
    p = np.array([3.0, -2.0, 4.0, 25.0, 40.0, 10.0, 0.5, 0.0])  # tumour einf, deps, tau (ps), sigma
    s = oracle.sim(cfg, p, exc=0)
    mat, oc = s.materials()
    tissues = {1: (36.0, 0.0, 0.0, 4.0), 2: (3.1, 1.6, 13e-12, 0.0), 3: (12.8, 24.6, 13e-12, 0.0),
               4: (25.0, 40.0, 10e-12, 0.5)}
    seen = 0
    for m, (ei, de, tau, sg) in tissues.items():
        sel = mat == m
        if not sel.any():
            continue
        seen += 1
        ceze, cezh, cezjp, kp, g = _debye_coefficients(ei, de, tau, sg, cfg.dt, cfg.dx)
        got = [oc[0][sel][0], oc[1][sel][0] * cfg.dx, oc[5][sel][0], oc[8][sel][0], oc[7][sel][0]]
        want = [ceze, cezh, cezjp, kp, g]
        for gv, wv, name in zip(got, want, ("Ceze", "Cezh*dx", "Cezjp", "Kp", "Bp/dt")):
            assert abs(gv - wv) <= 2e-7 * abs(wv) + 1e-30, (m, name, gv, wv)
    assert seen == 4


def _cpml_reflection(oracle, **mod):
    """Relative difference of a probe trace 5 cells in front of the right CPML
    of a 128^2 grid against the same source/probe offsets in a 512^2 grid whose
    boundary is too far away for a reflection to return within the run."""
    p = PHYS_PARAMS.copy()
    p[2] = 0.0
    steps = 260
    traces, offset = [], None
    for n in (128, 512):
        cfg = phys_small(steps=steps, n=n, n_exc=1).copy(breast_radius=0.0002, skin_thickness=0.0001,
                                                        src_radius=0.010, **mod)
        sx, sy = oracle.source_cell(cfg, 0)
        if offset is None:
            offset = (n - cfg.ncpml - 5 - 1) - sx  # probe - source, in cells, fixed by the small grid
        s = oracle.sim(cfg, p, exc=0, fp64=True)
        tr = []
        for _ in range(steps):
            s.step(1)
            tr.append(s.probe(sx + offset, sy))
        traces.append(np.array(tr))
    small, big = traces
    assert np.abs(big).max() > 0 and offset > 10
    return np.abs(small - big).max() / np.abs(big).max()


def test_physical_cpml_reflection_is_small(oracle):
    """Closed form by construction: the only difference between a small grid
    and a large one, at a probe in front of the small grid's CPML and before
    any reflection can return in the large grid, is the CPML's reflection.  The
    graded profile (m = 3, R0 = 1e-6) measures 3.5e-6 of the peak field here;
    the bar is 1e-4.  Planted mistake: the same profile without its loss
    (sigma_max = 0, a lossless coordinate stretch) must reflect visibly."""
    assert _cpml_reflection(oracle) < 1e-4
    assert _cpml_reflection(oracle, sigma_max=0.0) > 1e-2


def _ref_gpu_cfg(flags):
    c = t.preset(t.PRESET_REF_GPU)
    return c.copy(flags=c.flags | flags)


def _fitness_case(seed, lo=0.01, hi=0.05):
    rng = np.random.default_rng(seed)
    L = t.preset(t.PRESET_REF_GPU).farfield_len()
    return rng.uniform(lo, hi, L).astype(np.float32), rng.uniform(lo, hi, L).astype(np.float32)


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built (no reference tree)")
def test_generic_fitness_matches_compiled_fdtd_common_fitness(oracle):
    """fitness() of FDTD_common.cxx:53-61, compiled from the reference, on
    random D and measurement arrays: the oracle with TOMO_FLAG_REL_FITNESS |
    TOMO_FLAG_FITNESS_PER_EXC agrees within the reference's float accumulation
    (4000 same-sign terms: relative bound 4000 * 2^-24 = 2.4e-4, observed far
    below).  Planted mistakes: the per-frequency average (REL alone) or no
    division by m^2 (PER_EXC alone) miss by far more."""
    ref = Reference()
    for seed in (1, 2, 3):
        D, m = _fitness_case(seed)
        want = ref.fitness(D, m)
        got = oracle.fitness(_ref_gpu_cfg(t.FLAG_REL_FITNESS | t.FLAG_FITNESS_PER_EXC), D, m)
        assert abs(got - want) <= 2.4e-4 * abs(want), (got, want)
        for wrong in (t.FLAG_REL_FITNESS, t.FLAG_FITNESS_PER_EXC):
            bad = oracle.fitness(_ref_gpu_cfg(wrong), D, m)
            assert abs(bad - want) > 0.5 * abs(want), (wrong, bad, want)


@pytest.mark.skipif(not Reference.available(), reason="oracle/_ref not built (no reference tree)")
def test_gpu_fitness_at_unit_measurement_is_compiled_fitness_over_n_freq(oracle):
    """FDTD_GPU.cu:1004-1008, -sum (D - m)^2 / (n_exc n_freq), reduced to the
    compiled fitness() of FDTD_common.cxx:53-61, -sum (m - D)^2 / (n_exc m^2),
    at m = 1: the two differ by exactly the factor 1 / n_freq.  What stays
    pinned by the source alone is that factor (FDTD_GPU.cu:1008 divides by
    numberoffrequencies as well as numberofexcitationangles) and the absence
    of m^2, which m = 1 cannot see."""
    ref = Reference()
    cfg = _ref_gpu_cfg(0)
    for seed in (4, 5):
        D, _ = _fitness_case(seed, 0.2, 1.8)
        ones = np.ones_like(D)
        want = ref.fitness(D, ones) / cfg.n_freq
        got = oracle.fitness(cfg, D, ones)
        assert abs(got - want) <= 2.4e-4 * abs(want), (got, want)


def test_cpu_fitness_matches_compiled_fdtd_cpu_golden_value(oracle):
    """The value the reference's whole FDTD_CPU() returned in a fresh process
    (tests/golden/ref_fdtd_cpu_fitness.json): its far-field loop never runs
    (FDTD_CPU.cxx:816), so it is the fitness of FDTD_CPU.cxx:841-845 at D = 0,
    -sum (0 - m)^2 / (m^2 n_exc n_freq) accumulated in float, = -100.0038.
    The oracle's FP64 form at D = 0 gives -100 exactly (4000 / 40), within
    the float accumulation bound 4000 * ulp(100) / 2 = 0.0153 of it; dividing
    by n_exc n_freq without m^2, or by m^2 n_exc alone, misses by orders of
    magnitude.  The float accumulator restated on the same measurement gives
    the golden value bit for bit."""
    import json
    with open(os.path.join(GOLD, "ref_fdtd_cpu_fitness.json")) as f:
        gold = json.load(f)
    cfg = ref_cpu_full()
    m = t.measurement("cpu")[:cfg.farfield_len()]
    D0 = np.zeros(cfg.farfield_len())
    got = oracle.fitness(cfg, D0, m)
    assert abs(got - (-100.0)) <= 1e-9
    assert abs(got - gold["fitness"]) <= 4000 * 2.0 ** -17 / 2, (got, gold["fitness"])
    for wrong in (cfg.flags & ~t.FLAG_REL_FITNESS, cfg.flags | t.FLAG_FITNESS_PER_EXC):
        bad = oracle.fitness(cfg.copy(flags=wrong), D0, m)
        assert abs(bad - gold["fitness"]) > 50, (wrong, bad)
    # the reference's float accumulator: fit -= (double) pow(D - m, 2) / (pow(m, 2) * 4 * 10)
    acc = np.float32(0.0)
    for x in m.astype(np.float64):
        acc = np.float32(np.float64(acc) - (0.0 - x) ** 2 / (x ** 2 * cfg.n_exc * cfg.n_freq))
    assert float(acc) == gold["fitness"]


def test_phase_fitness_term_special_cases(oracle):
    """The phase term sketched at FDTD_GPU.cu:1009-1018: with D equal to the
    target, each entry adds -dphi^2 / (10 pi^2 n_exc n_freq) with dphi = P -
    P_target wrapped once into [-pi, pi] (dphi > pi: -2 pi; dphi < -pi: +2 pi).
    Offsets 0.3, -2.0, 3 pi / 2 (-> -pi / 2) and -3 pi / 2 (-> pi / 2)."""
    cfg = _ref_gpu_cfg(t.FLAG_PHASE_FITNESS)
    L = cfg.farfield_len()
    n = cfg.n_exc * cfg.n_freq
    rng = np.random.default_rng(9)
    D = rng.uniform(0.01, 0.05, L).astype(np.float32)
    tp = rng.uniform(-1.0, 1.0, L).astype(np.float32)
    for off, wrapped in ((0.3, 0.3), (-2.0, -2.0), (1.5 * PI, -0.5 * PI), (-1.5 * PI, 0.5 * PI)):
        P = tp.astype(np.float64) + off
        got = oracle.fitness(cfg, D, D, P=P, target_phase=tp)
        want = -L * wrapped ** 2 / (10 * PI ** 2 * n)
        assert abs(got - want) <= 1e-6 * abs(want), (off, got, want)
    # without the flag the phase is ignored
    assert oracle.fitness(_ref_gpu_cfg(0), D, D, P=tp.astype(np.float64) + 1.0, target_phase=tp) == 0.0


FAT_FREQS = (1.5e9, 2.0e9, 2.5e9, 3.0e9, 3.5e9, 4.0e9)
FAT_MAG_TOL, FAT_PHASE_TOL = 1e-3, 2e-3  # relative |ratio|, radians


def test_physical_debye_medium_propagation_matches_hankel_closed_form(oracle):
    """Closed form for the whole physical-mode update (Yee leapfrog, Debye ADE,
    soft source): in a homogeneous Debye medium a line current radiates
    H0^(2)(k rho) with k = w sqrt(mu0 eps(w)), so the ratio of the field's
    spectra 40 and 20 cells from the source depends only on eps(w).  The FP64
    oracle meets it to 2e-4 (|ratio|) and 7e-4 rad; the same medium without its
    dispersion (eps_inf alone, or the static eps_inf + delta_eps) misses by
    1-9% and up to 0.3 rad, so a dropped or mis-signed polarisation current
    fails here."""
    cfg, p = fat_medium()
    s = oracle.sim(cfg, p, exc=0, fp64=True)
    mat, _ = s.materials()
    assert np.all(mat == 2), "the grid must hold the fat tissue only"
    sx, sy = oracle.source_cell(cfg, 0)
    d1, d2, every = 20, 40, 4
    tr = []
    for _ in range(cfg.steps // every):
        s.step(every)
        tr.append((s.probe(sx + d1, sy), s.probe(sx + d2, sy), s.probe(sx, sy + d1), s.probe(sx, sy + d2)))
    tr = np.array(tr)
    assert np.abs(tr[-50:]).max() < 1e-4 * np.abs(tr).max()  # the pulse has passed both probes
    for f in FAT_FREQS:
        want = line_source_ratio(f, d1, d2, cfg.dx, *FAT)
        for a, b in ((0, 1), (2, 3)):  # along x and along y
            got = spectrum(tr[:, b], f, every * cfg.dt) / spectrum(tr[:, a], f, every * cfg.dt)
            assert abs(abs(got) / abs(want) - 1) < FAT_MAG_TOL, (f, got, want)
            assert abs(np.angle(got / want)) < FAT_PHASE_TOL, (f, got, want)
        for wrong in ((FAT[0], 0.0, FAT[2]), (FAT[0] + FAT[1], 0.0, FAT[2])):
            bad = line_source_ratio(f, d1, d2, cfg.dx, *wrong)
            got = spectrum(tr[:, 1], f, every * cfg.dt) / spectrum(tr[:, 0], f, every * cfg.dt)
            assert abs(abs(got) / abs(bad) - 1) > FAT_MAG_TOL or abs(np.angle(got / bad)) > FAT_PHASE_TOL



def test_physical_debye_cylinder_scattering_matches_mie_series(oracle):
    """Closed form with material interfaces: a Debye cylinder (radius 6 mm,
    12 cells) in the fat medium, lit by the antenna 10 mm away.  The field
    spectra at six probes around it (behind, beside, in front), each over a
    reference probe, match the cylinder's series solution (Bessel/Hankel, both
    media dispersive; radius = the area-equivalent radius of the rasterised
    disc) within 1% at 1.5-3.5 GHz (measured: <= 0.3%).  Without the cylinder,
    or with the cylinder's dispersion dropped, the series misses by 5-140%, and
    the test asserts that it does."""
    cfg, p = fat_with_cylinder()
    s = oracle.sim(cfg, p, exc=0, fp64=True)
    mat, _ = s.materials()
    assert set(np.unique(mat)) == {2, 4}
    a = np.sqrt((mat == 4).sum() / np.pi) * cfg.dx  # area-equivalent radius of the staircased disc
    sx, sy = oracle.source_cell(cfg, 0)
    every = 4
    pts = CYL_PROBES + [CYL_REF]
    tr = []
    for _ in range(cfg.steps // every):
        s.step(every)
        tr.append([s.probe(sx + u, sy + v) for u, v in pts])
    tr = np.array(tr)
    centre = (10e-3, 0.0)
    for f in (1.5e9, 2.5e9, 3.5e9):
        ref = spectrum(tr[:, -1], f, every * cfg.dt)
        for i, (u, v) in enumerate(CYL_PROBES):
            got = spectrum(tr[:, i], f, every * cfg.dt) / ref
            pos, rpos = (u * cfg.dx, v * cfg.dx), (CYL_REF[0] * cfg.dx, CYL_REF[1] * cfg.dx)
            want = cylinder_field(f, pos, centre, a, FAT, CYL) / cylinder_field(f, rpos, centre, a, FAT, CYL)
            assert abs(got / want - 1) < 1e-2, (f, (u, v), got, want)
            for wrong in (None, (CYL[0], 0.0, CYL[2])):
                bad = cylinder_field(f, pos, centre, a, FAT, wrong) / cylinder_field(f, rpos, centre, a, FAT, wrong)
                assert abs(got / bad - 1) > 3e-2, (f, (u, v), wrong)



def check_cylinder_far_field(D, cfg, mat, src, tol_l2=1e-2, tol_max=2e-2):
    """D[freq][angle] of the skin cylinder against its series far field, both
    normalised to their mean over the angles (the source spectrum and the far-
    field constants cancel; so does the choice of phase origin)."""
    a = np.sqrt((mat != 0).sum() / np.pi) * cfg.dx  # area-equivalent radius of the staircased disc
    phi = 2 * np.pi * np.arange(cfg.n_angles) / cfg.n_angles
    for fi in range(cfg.n_freq):
        f = cfg.f_start + fi * (cfg.f_stop - cfg.f_start) / (cfg.n_freq - 1)
        eps = 36.0 - 1j * 4.0 / (2 * np.pi * f * 8.854187817e-12)
        g = D[fi] / D[fi].mean()
        w = cylinder_far_pattern(f, phi, src, a, eps)
        w = w / w.mean()
        assert np.linalg.norm(g - w) / np.linalg.norm(w) < tol_l2, (f, np.linalg.norm(g - w) / np.linalg.norm(w))
        assert np.abs(g / w - 1).max() < tol_max, f
        for wrong in (None, 36.0 + 0j):  # no cylinder; the cylinder without its conductivity
            b = cylinder_far_pattern(f, phi, src, a, wrong)
            b = b / b.mean()
            assert np.linalg.norm(g - b) / np.linalg.norm(b) > 5 * tol_l2, (f, wrong)


def test_physical_far_field_of_lossy_cylinder_matches_series(oracle):
    """Closed form for the whole far-field chain (fields, running DFT on the
    closed Huygens box, projection, L and eta0 N combined): a lossy skin
    cylinder (radius 6 mm) in vacuum lit by the antenna 16 mm away.  D(phi),
    normalised over the angles, matches the cylinder's series far field within
    1% (relative L2; measured <= 0.13%, and <= 0.5% at any angle) at every
    frequency; without the cylinder, or without its conductivity, it misses by
    8-76%, and the test asserts that it does."""
    cfg, p = skin_cylinder()
    s = oracle.sim(cfg, p, exc=0, fp64=True)
    mat, _ = s.materials()
    assert (mat == 1).sum() > 400 and (mat > 1).sum() <= 1  # skin, plus at most the centre node
    sx, sy = oracle.source_cell(cfg, 0)
    s.step(cfg.steps)
    D, _ = s.farfield()
    check_cylinder_far_field(D, cfg, mat, ((sx - cfg.nx // 2) * cfg.dx, (sy - cfg.ny // 2) * cfg.dy))


DISC_FREQS = (2.0e9, 2.5e9, 3.0e9)


def check_disc_scattering(tr, cfg, mat, every, tol=1.5e-2):
    """Probe records of the Debye disc (one column per DISC_PROBES entry, a
    sample every `every` steps from the start) against the plane-wave series.

    * Shape: each probe's spectrum over the first probe's, against the same
      ratio of the series: within `tol`.  The disc without its dispersion
      (static eps_inf + delta_eps), without delta_eps, without sigma, with
      delta_eps x 0.9 or with tau x 1.2 misses by more than 2 tol.
    * Absolute: the spectrum itself over the incident pulse's spectrum at the
      disc centre, against the series: within `tol`.  After m steps Ez is the
      field at (m - 1/2) dt (the E update of step i uses the incident wave at
      i + 1/2), so sample n sits at (every (n + 1) - 1/2) dt.  The incident
      wave's origin at the grid corner instead of its centre misses by far more.
    """
    a = np.sqrt((mat != 0).sum() / np.pi) * cfg.dx  # area-equivalent radius of the staircased disc
    tr = tapered(tr)
    dts = every * cfg.dt
    pos = [(u * cfg.dx, v * cfg.dx) for u, v in DISC_PROBES]
    ei, de, sg = DISC
    wrongs = ((ei + de, 0, sg, DISC_TAU), (ei, 0, sg, DISC_TAU), (ei, de, 0, DISC_TAU), (ei, 0.9 * de, sg, DISC_TAU),
              (ei, de, sg, 1.2 * DISC_TAU))
    miss = np.zeros(len(wrongs))
    corner = 0.0
    for f in DISC_FREQS:
        w = 2 * np.pi * f
        X = np.array([spectrum(tr[:, i], f, dts) for i in range(len(pos))]) * dts * np.exp(-1j * w * (every - 0.5) * cfg.dt)
        S = np.array([plane_wave_scattered(f, q, a, debye_eps(f, *DISC, DISC_TAU)) for q in pos])
        shape = np.abs((X[1:] / X[0]) / (S[1:] / S[0]) - 1)
        assert shape.max() < tol, (f, shape)
        absolute = np.abs(X / (disc_incident_spectrum(cfg, f) * S) - 1)
        assert absolute.max() < tol, (f, absolute)
        corner = max(corner, np.abs(X / (disc_incident_spectrum(cfg, f, origin_cells=cfg.nx // 2) * S) - 1).max())
        for k, wrong in enumerate(wrongs):
            Sw = np.array([plane_wave_scattered(f, q, a, debye_eps(f, *wrong)) for q in pos])
            miss[k] = max(miss[k], np.abs((X[1:] / X[0]) / (Sw[1:] / Sw[0]) - 1).max())
    assert corner > 10 * tol, corner
    assert miss.min() > 2 * tol, miss


def test_ref_gpu_debye_disc_scattering_matches_plane_wave_series(oracle):
    """Closed form for the FDTD_GPU.cu update (scattered-field formulation,
    Debye polarisation current Jp, incident wave fwf with its origin at the
    grid centre): a homogeneous Debye disc lit by the plane wave of excitation
    0.  Ten probes around the disc (forward, sideways, backward) match the
    disc's plane-wave series within 1.5% at 2-3 GHz, both as ratios between
    probes and absolutely against the incident pulse's spectrum (measured:
    <= 0.9% and <= 0.75%).  Dropping the disc's dispersion, its delta_eps or
    its sigma, or scaling delta_eps by 0.9 or tau by 1.2, misses by 16-960%,
    and the incident wave with its origin at the grid corner misses absolutely
    by 200% at 2.5 GHz (300 cells are two and three wavelengths at 2 and 3 GHz)."""
    cfg, p = ref_debye_disc()
    s = oracle.sim(cfg, p, exc=0, fp64=True)
    mat, _ = s.materials()
    assert set(np.unique(mat)) == set(range(10))
    c, every = cfg.nx // 2, 2
    tr = []
    for _ in range(cfg.steps // every):
        s.step(every)
        tr.append([s.probe(c + u, c + v) for u, v in DISC_PROBES])
    check_disc_scattering(np.array(tr), cfg, mat, every)


def check_disc_far_field(D, P, cfg, mat, tol=1.5e-2):
    """The complex far field sqrt(D) e^{jP} of the Debye disc against the
    series far field -sum_n c_n e^{jn phi} (the far-field form of
    plane_wave_scattered), up to one complex constant per frequency (fitted by
    least squares), at 2-5 GHz.  N2FPostProcess places a contour point at
    (x - nx/2 + 1) dx with nx/2 in float, while the disc is centred on node
    nx/2 in integers, so its phase centre sits o = (nx//2 - nx/2 + 1) dx off the
    disc's centre along both axes and the series carries e^{j k0 o (cos phi +
    sin phi)}.  Without that factor, or with the opposite phase convention
    (the conjugate), the fit misses by more than 2 tol."""
    a = np.sqrt((mat != 0).sum() / np.pi) * cfg.dx
    phi = 2 * np.pi * np.arange(cfg.n_angles) / cfg.n_angles
    o = (cfg.nx // 2 - np.float32(cfg.nx) / 2 + 1) * cfg.dx
    worst_plain = worst_conj = 0.0
    for fi in range(1, 5):
        f = cfg.f_start + fi * (cfg.f_stop - cfg.f_start) / (cfg.n_freq - 1)
        k0 = 2 * np.pi * f / C0
        E = np.sqrt(np.asarray(D[fi], np.float64)) * np.exp(1j * np.asarray(P[fi], np.float64))
        F = plane_wave_far_field(f, phi, a, debye_eps(f, *DISC, DISC_TAU))
        shift = np.exp(1j * k0 * o * (np.cos(phi) + np.sin(phi)))

        def miss(T):
            c = np.vdot(T, E) / np.vdot(T, T)
            return np.linalg.norm(E - c * T) / np.linalg.norm(E)
        assert miss(F * shift) < tol, (f, miss(F * shift))
        worst_plain = max(worst_plain, miss(F))
        worst_conj = max(worst_conj, miss(np.conj(F * shift)))
    assert worst_plain > 2 * tol and worst_conj > 2 * tol, (worst_plain, worst_conj)


def test_ref_gpu_far_field_phase_of_debye_disc_matches_series(oracle):
    """Closed form for N2FPostProcess's phase output P together with D: the
    Debye disc's far field after 2500 steps (the scattered field has crossed
    the contour) against its series far field.  Measured: <= 0.9% (complex
    relative L2, 2-5 GHz; bar 1.5%); without the reference's half-cell phase
    centre offset 1.9-4.1%, conjugated 4-76%."""
    cfg, p = ref_debye_disc(steps=2500)
    s = oracle.sim(cfg, p, exc=0, fp64=True)
    mat, _ = s.materials()
    s.step(cfg.steps)
    D, P = s.farfield()
    check_disc_far_field(np.asarray(D).reshape(cfg.n_freq, cfg.n_angles),
                         np.asarray(P).reshape(cfg.n_freq, cfg.n_angles), cfg, mat)
