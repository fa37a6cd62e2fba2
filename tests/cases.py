"""Small configurations shared by the parity tests (sizes the oracle runs in seconds)."""
import numpy as np

import tomography_amd as t


def ref_small(steps=300, n=101, flags=0):
    """REF_GPU physics on an n x n grid (disc r = 31.5 mm, dx = 1 mm)."""
    c = t.preset(t.PRESET_REF_GPU)
    return c.copy(nx=n, ny=n, ntff_offset=12, steps=steps, flags=flags)


def ref_cpu_small(steps=300, n=101):
    c = t.preset(t.PRESET_REF_CPU)
    return c.copy(nx=n, ny=n, ntff_offset=12, steps=steps)


def phys_small(steps=512, n=128, n_exc=2, flags=0):
    """Physical mode on a 64 mm domain (dx = 0.5 mm) with a small phantom."""
    c = t.preset(t.PRESET_CFG1)
    return c.copy(nx=n, ny=n, steps=steps, n_exc=n_exc, n_freq=4, n_angles=36, breast_radius=0.018,
                  skin_thickness=0.0015, src_radius=0.022, flags=flags)


def ref_params(seed=1, npch=3):
    rng = np.random.default_rng(seed)
    n2 = npch * npch
    return np.concatenate([rng.uniform(1.0, 7.0, n2), rng.uniform(1.5, 41.0, n2), rng.uniform(0.0, 0.3, n2)])


def ref_cpu_params(seed=1):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.uniform(1.0, 40.0, 81), rng.uniform(0.0, 1.1, 81)])


PHYS_PARAMS = np.array([3.0, -2.0, 4.0, 20.0, 34.0, 13.0, 0.2, 0.0])


def rel_l2(a, b):
    """||a - b|| / ||b|| in double precision (complex-aware)."""
    a = np.asarray(a).astype(np.complex128)
    b = np.asarray(b).astype(np.complex128)
    n = np.sqrt((np.abs(b) ** 2).sum())
    return float(np.sqrt((np.abs(a - b) ** 2).sum()) / (n if n > 0 else 1.0))


# A 200^2 grid (dx = 0.5 mm) filled with the Debye fat tissue and nothing
# else: the breast outline lies outside the grid, no tumour block, and phantom
# seed offset 126 puts all three fibroglandular ellipses outside the grid too
# (the tests check the material map).  The antenna sits at the centre node.
FAT = (3.1, 1.6, 13e-12)  # eps_inf, delta_eps, tau of the fat tissue (sigma = 0)
FAT_OFFSET = 126


def fat_medium(steps=4000):
    c = phys_small(steps=steps, n=200, n_exc=1)
    p = np.array([0.0, 0.0, 0.0, 20.0, 34.0, 13.0, 0.0, float(FAT_OFFSET)])
    return c.copy(breast_radius=0.08, skin_thickness=0.0001, src_radius=0.0), p


def line_source_ratio(f, d1, d2, dx, eps_inf, d_eps, tau):
    """E(d2)/E(d1) of a time-harmonic line current in a homogeneous Debye medium:
    H0^(2)(k d2) / H0^(2)(k d1), k = w sqrt(mu0 eps0 (eps_inf + d_eps / (1 + j w tau)))
    (phasors e^{+jwt}, the convention of a DFT with kernel e^{-jwt})."""
    from scipy.special import hankel2
    w = 2 * np.pi * f
    k = w * np.sqrt(4e-7 * np.pi * 8.854187817e-12 * (eps_inf + d_eps / (1 + 1j * w * tau)))
    return hankel2(0, k * d2 * dx) / hankel2(0, k * d1 * dx)


def spectrum(samples, f, dt_sample):
    n = np.arange(len(samples))
    return np.sum(np.asarray(samples, np.float64) * np.exp(-2j * np.pi * f * n * dt_sample))


def debye_k(f, eps_inf, d_eps, tau):
    w = 2 * np.pi * f
    return w * np.sqrt(4e-7 * np.pi * 8.854187817e-12 * (eps_inf + d_eps / (1 + 1j * w * tau)))


def cylinder_field(f, pos, centre, a, background, cylinder, n_terms=30):
    """Ez (up to a constant factor) at `pos` (m, relative to a line current at
    the origin) in a homogeneous Debye background holding a Debye cylinder of
    radius a centred at `centre` (m), e^{+jwt} phasors:
        E = H0^(2)(kb |r|) + sum_n a_n H_n^(2)(kb rho) e^{jn(phi - phi_s)},
        a_n = -H_n^(2)(kb d) [kb J_n'(kb a) J_n(k1 a) - k1 J_n(kb a) J_n'(k1 a)]
                            / [kb H_n^(2)'(kb a) J_n(k1 a) - k1 H_n^(2)(kb a) J_n'(k1 a)]
    with (rho, phi) the probe and (d, phi_s) the source in cylinder-centred polar
    coordinates (continuity of Ez and of dEz/drho at rho = a).  `cylinder` =
    None gives the background alone."""
    from scipy.special import h2vp, hankel2, jv, jvp
    kb = debye_k(f, *background)
    E = hankel2(0, kb * np.hypot(*pos))
    if cylinder is None:
        return E
    k1 = debye_k(f, *cylinder)
    rs = (-centre[0], -centre[1])
    d, phs = np.hypot(*rs), np.arctan2(rs[1], rs[0])
    r = (pos[0] - centre[0], pos[1] - centre[1])
    rho, ph = np.hypot(*r), np.arctan2(r[1], r[0])
    for n in range(-n_terms, n_terms + 1):
        cn = (kb * jvp(n, kb * a) * jv(n, k1 * a) - k1 * jv(n, kb * a) * jvp(n, k1 * a)) / \
             (kb * h2vp(n, kb * a) * jv(n, k1 * a) - k1 * hankel2(n, kb * a) * jvp(n, k1 * a))
        E = E + (-hankel2(n, kb * d) * cn) * hankel2(n, kb * rho) * np.exp(1j * n * (ph - phs))
    return E


# A Debye cylinder (a tumour block: centre 10 mm from the antenna, radius 6 mm,
# eps_inf 8, delta_eps 4, tau 13 ps, sigma 0) in the fat medium above, and the
# probe offsets (cells from the antenna) around it: behind, beside, in front.
CYL = (8.0, 4.0, 13e-12)
CYL_PROBES = [(40, 0), (20, 24), (20, -24), (-20, 0), (30, 18), (50, 10)]
CYL_REF = (0, -30)


def fat_with_cylinder(steps=4000):
    cfg, p = fat_medium(steps)
    p = p.copy()
    p[:7] = [10.0, 0.0, 6.0, CYL[0], CYL[1], CYL[2] * 1e12, 0.0]
    return cfg, p


def skin_cylinder(steps=2000):
    """Vacuum 128^2 grid holding one cylinder of the skin tissue (eps_r 36,
    sigma 4 S/m): breast radius 6 mm, skin 5.9 mm (the interior shrinks to the
    centre node), no tumour block; the antenna 16 mm from its centre."""
    c = phys_small(steps=steps, n=128, n_exc=1)
    p = PHYS_PARAMS.copy()
    p[2] = 0.0
    return c.copy(breast_radius=0.006, skin_thickness=0.0059, src_radius=0.016), p


def cylinder_far_pattern(f, phi, src, a, eps_c):
    """|F(phi)|^2 of a line current at `src` (m, relative to the centre of a
    cylinder of radius a and complex relative permittivity eps_c, None = no
    cylinder) in vacuum: F = e^{j k0 d cos(phi - phi_s)} + sum_n a_n j^n e^{jn(phi - phi_s)},
    a_n as in cylinder_field (the far-field forms of H0^(2)(k0|r - r_s|) and
    H_n^(2)(k0 rho))."""
    from scipy.special import h2vp, hankel2, jv, jvp
    w = 2 * np.pi * f
    k0 = w * np.sqrt(4e-7 * np.pi * 8.854187817e-12)
    d, phs = np.hypot(*src), np.arctan2(src[1], src[0])
    F = np.exp(1j * k0 * d * np.cos(phi - phs)).astype(complex)
    if eps_c is not None:
        k1 = k0 * np.sqrt(eps_c)
        for n in range(-40, 41):
            cn = (k0 * jvp(n, k0 * a) * jv(n, k1 * a) - k1 * jv(n, k0 * a) * jvp(n, k1 * a)) / \
                 (k0 * h2vp(n, k0 * a) * jv(n, k1 * a) - k1 * hankel2(n, k0 * a) * jvp(n, k1 * a))
            F = F + (-hankel2(n, k0 * d) * cn) * (1j ** n) * np.exp(1j * n * (phi - phs))
    return np.abs(F) ** 2


# A homogeneous Debye disc under the FDTD_GPU.cu physics (scattered-field
# formulation; excitation 0 is a Gaussian plane wave travelling along +x):
# radius 15.75 mm (31.5 cells of 1 mm), eps_inf 4, delta_eps 10, sigma 0.2 S/m
# in all nine patches, and a relaxation time of 40 ps, which makes the disc
# strongly dispersive over 2-3 GHz.  The 601^2 grid puts the CPML far enough
# out that its echoes reach the probes (30 cells from the centre) only in the
# last part of the 1300-step record, which a cosine taper removes.
DISC = (4.0, 10.0, 0.2)  # eps_inf, delta_eps, sigma (S/m)
DISC_TAU = 40e-12
DISC_PROBES = [(30, 0), (21, 21), (0, 30), (-21, 21), (-30, 0), (-21, -21), (0, -30), (21, -21), (26, 15),
               (-28, -10)]
C0 = 1.0 / np.sqrt(4e-7 * np.pi * 8.854187817e-12)


def ref_debye_disc(steps=1300, n=601):
    c = ref_small(steps=steps, n=n).copy(breast_radius=0.01575, debye_tau=DISC_TAU)
    p = np.concatenate([np.full(9, DISC[0]), np.full(9, DISC[1]), np.full(9, DISC[2])])
    return c, p


def debye_eps(f, eps_inf, d_eps, sigma, tau):
    """Complex relative permittivity eps_inf + d_eps / (1 + j w tau) - j sigma / (w eps0)."""
    w = 2 * np.pi * f
    return eps_inf + d_eps / (1 + 1j * w * tau) - 1j * sigma / (w * 8.854187817e-12)


def plane_wave_scattered(f, pos, a, eps_c, n_terms=40):
    """Ez scattered at `pos` (m from the disc centre) by a disc of radius a and
    complex relative permittivity eps_c in vacuum, for a unit plane wave
    e^{-j k0 x} (e^{+jwt} phasors, travelling along +x, phase zero at the
    centre): sum_n (-j)^n (-c_n) H_n^(2)(k0 rho) e^{jn phi}, c_n as in
    cylinder_field (the plane wave's Jacobi-Anger coefficients (-j)^n J_n)."""
    from scipy.special import h2vp, hankel2, jv, jvp
    k0 = 2 * np.pi * f / C0
    k1 = k0 * np.sqrt(eps_c)
    rho, ph = np.hypot(*pos), np.arctan2(pos[1], pos[0])
    S = 0j
    for m in range(-n_terms, n_terms + 1):
        cn = (k0 * jvp(m, k0 * a) * jv(m, k1 * a) - k1 * jv(m, k0 * a) * jvp(m, k1 * a)) / \
             (k0 * h2vp(m, k0 * a) * jv(m, k1 * a) - k1 * hankel2(m, k0 * a) * jvp(m, k1 * a))
        S = S + (-1j) ** m * (-cn) * hankel2(m, k0 * rho) * np.exp(1j * m * ph)
    return S


def plane_wave_far_field(f, phi, a, eps_c, n_terms=40):
    """Far-field form of plane_wave_scattered without its sqrt(2 / (pi k0 rho))
    e^{-j(k0 rho - pi/4)} factor: H_n^(2)(k0 rho) -> j^n, so the pattern is
    -sum_n c_n e^{jn phi}."""
    from scipy.special import h2vp, hankel2, jv, jvp
    k0 = 2 * np.pi * f / C0
    k1 = k0 * np.sqrt(eps_c)
    F = np.zeros(np.shape(phi), complex)
    for m in range(-n_terms, n_terms + 1):
        cn = (k0 * jvp(m, k0 * a) * jv(m, k1 * a) - k1 * jv(m, k0 * a) * jvp(m, k1 * a)) / \
             (k0 * h2vp(m, k0 * a) * jv(m, k1 * a) - k1 * hankel2(m, k0 * a) * jvp(m, k1 * a))
        F = F - cn * np.exp(1j * m * np.asarray(phi))
    return F


def disc_incident_spectrum(cfg, f, origin_cells=0.0):
    """Spectrum at the disc centre of the incident wave exp(-((t - t0 - (x + R)/c0)/tau)^2)
    with x measured from the grid centre (origin_cells shifts that origin):
    tau sqrt(pi) e^{-(w tau / 2)^2} e^{-jw (t0 + (R + origin_cells dx)/c0)}."""
    w = 2 * np.pi * f
    tp = cfg.pulse_tau
    delay = cfg.pulse_t0 + (cfg.breast_radius + origin_cells * cfg.dx) / C0
    return tp * np.sqrt(np.pi) * np.exp(-(w * tp) ** 2 / 4) * np.exp(-1j * w * delay)


def tapered(tr, frac=0.3):
    """The record with a cosine taper over its last `frac`."""
    tr = np.asarray(tr, np.float64)
    n = len(tr)
    m = int(n * frac)
    w = np.ones(n)
    w[n - m:] = 0.5 * (1 + np.cos(np.pi * np.arange(m) / m))
    return tr * w[:, None]
