// setup.cpp — presets and per-simulation host setup.
//
// Reference correspondences:
//   presets REF_GPU/REF_CPU : the macros of FDTD_common.hxx:3-65
//   unpack_ref              : FDTD_GPU.cu:655-693 (27 values), FDTD_CPU.cxx:681-698 (162 values)
//   ref raster inputs       : disc test FDTD_GPU.cu:724-727, getOptimizationCell FDTD_GPU.cu:544-551
//   ref coefficients        : FDTD_GPU.cu:739-741 (host) and 1118-1120 (scattered_parameter_init);
//                             CPU forms FDTD_CPU.cxx:893-905, 936-959
//   ref CPML profiles       : FDTD_GPU.cu:752-806
//   ref far-field normaliser: calc_incident_power FDTD_common.cxx:48-51, 33.329 at FDTD_GPU.cu:626
// The physical-mode phantom, CPML grading and point source are this
// library's own (they have no reference counterpart).
#include "setup.h"

#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

namespace tomo {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

double c0() { return 1.0 / std::sqrt(kMU0 * kEPS0); }
double eta0() { return std::sqrt(kMU0 / kEPS0); }

int contour_len(const tomo_sim_config& c) { return 2 * c.nx + 2 * c.ny - 8 * c.ntff_offset - 4; }

static void ref_base(tomo_sim_config* c) {
    std::memset(c, 0, sizeof(*c));
    const double breast_radius = 0.0315, dx = 0.001;
    c->mode = TOMO_MODE_REF_COMPAT;
    c->nx = c->ny = (int)std::ceil(12.7 * breast_radius / dx);     // 401
    c->steps = 2000;
    c->n_exc = 4;
    c->ncpml = 10;
    c->ntff_offset = (int)std::floor(3.2 * breast_radius / dx);    // 100
    c->n_freq = 10;
    c->n_angles = 100;
    c->numpatches = 3;
    c->n_params = 27;
    c->dx = c->dy = dx;
    c->dt = dx / c0() / 2;
    c->f_start = 1e9;
    c->f_stop = 1e10;
    c->cpml_order = 2;
    c->kappa_max = 10;
    c->alpha_max = 0.01;
    c->sigma_max = (c->cpml_order + 1) / (150 * kPI * dx);
    c->pulse_tau = 3.3445267e-11;
    c->pulse_t0 = std::sqrt(20.0) * c->pulse_tau;
    c->debye_tau = 7e-12;
    c->breast_radius = breast_radius;
    c->ff_norm = 33.329;
}

static void phys_base(tomo_sim_config* c, int n, double dx, int steps, uint64_t seed) {
    std::memset(c, 0, sizeof(*c));
    c->mode = TOMO_MODE_PHYSICAL;
    c->nx = c->ny = n;
    c->steps = steps;
    c->n_exc = 1;
    c->ncpml = 10;
    c->ntff_offset = c->ncpml + 5;          // contour 5 cells inside the CPML
    c->n_freq = 8;
    c->n_angles = 180;
    c->numpatches = 0;
    c->n_params = 8;
    c->dx = c->dy = dx;
    c->dt = 0.99 * dx / (c0() * std::sqrt(2.0));
    c->f_start = 1e9;
    c->f_stop = 4e9;
    c->cpml_order = 3;
    c->kappa_max = 5;
    c->alpha_max = 0.05;
    const double R0 = 1e-6;
    c->sigma_max = -(c->cpml_order + 1) * std::log(R0) / (2 * eta0() * c->ncpml * dx);
    const double fc = 2.5e9;
    c->pulse_tau = 1.0 / (std::sqrt(2.0) * kPI * fc);   // spectrum peak of the differentiated Gaussian at fc
    c->pulse_t0 = 4.0 * c->pulse_tau;
    c->debye_tau = 13e-12;
    c->breast_radius = 0.040;
    c->skin_thickness = 0.002;
    c->src_radius = 0.050;
    c->ff_norm = 1.0;
    c->seed = seed;
}

int preset(int id, int grid, tomo_sim_config* out) {
    if (!out) return fail(TOMO_E_INVALID, "null output");
    switch (id) {
        case TOMO_PRESET_REF_GPU:
            ref_base(out);
            return TOMO_OK;
        case TOMO_PRESET_REF_CPU:
            ref_base(out);
            out->flags = TOMO_FLAG_CPU_REF_VARIANT;
            out->numpatches = 9;
            out->n_params = 162;
            return TOMO_OK;
        case TOMO_PRESET_CFG1:
            phys_base(out, 256, 0.5e-3, 2048, 20261202ull);
            return TOMO_OK;
        case TOMO_PRESET_CFG2:
            phys_base(out, 1024, 0.25e-3, 8192, 20261202ull);
            return TOMO_OK;
        case TOMO_PRESET_CFG3:
            phys_base(out, 512, 0.5e-3, 4096, 20261203ull);
            out->n_exc = 16;
            out->src_radius = 0.055;
            out->n_freq = 16;
            out->n_angles = 360;
            return TOMO_OK;
        case TOMO_PRESET_CFG4:
        case TOMO_PRESET_CFG5: {
            int n = grid > 0 ? grid : 512;
            if (n < 64 || n > 8192) return fail(TOMO_E_INVALID, "grid out of range");
            int steps = 8 * n;
            steps = steps < 2048 ? 2048 : (steps > 16384 ? 16384 : steps);
            phys_base(out, n, 0.5e-3, steps, id == TOMO_PRESET_CFG4 ? 20261204ull : 20261205ull);
            return TOMO_OK;
        }
        default:
            return fail(TOMO_E_INVALID, "unknown preset");
    }
}

int validate(const tomo_sim_config& c) {
    if (c.mode != TOMO_MODE_REF_COMPAT && c.mode != TOMO_MODE_PHYSICAL) return fail(TOMO_E_INVALID, "bad mode");
    if (c.nx < 16 || c.ny < 16 || c.nx > 32768 || c.ny > 32768) return fail(TOMO_E_INVALID, "grid size out of range");
    if (c.ncpml < 1 || c.ncpml > TOMO_MAX_NC - 1 || 2 * c.ncpml + 4 > c.nx || 2 * c.ncpml + 4 > c.ny)
        return fail(TOMO_E_INVALID, "bad CPML thickness");
    if (c.n_freq < 1 || c.n_freq > TOMO_MAX_FREQ) return fail(TOMO_E_INVALID, "n_freq out of range");
    if (c.n_angles < 1 || c.n_angles > 4096) return fail(TOMO_E_INVALID, "n_angles out of range");
    if (c.n_exc < 1 || c.n_exc > 64) return fail(TOMO_E_INVALID, "n_exc out of range");
    if (c.steps < 1 || c.steps > (1 << 24)) return fail(TOMO_E_INVALID, "steps out of range");
    if (c.ntff_offset < 1 || c.nx - 2 * c.ntff_offset - 2 < 1 || c.ny - 2 * c.ntff_offset - 2 < 1)
        return fail(TOMO_E_INVALID, "NTFF contour does not fit");
    if (!(c.dx > 0) || !(c.dy > 0) || !(c.dt > 0)) return fail(TOMO_E_INVALID, "non-positive dx/dy/dt");
    if (c.dx != c.dy) return fail(TOMO_E_INVALID, "square cells required (dx == dy)");
    if (c.mode == TOMO_MODE_REF_COMPAT) {
        if (c.nx != c.ny) return fail(TOMO_E_INVALID, "ref-compat mode requires nx == ny");
        if (c.numpatches < 1 || 1 + c.numpatches * c.numpatches > TOMO_MAX_MAT)
            return fail(TOMO_E_INVALID, "numpatches out of range for the device material table");
        if (!(c.debye_tau > 0)) return fail(TOMO_E_INVALID, "debye_tau must be positive");
    } else {
        if (!(c.breast_radius > c.skin_thickness) || !(c.skin_thickness >= 0))
            return fail(TOMO_E_INVALID, "bad phantom radii");
    }
    return TOMO_OK;
}

// ---------------------------------------------------------------- seeded phantom
namespace {
struct SplitMix {
    uint64_t s;
    explicit SplitMix(uint64_t seed) : s(seed) {}
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};
const Tissue kVac{1.0, 0.0, 0.0, 0.0};
const Tissue kSkin{36.0, 0.0, 0.0, 4.0};
const Tissue kFat{3.1, 1.6, 13e-12, 0.0};
const Tissue kFibro{12.8, 24.6, 13e-12, 0.0};
const Tissue kTum{20.0, 34.0, 13e-12, 0.0};
}  // namespace

std::vector<double> target_vector(const tomo_sim_config& c) {
    SplitMix r(c.seed ^ 0x7A26E7ull);
    std::vector<double> v(8, 0.0);
    v[0] = -15.0 + 30.0 * r.u01();
    v[1] = -15.0 + 30.0 * r.u01();
    v[2] = 2.0 + 4.0 * r.u01();
    v[3] = kTum.einf;
    v[4] = kTum.deps;
    v[5] = kTum.tau * 1e12;
    v[6] = 0.0;
    v[7] = 0.0;
    return v;
}

// Phantom from an 8k-dim vector: block k = tumour k (cx mm, cy mm, r mm,
// eps_inf, delta_eps, tau ps, sigma S/m, [block 0 only] ellipse seed offset).
static int build_phantom(const tomo_sim_config& c, const double* p, size_t np, SimSetup& s) {
    if (np < 8) return fail(TOMO_E_SHORT, "physical mode needs at least 8 parameters");
    for (size_t i = 0; i < np; ++i)
        if (!std::isfinite(p[i])) return fail(TOMO_E_INVALID, "non-finite parameter");
    const double Ro = c.breast_radius, Ri = c.breast_radius - c.skin_thickness;
    PhantomGeom& g = s.geom;
    g.ro2 = (float)(Ro * Ro);
    g.ri2 = (float)(Ri * Ri);
    uint64_t off = (p[7] > 0) ? (uint64_t)std::floor(p[7]) : 0;
    SplitMix r((c.seed ^ 0xE111B5E5ull) + off);
    for (int e = 0; e < 3; ++e) {
        const double rho = (Ri - 0.012) * std::sqrt(r.u01());
        const double ph = 2.0 * kPI * r.u01();
        const double a = 0.004 + 0.008 * r.u01();
        const double b = 0.004 + 0.008 * r.u01();
        const double th = kPI * r.u01();
        g.ell[e][0] = (float)(rho * std::cos(ph));
        g.ell[e][1] = (float)(rho * std::sin(ph));
        g.ell[e][2] = (float)std::cos(th);
        g.ell[e][3] = (float)std::sin(th);
        g.ell[e][4] = (float)(1.0 / (a * a));
        g.ell[e][5] = (float)(1.0 / (b * b));
    }
    s.tissues = {kVac, kSkin, kFat, kFibro};
    g.ntum = 0;
    for (int k = 0; k < 4; ++k) {
        const size_t b = 8 * (size_t)k;
        if (np <= b) break;
        auto at = [&](size_t i, double d) { return b + i < np ? p[b + i] : d; };
        double rmm = at(2, 4.0);
        if (!(rmm > 0.0)) continue;
        rmm = rmm < 2.0 ? 2.0 : (rmm > 6.0 ? 6.0 : rmm);
        Tissue t{at(3, kTum.einf), at(4, kTum.deps), at(5, kTum.tau * 1e12) * 1e-12, at(6, kTum.sigma)};
        if (!(t.einf >= 1.0) || !(t.deps >= 0.0) || !(t.tau > 0.0) || !(t.sigma >= 0.0))
            return fail(TOMO_E_INVALID, "tumour Debye parameters out of range");
        const double rr = rmm * 1e-3;
        g.tum[g.ntum][0] = (float)(at(0, 0.0) * 1e-3);
        g.tum[g.ntum][1] = (float)(at(1, 0.0) * 1e-3);
        g.tum[g.ntum][2] = (float)(rr * rr);
        s.tissues.push_back(t);
        g.ntum++;
    }
    return TOMO_OK;
}

// Exact FP32 disc test of the device rasteriser (no contraction: explicit
// rounding at every step, identical on host and device).
static inline bool in_disc_f32(int x, int y, const tomo_sim_config& c, float r2) {
    volatile float X = (float)(x - c.nx / 2) * (float)c.dx;
    volatile float Y = (float)(y - c.ny / 2) * (float)c.dy;
    volatile float a = X * X;
    volatile float b = Y * Y;
    volatile float s = a + b;
    return s <= r2;
}

static void physical_spans(const tomo_sim_config& c, SimSetup& s) {
    // Dispersive tissue is painted only inside the inner (fat) disc.
    s.span.assign(3 * (c.ny + 1), 0);
    int off = 0;
    const int cx = c.nx / 2;
    for (int y = 0; y <= c.ny; ++y) {
        int lo = 0, hi = 0;
        if (in_disc_f32(cx, y, c, s.geom.ri2)) {
            int l = cx, h = cx;
            while (l - 1 >= 0 && in_disc_f32(l - 1, y, c, s.geom.ri2)) --l;
            while (h + 1 <= c.nx && in_disc_f32(h + 1, y, c, s.geom.ri2)) ++h;
            lo = l; hi = h + 1;
        }
        s.span[3 * y] = lo;
        s.span[3 * y + 1] = hi;
        s.span[3 * y + 2] = off;
        off += hi - lo;
    }
    s.jp_len = off;
}

// Reference disc test and patch index, evaluated with the reference's own
// mixture of float storage and double arithmetic (FDTD_GPU.cu:724-731).
static int ref_raster(const tomo_sim_config& c, SimSetup& s) {
    const int nx = c.nx, ny = c.ny, np = c.numpatches;
    const double dx = c.dx, dy = c.dy, R = c.breast_radius;
    RefRaster& rr = s.ref;
    rr.lo.assign(ny + 1, 0);
    rr.hi.assign(ny + 1, -1);
    rr.xc.assign(nx + 1, 0);
    rr.yc.assign(ny + 1, 0);
    for (int i = 0; i <= nx; ++i) rr.xc[i] = (int)((i - (nx / 2 - (int)(R / dx))) / (2 * R / (np * dx)));
    for (int j = 0; j <= ny; ++j) rr.yc[j] = (int)((j - (ny / 2 - R / dy)) / (2 * R / (np * dy)));
    s.span.assign(3 * (ny + 1), 0);
    int off = 0;
    for (int j = 0; j <= ny; ++j) {
        int lo = -1, hi = -2;
        for (int i = 0; i <= nx; ++i) {
            float radius = (float)std::sqrt(std::pow(((float)i - nx / 2) * dx, 2) + std::pow(((float)j - ny / 2) * dy, 2));
            if (radius <= R) {
                if (lo < 0) lo = i;
                hi = i;
                const int patch = rr.xc[i] + np * rr.yc[j];
                if (rr.xc[i] < 0 || rr.xc[i] >= np || rr.yc[j] < 0 || rr.yc[j] >= np || patch < 0 || patch >= np * np)
                    return fail(TOMO_E_INVALID, "patch index out of range");
            }
        }
        if (lo >= 0) { rr.lo[j] = lo; rr.hi[j] = hi; }
        const int slo = lo >= 0 ? lo : 0, shi = lo >= 0 ? hi + 1 : 0;
        s.span[3 * j] = slo;
        s.span[3 * j + 1] = shi;
        s.span[3 * j + 2] = off;
        off += shi - slo;
    }
    s.jp_len = off;
    return TOMO_OK;
}

static void set_coef(SimSetup& s, int m, double ceze, double cezh, double cezeic, double cezeip, double cezjp,
                     double kp, double bpdt, bool disp) {
    float* q = s.coef + m * TOMO_NCOEF;
    q[0] = (float)ceze; q[1] = (float)cezh; q[2] = (float)cezeic; q[3] = (float)cezeip;
    q[4] = (float)cezjp; q[5] = (float)kp; q[6] = (float)bpdt; q[7] = disp ? 1.f : 0.f;
}

// Reference coefficient table, one row per material (0 = vacuum, 1 + patch).
static int ref_coefs(const tomo_sim_config& c, const double* p, size_t np, SimSetup& s) {
    const int npch = c.numpatches, n2 = npch * npch;
    const bool cpu = (c.flags & TOMO_FLAG_CPU_COEFS) != 0;
    const bool eps_sigma = (c.flags & TOMO_FLAG_PARAMS_EPS_SIGMA) != 0;
    const size_t need = (size_t)n2 * (eps_sigma ? 2 : 3);
    if (np < need) return fail(TOMO_E_SHORT, "parameter vector shorter than " + std::to_string(need));
    s.image.resize(need);
    for (size_t k = 0; k < need; ++k) {
        if (!std::isfinite(p[k])) return fail(TOMO_E_INVALID, "non-finite parameter");
        s.image[k] = (float)p[k];
    }
    const double dt = c.dt, dx = c.dx, dy = c.dy, tr = c.debye_tau;
    const double Kp = (1 - dt / (2 * tr)) / (1 + dt / (2 * tr));
    s.nmat = 1 + n2;
    for (int m = 0; m < s.nmat; ++m) {
        if (!cpu) {
            float ei = 1, de = 0, sg = 0;
            if (m) { ei = s.image[m - 1]; de = s.image[m - 1 + n2]; sg = s.image[m - 1 + 2 * n2]; }
            if (!(ei > 0) || !(sg >= 0) || !(de >= 0)) return fail(TOMO_E_INVALID, "eps_inf must be > 0, sigma and delta_eps >= 0");
            const float Bp = (float)(kEPS0 * de * dt / (tr * (1 + dt / (2 * tr))));
            const double dh = 2 * ei * kEPS0 + sg * dt + (double)Bp;
            const double dd = 2 * kEPS0 * ei + sg * dt + (double)Bp;
            set_coef(s, m, (2 * ei * kEPS0 - sg * dt + (double)Bp) / dh, (2 * dt / dx) / dh,
                     (-2 * kEPS0 * ei + 2 * kEPS0 - sg * dt - (double)Bp) / dd,
                     (2 * kEPS0 * ei - 2 * kEPS0 - sg * dt + (double)Bp) / dd, ((1 + Kp) * dt) / dd, Kp,
                     (double)Bp / dt, de > 0);
        } else {
            float er = 1, sg = 0;
            if (m) { er = s.image[m - 1]; sg = s.image[m - 1 + n2]; }
            if (!(er > 0) || !(sg >= 0)) return fail(TOMO_E_INVALID, "eps_r must be > 0 and sigma >= 0");
            const double den = 2 * er * kEPS0 + dt * sg;
            const double den2 = 2 * kEPS0 * er + sg * dt;
            set_coef(s, m, (2 * er * kEPS0 - dt * sg) / den, (2 * dt / dx) / den,
                     (2 * (kEPS0 - kEPS0 * er) - sg * dt) / den2, -1 * (2 * (kEPS0 - kEPS0 * er) + sg * dt) / den2,
                     0.0, 0.0, 0.0, false);
        }
    }
    (void)dy;
    return TOMO_OK;
}

static void phys_coefs(const tomo_sim_config& c, SimSetup& s) {
    const double dt = c.dt;
    s.nmat = (int)s.tissues.size();
    for (int m = 0; m < s.nmat; ++m) {
        const Tissue& t = s.tissues[m];
        double Bp = 0, Kp = 0;
        if (t.deps > 0 && t.tau > 0) {
            Bp = kEPS0 * t.deps * dt / (t.tau * (1 + dt / (2 * t.tau)));
            Kp = (1 - dt / (2 * t.tau)) / (1 + dt / (2 * t.tau));
        }
        const double den = 2 * t.einf * kEPS0 + t.sigma * dt + Bp;
        set_coef(s, m, (2 * t.einf * kEPS0 - t.sigma * dt + Bp) / den, 2 * dt / (c.dx * den), 0, 0,
                 (1 + Kp) * dt / den, Kp, Bp / dt, Bp > 0);
    }
}

// CPML profiles of FDTD_GPU.cu:752-806 with its float/double mix.
static void ref_cpml(const tomo_sim_config& c, SimSetup& s) {
    const int nc = c.ncpml;
    const double dx = c.dx, dt = c.dt, m = c.cpml_order, kmax = c.kappa_max, amax = c.alpha_max;
    const float smax = (float)c.sigma_max;
    for (int i = 0; i < nc; ++i) {
        float rho = (float)(((float)i + 0.25) / nc);
        const float sig_e = (float)(smax * 1.0 * std::pow((double)rho, m));
        const float ke = (float)(std::pow((double)rho, m) * (kmax - 1) + 1);
        const float al_e = (float)(0.0 + (amax - 0.0) * rho);
        rho = (float)(((float)i + 0.75) / nc);
        const float sig_m = (float)((kMU0 / kEPS0) * smax * 1.0 * std::pow((double)rho, m));
        const float km = (float)(std::pow((double)rho, m) * (kmax - 1) + 1);
        const float al_m = (float)((kMU0 / kEPS0) * (0.0 + (amax - 0.0) * rho));
        const float be = (float)std::exp(-1 * (dt / kEPS0) * (double)(sig_e / ke + al_e));
        const float ae = (float)((double)((be - 1) * sig_e) / (dx * (double)(sig_e * ke + al_e * ke * ke)));
        const float arg = (float)(-1 * (dt / kMU0) * (double)((sig_m / km) + al_m));
        const float bm = std::exp(arg);  // single-precision exp of a float argument, as the reference host code
        const float am = (float)((double)((bm - 1) * sig_m) / (dx * (double)(sig_m * km + al_m * km * km)));
        s.ke[i] = ke; s.km[i] = km; s.ike[i] = 1.0f / ke; s.ikm[i] = 1.0f / km;
        s.ae[i] = ae; s.be[i] = be; s.am[i] = am; s.bm[i] = bm;
    }
}

static void phys_cpml(const tomo_sim_config& c, SimSetup& s) {
    const int nc = c.ncpml;
    for (int i = 0; i < nc; ++i) {
        for (int h = 0; h < 2; ++h) {
            const double rho = ((double)i + (h ? 0.75 : 0.25)) / nc;
            const double g = std::pow(rho, c.cpml_order);
            const double sig = c.sigma_max * g;
            const double kap = 1.0 + (c.kappa_max - 1.0) * g;
            const double al = c.alpha_max * (1.0 - rho);
            const double b = std::exp(-(sig / kap + al) * c.dt / kEPS0);
            const double a = (sig == 0.0 && al == 0.0) ? 0.0 : sig * (b - 1.0) / (c.dx * (sig * kap + al * kap * kap));
            if (h) { s.km[i] = (float)kap; s.ikm[i] = (float)(1.0 / kap); s.bm[i] = (float)b; s.am[i] = (float)a; }
            else { s.ke[i] = (float)kap; s.ike[i] = (float)(1.0 / kap); s.be[i] = (float)b; s.ae[i] = (float)a; }
        }
    }
}

// Products of the configuration alone (shared by every simulation of it):
// the physical source waveform and its spectrum normaliser, the reference
// disc extents and patch coordinates.
struct ConfigProducts {
    std::vector<float> srcw;
    double inv_norm[TOMO_MAX_FREQ] = {};
    RefRaster ref;
    std::vector<int> span;
    int jp_len = 0;
    bool has_ref = false, has_phys = false;
};

static std::mutex g_cp_mu;
static std::map<std::string, std::shared_ptr<ConfigProducts>> g_cp;

static std::shared_ptr<ConfigProducts> config_products(const tomo_sim_config& c) {
    tomo_sim_config k = c;
    k.n_params = 0;
    k.n_exc = 0;  // per-excitation quantities are not cached
    const std::string key((const char*)&k, sizeof(k));
    std::lock_guard<std::mutex> g(g_cp_mu);
    auto it = g_cp.find(key);
    if (it != g_cp.end()) return it->second;
    if (g_cp.size() > 256) g_cp.clear();
    auto p = std::make_shared<ConfigProducts>();
    g_cp[key] = p;
    return p;
}

static double phys_source_value(const tomo_sim_config& c, int n) {
    const double t = (n + 1) * c.dt;
    const double u = (t - c.pulse_t0) / c.pulse_tau;
    return -u * std::exp(-u * u);
}

int build_setup(const tomo_sim_config& c, const double* p, size_t np, int exc, SimSetup& s) {
    int rc = validate(c);
    if (rc) return rc;
    if (exc < 0 || exc >= c.n_exc) return fail(TOMO_E_INVALID, "excitation index out of range");
    if (!p && np) return fail(TOMO_E_INVALID, "null parameters");
    s.cfg = c;
    s.exc = exc;
    s.pitch = ((c.nx + 1) + 31) & ~31;
    const bool ref = c.mode == TOMO_MODE_REF_COMPAT;
    // NTFF contour (FDTD_common.cxx:4-46).  The reference's x faces start one
    // cell below the bottom face and its y faces stop one cell short of the
    // right face; physical mode closes the box with the same point count.
    const int D = c.ntff_offset;
    s.cx_hi = ref ? c.nx - D - 2 : c.nx - D - 1;
    s.cy_lo = ref ? D : D + 1;
    s.L1 = s.cx_hi - D;
    s.L2 = s.L1 + (c.ny - D - s.cy_lo);
    s.L3 = s.L2 + (s.cx_hi - D);
    s.P = s.L3 + (c.ny - D - s.cy_lo);
    const double dt = c.dt;
    for (int f = 0; f < c.n_freq; ++f) {
        const double df = c.n_freq > 1 ? (c.f_stop - c.f_start) / (c.n_freq - 1) : 0.0;
        double fr = c.f_start + df * f;
        if (ref) fr = (double)(float)fr;  // the reference keeps the frequency in a float
        s.freq[f] = fr;
        s.wdt[f] = 2 * kPI * fr * dt;
    }
    if (ref) {
        if ((rc = ref_coefs(c, p, np, s))) return rc;
        auto cp = config_products(c);
        {
            std::lock_guard<std::mutex> g(g_cp_mu);
            if (!cp->has_ref) {
                SimSetup tmp;
                tmp.cfg = c;
                if ((rc = ref_raster(c, tmp))) return rc;
                cp->ref = tmp.ref;
                cp->span = tmp.span;
                cp->jp_len = tmp.jp_len;
                cp->has_ref = true;
            }
            s.ref = cp->ref;
            s.span = cp->span;
            s.jp_len = cp->jp_len;
        }
        ref_cpml(c, s);
        const float chez = (float)((dt / c.dx) / kMU0);
        s.chx = s.chy = chez;
        s.chx_d = (float)((double)chez * c.dy);
        s.chy_d = (float)((double)chez * c.dx);
        // incident plane wave, fwf (FDTD_GPU.cu:283-296); CPU variant FDTD_CPU.cxx:211-225
        const float Phi = (float)(exc * 2 * kPI / c.n_exc);
        float ky, kx;
        if (c.flags & TOMO_FLAG_FWF_ABS_ORIGIN) { ky = (float)std::sin((double)Phi); kx = (float)std::cos((double)Phi); }
        else { ky = std::sin(Phi); kx = std::cos(Phi); }
        const double l = (double)(float)(-c.breast_radius);
        const double ic0 = (double)(1 / (float)c0());
        s.inc_dt = (float)dt;
        s.inc_c = (float)(-(double)(float)c.pulse_t0 + ic0 * l);
        s.inc_ay = (float)(ic0 * ky * c.dx);
        s.inc_ax = (float)(ic0 * kx * c.dy);
        s.inc_itau = (float)(1.0 / c.pulse_tau);
        s.inc_abs = (c.flags & TOMO_FLAG_FWF_ABS_ORIGIN) ? 1 : 0;
        s.tJ = 0.0;
        s.tM = 0.5;
        for (int f = 0; f < c.n_freq; ++f) {
            const double fr = s.freq[f], tau = c.pulse_tau;
            const float Prad = (float)((0.5 / eta0()) *
                                       std::pow(tau * std::sqrt(kPI) * std::exp(-tau * tau * 2 * kPI * fr * 2 * kPI * fr / 4), 2));
            s.inv_norm[f] = 1.0 / (8 * kPI * eta0() * (double)Prad * c.ff_norm);
        }
    } else {
        if ((rc = build_phantom(c, p, np, s))) return rc;
        if ((int)s.tissues.size() > TOMO_MAX_MAT) return fail(TOMO_E_INVALID, "too many materials");
        phys_coefs(c, s);
        physical_spans(c, s);
        phys_cpml(c, s);
        s.chy = (float)(dt / (kMU0 * c.dx));
        s.chx = (float)(dt / (kMU0 * c.dy));
        s.chy_d = (float)((double)s.chy * c.dx);
        s.chx_d = (float)((double)s.chx * c.dy);
        const double th = 2.0 * kPI * exc / c.n_exc;
        s.sx = c.nx / 2 + (int)std::lround(c.src_radius * std::cos(th) / c.dx);
        s.sy = c.ny / 2 + (int)std::lround(c.src_radius * std::sin(th) / c.dy);
        if (s.sx < 1 || s.sx >= c.nx || s.sy < 1 || s.sy >= c.ny) return fail(TOMO_E_INVALID, "source outside grid");
        auto cp = config_products(c);
        std::lock_guard<std::mutex> g(g_cp_mu);
        if (!cp->has_phys) {
            cp->srcw.resize(c.steps);
            for (int n = 0; n < c.steps; ++n) cp->srcw[n] = (float)phys_source_value(c, n);
            // S(f) = sum_n s(n) dt exp(-j w (n+1) dt), in FP64 with an exact reseed every 256 steps
            for (int f = 0; f < c.n_freq; ++f) {
                const double w = 2.0 * kPI * s.freq[f];
                const std::complex<double> rot = std::exp(std::complex<double>(0.0, -w * dt));
                std::complex<double> S = 0, z = 0;
                for (int n = 0; n < c.steps; ++n) {
                    if ((n & 255) == 0) z = std::exp(std::complex<double>(0.0, -w * (n + 1) * dt));
                    S += phys_source_value(c, n) * dt * z;
                    z *= rot;
                }
                cp->inv_norm[f] = 1.0 / (8 * kPI * eta0() * std::norm(S));
            }
            cp->has_phys = true;
        }
        s.srcw = cp->srcw;
        for (int f = 0; f < c.n_freq; ++f) s.inv_norm[f] = cp->inv_norm[f];
        s.tJ = 0.5;
        s.tM = 1.0;
    }
    return TOMO_OK;
}

}  // namespace tomo
