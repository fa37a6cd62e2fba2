// fdtd_oracle.cpp — CPU restatement of the reference's 2D TMz FDTD fitness path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
// to check results (or as the timed CPU baseline), never as the product.
//
// Two kinds of arithmetic, chosen by the template parameter R:
//   R = float  : restates the reference's own FP32 storage with the FP64
//                promotions its C++ expressions imply (double macros such as
//                dt, dx, Kp, eps0), so that with the CPU-variant flags the
//                fields agree bit-for-bit with the compiled FDTD_CPU.cxx;
//   R = double : the FP64 oracle used to state every FP32 result's error.
//
// Reference functions restated (all citations are to the reference tree):
//   parameter unpack          FDTD_GPU.cu:655-693,   FDTD_CPU.cxx:681-698
//   getOptimizationCell       FDTD_GPU.cu:544-551,   FDTD_CPU.cxx:530-537
//   disc rasterisation/coefs  FDTD_GPU.cu:718-743,   FDTD_CPU.cxx:936-971,1078-1164
//   scattered_parameter_init  FDTD_GPU.cu:1112-1123, FDTD_CPU.cxx:893-905
//   CPML profiles             FDTD_GPU.cu:748-806,   FDTD_CPU.cxx:1328-1479
//   fwf incident wave         FDTD_GPU.cu:283-296,   FDTD_CPU.cxx:211-225
//   H_field_update            FDTD_GPU.cu:298-355,   FDTD_CPU.cxx:227-285
//   E_field_update (+Debye)   FDTD_GPU.cu:358-434,   FDTD_CPU.cxx:288-373
//   contour index map         FDTD_common.cxx:4-46
//   calculate_JandM           FDTD_GPU.cu:231-280,   FDTD_CPU.cxx:146-197
//   calc_incident_power       FDTD_common.cxx:48-51
//   N2FPostProcess            FDTD_GPU.cu:553-634,   FDTD_CPU.cxx:538-614
//   fitness                   FDTD_GPU.cu:1004-1018, FDTD_CPU.cxx:841-845,
//                             FDTD_common.cxx:53-61 (TOMO_FLAG_FITNESS_PER_EXC)
// Deviation from FDTD_CPU.cxx: the far-field loop at FDTD_CPU.cxx:816 reads an
// uninitialised freq_index (so D stays zero there); the restatement starts
// that loop at 0, which is what FDTD_GPU.cu:971 does.
//
// Physical mode (the benchmark workloads) has no reference counterpart; it
// keeps the same update structure (Yee leapfrog, Roden-Gedney CPML, the same
// two-level Debye polarisation current, the same running DFT and projection)
// with the reference's non-physical quirks removed.
//
// How each part is pinned (tests/test_oracle.py):
//   CPU-path fields, CPML tables, contour, materials, running DFT:
//       bit-exact against fixtures from the compiled FDTD_CPU.cxx;
//   GPU-path E update: the delta_eps = 0 special case reduces it to the CPU
//       fixture bit for bit; a vacuum phantom gives D == 0;
//   calc_incident_power: numerical transform of the Gaussian pulse;
//   physical Debye coefficients: an independent 2x2 solve of the discretised
//       PDEs; physical CPML: reflection against a larger grid; physical running
//       DFT, closed contour and projection: point-source isotropy and a direct
//       transform of the samples.
//   GPU-path Debye current with delta_eps > 0 and the centred fwf origin: a
//       homogeneous Debye disc lit by the plane wave matches its plane-wave
//       series (Bessel/Hankel) within 1.5%, as ratios between probes and
//       absolutely against the incident pulse's spectrum.
//   Phase output P of N2FPostProcess: the same disc's complex far field
//       sqrt(D) e^{jP} matches its series far field (with the reference's
//       phase centre, half a cell off the disc centre for odd nx) within 1.5%.
//   Fitness: fitness() of FDTD_common.cxx:53-61 against the compiled function
//       on random arrays; FDTD_CPU.cxx:841-845 against the value the whole
//       compiled FDTD_CPU() returns (D = 0, its far-field loop never runs);
//       FDTD_GPU.cu:1004-1008 reduced to the compiled fitness() at m = 1,
//       which leaves its 1/n_freq factor pinned by the source alone
//       (FDTD_GPU.cu cannot be compiled here); the phase term of
//       FDTD_GPU.cu:1009-1018 by its special cases.

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/tomo.h"

namespace orc {

// Physical constants exactly as FDTD_common.hxx:3-12 spells them.
static const double kPI = 3.141592653589793238;
static const double kEPS0 = 8.85418e-12;
static const double kMU0 = kPI * 4e-7;
static inline double c0() { return 1.0 / std::sqrt(kMU0 * kEPS0); }
static inline double eta0() { return std::sqrt(kMU0 / kEPS0); }

static thread_local std::string g_err;

// ---------------------------------------------------------------- contour
// FDTD_common.cxx:4-46: contour index -> (x, y) over four legs
// yn (y = D, x ascending), xp (x = nx-D-1, y ascending),
// yp (y = ny-D-1, x descending), xn (x = D, y descending).
struct Contour {
    int P = 0, L1 = 0, L2 = 0, L3 = 0;
    std::vector<int> x, y, leg;  // leg: 0 yn, 1 xp, 2 yp, 3 xn
};

// closed = false: the reference's map, whose x faces start one cell below the
// bottom face (y = D) and whose y faces stop one cell short of the right face
// (x = nx-D-2).  closed = true (physical mode): the same number of points on
// a closed box: y faces x in [D+1, nx-D-1], x faces y in [D+1, ny-D-1].
static Contour make_contour(int nx, int ny, int D, bool closed = false) {
    Contour c;
    const int xhi = closed ? nx - D - 1 : nx - D - 2;  // last x of the y faces
    const int ylo = closed ? D + 1 : D;                 // first y of the x faces
    c.L1 = xhi - D;
    c.L2 = c.L1 + (ny - D - ylo);
    c.L3 = c.L2 + (xhi - D);
    c.P = c.L3 + (ny - D - ylo);
    for (int i = 0; i < c.P; ++i) {
        int x, y, g;
        if (i < c.L1) {
            x = i + D + 1; y = D; g = 0;
        } else if (i < c.L2) {
            x = nx - D - 1; y = (i - c.L1) + ylo; g = 1;
        } else if (i < c.L3) {
            x = xhi - (i - c.L2); y = ny - D - 1; g = 2;
        } else {
            x = D; y = ny - D - (i - c.L3) - 1; g = 3;
        }
        c.x.push_back(x); c.y.push_back(y); c.leg.push_back(g);
    }
    return c;
}

// ---------------------------------------------------------------- phantom (physical mode)
// Seeded generator shared in specification with the product's host setup:
// splitmix64 stream, a skin ring, a Debye fat disc, three fibroglandular
// ellipses and up to four tumour discs taken from the parameter vector.
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

struct Tissue { double einf, deps, tau, sigma; };
struct Circle { float cx, cy, r2; };
struct Ellipse { float cx, cy, cs, sn, ia2, ib2; };

struct PhysPhantom {
    float ro2, ri2;          // outer / inner radius squared
    Ellipse ell[3];
    int ntum = 0;
    Circle tum[4];
    std::vector<Tissue> tissue;  // material id -> tissue
};

static const Tissue kVacuum{1.0, 0.0, 0.0, 0.0};
static const Tissue kSkin{36.0, 0.0, 0.0, 4.0};
static const Tissue kFat{3.1, 1.6, 13e-12, 0.0};
static const Tissue kFibro{12.8, 24.6, 13e-12, 0.0};
static const Tissue kTumourDefault{20.0, 34.0, 13e-12, 0.0};

static bool build_phys_phantom(const tomo_sim_config& c, const double* p, size_t np, PhysPhantom& ph) {
    const double Ro = c.breast_radius, Ri = c.breast_radius - c.skin_thickness;
    ph.ro2 = (float)(Ro * Ro);
    ph.ri2 = (float)(Ri * Ri);
    uint64_t off = 0;
    if (np > 7 && p[7] > 0) off = (uint64_t)std::floor(p[7]);
    SplitMix rng((c.seed ^ 0xE111B5E5ull) + off);
    for (int e = 0; e < 3; ++e) {
        double rho = (Ri - 0.012) * std::sqrt(rng.u01());
        double phi = 2.0 * kPI * rng.u01();
        double a = 0.004 + 0.008 * rng.u01();
        double b = 0.004 + 0.008 * rng.u01();
        double th = kPI * rng.u01();
        Ellipse& E = ph.ell[e];
        E.cx = (float)(rho * std::cos(phi));
        E.cy = (float)(rho * std::sin(phi));
        E.cs = (float)std::cos(th);
        E.sn = (float)std::sin(th);
        E.ia2 = (float)(1.0 / (a * a));
        E.ib2 = (float)(1.0 / (b * b));
    }
    ph.tissue = {kVacuum, kSkin, kFat, kFibro};
    ph.ntum = 0;
    for (int k = 0; k < 4; ++k) {
        size_t b = 8 * (size_t)k;
        if (np <= b) break;
        auto get = [&](size_t i, double dflt) { return (b + i < np) ? p[b + i] : dflt; };
        double r_mm = get(2, 4.0);
        if (!(r_mm > 0.0)) continue;  // radius <= 0: no tumour in this block
        if (r_mm < 2.0) r_mm = 2.0;
        if (r_mm > 6.0) r_mm = 6.0;
        Tissue t{get(3, kTumourDefault.einf), get(4, kTumourDefault.deps),
                 get(5, kTumourDefault.tau * 1e12) * 1e-12, get(6, kTumourDefault.sigma)};
        if (!(t.einf >= 1.0) || !(t.deps >= 0.0) || !(t.tau > 0.0) || !(t.sigma >= 0.0) ||
            !std::isfinite(get(0, 0.0)) || !std::isfinite(get(1, 0.0))) {
            g_err = "invalid tumour parameters";
            return false;
        }
        double r = r_mm * 1e-3;
        Circle& C = ph.tum[ph.ntum];
        C.cx = (float)(get(0, 0.0) * 1e-3);
        C.cy = (float)(get(1, 0.0) * 1e-3);
        C.r2 = (float)(r * r);
        ph.tissue.push_back(t);
        ph.ntum++;
    }
    return true;
}

// Strictly ordered FP32 arithmetic (the file is built with -ffp-contract=off).
static inline bool in_circle(float X, float Y, float cx, float cy, float r2) {
    volatile float dX = X - cx, dY = Y - cy;
    volatile float a = dX * dX, b = dY * dY;
    volatile float s = a + b;
    return s <= r2;
}
static inline bool in_ellipse(float X, float Y, const Ellipse& E) {
    volatile float dX = X - E.cx, dY = Y - E.cy;
    volatile float u1 = dX * E.cs, u2 = dY * E.sn;
    volatile float u = u1 + u2;
    volatile float v1 = dY * E.cs, v2 = dX * E.sn;
    volatile float v = v1 - v2;
    volatile float uu = u * u, vv = v * v;
    volatile float q1 = uu * E.ia2, q2 = vv * E.ib2;
    volatile float q = q1 + q2;
    return q <= 1.0f;
}

static uint8_t phys_material(const PhysPhantom& ph, int x, int y, int nx, int ny, double dx, double dy) {
    volatile float Xi = (float)(x - nx / 2), Yi = (float)(y - ny / 2);
    volatile float X = Xi * (float)dx, Y = Yi * (float)dy;
    uint8_t m = 0;
    if (!in_circle(X, Y, 0.f, 0.f, ph.ro2)) return m;
    m = 1;
    if (!in_circle(X, Y, 0.f, 0.f, ph.ri2)) return m;
    m = 2;
    for (int e = 0; e < 3; ++e)
        if (in_ellipse(X, Y, ph.ell[e])) m = 3;
    for (int k = 0; k < ph.ntum; ++k)
        if (in_circle(X, Y, ph.tum[k].cx, ph.tum[k].cy, ph.tum[k].r2)) m = (uint8_t)(4 + k);
    return m;
}

// Target vector of the seeded target phantom (physical mode).
static std::vector<double> target_vector(const tomo_sim_config& c) {
    SplitMix rng(c.seed ^ 0x7A26E7ull);
    std::vector<double> v(8, 0.0);
    v[0] = -15.0 + 30.0 * rng.u01();
    v[1] = -15.0 + 30.0 * rng.u01();
    v[2] = 2.0 + 4.0 * rng.u01();
    v[3] = kTumourDefault.einf;
    v[4] = kTumourDefault.deps;
    v[5] = kTumourDefault.tau * 1e12;
    v[6] = 0.0;
    v[7] = 0.0;
    return v;
}

// ---------------------------------------------------------------- simulation
template <class R>
struct Sim {
    tomo_sim_config c{};
    int nx = 0, ny = 0, nc = 0, exc = 0;
    bool ref = true;
    Contour ct;
    // material map and per-cell coefficients, (nx+1)*(ny+1), index x + y*(nx+1)
    std::vector<uint8_t> mat;
    std::vector<R> Ceze, Cezhy, Cezhx, Cezeic, Cezeip, Cezjp, Beta_p, BpDt, Kpc;
    bool dispersive_any = false;
    // CPML 1-D profiles (ncells entries each)
    std::vector<R> kex, kmx, aex, bex, amx, bmx;
    // fields
    std::vector<R> Ez, Jp, Hx, Hy;                 // Ez, Jp: (nx+1)(ny+1); H: nx*ny
    std::vector<R> psi_ezx, psi_hyx, psi_ezy, psi_hxy;  // [row][2nc] / [2nc][col]
    // running DFT: [freq][point]
    std::vector<std::complex<R>> J, M;
    std::vector<double> freq;
    // physical source
    int sx = -1, sy = -1;
    std::vector<R> srcw;
    int n = 0;  // steps done
    // incident wave (ref)
    float Phi = 0.f;

    int E(int x, int y) const { return x + y * (nx + 1); }
    int H(int x, int y) const { return x + y * nx; }
};

// CPML profiles exactly as FDTD_GPU.cu:752-806 / FDTD_CPU.cxx:1328-1479,
// including their float/double mix when R = float.
template <class R>
static void ref_cpml(Sim<R>& s) {
    const tomo_sim_config& c = s.c;
    const int nc = s.nc;
    const double dx = c.dx, dt = c.dt, npml = c.cpml_order, kmax = c.kappa_max;
    const double alpha_max = c.alpha_max, alpha_min = 0.0, sigma_factor = 1.0;
    s.kex.assign(nc, 0); s.kmx.assign(nc, 0); s.aex.assign(nc, 0);
    s.bex.assign(nc, 0); s.amx.assign(nc, 0); s.bmx.assign(nc, 0);
    R sigma_max = (R)c.sigma_max;
    for (int i = 0; i < nc; ++i) {
        R rho = (R)(((R)i + 0.25) / nc);
        R sig_e = (R)(sigma_max * sigma_factor * std::pow((double)rho, npml));
        R ke = (R)(std::pow((double)rho, npml) * (kmax - 1) + 1);
        R al_e = (R)(alpha_min + (alpha_max - alpha_min) * rho);
        R rhom = (R)(((R)i + 0.75) / nc);
        R sig_m = (R)((kMU0 / kEPS0) * sigma_max * sigma_factor * std::pow((double)rhom, npml));
        R km = (R)(std::pow((double)rhom, npml) * (kmax - 1) + 1);
        R al_m = (R)((kMU0 / kEPS0) * (alpha_min + (alpha_max - alpha_min) * rhom));
        R be = (R)std::exp(-1 * (dt / kEPS0) * (double)(sig_e / ke + al_e));
        R ae = (R)((double)((be - 1) * sig_e) / (dx * (double)(sig_e * ke + al_e * ke * ke)));
        R arg = (R)(-1 * (dt / kMU0) * (double)((sig_m / km) + al_m));
        R bm = (R)std::exp(arg);  // std::exp(float) when R = float, as the reference
        R am = (R)((double)((bm - 1) * sig_m) / (dx * (double)(sig_m * km + al_m * km * km)));
        s.kex[i] = ke; s.kmx[i] = km; s.aex[i] = ae; s.bex[i] = be; s.amx[i] = am; s.bmx[i] = bm;
    }
}

// Physical CPML: polynomial grading of order m, sigma_max from R0 (stored in
// cfg.sigma_max by the preset), kappa = 1 + (kmax-1) rho^m, alpha = amax (1-rho).
template <class R>
static void phys_cpml(Sim<R>& s) {
    const tomo_sim_config& c = s.c;
    const int nc = s.nc;
    s.kex.assign(nc, 0); s.kmx.assign(nc, 0); s.aex.assign(nc, 0);
    s.bex.assign(nc, 0); s.amx.assign(nc, 0); s.bmx.assign(nc, 0);
    for (int i = 0; i < nc; ++i) {
        for (int h = 0; h < 2; ++h) {
            double rho = ((double)i + (h ? 0.75 : 0.25)) / nc;
            double g = std::pow(rho, c.cpml_order);
            double sig = c.sigma_max * g;
            double kap = 1.0 + (c.kappa_max - 1.0) * g;
            double al = c.alpha_max * (1.0 - rho);
            double b = std::exp(-(sig / kap + al) * c.dt / kEPS0);
            double a = (sig == 0.0 && al == 0.0) ? 0.0 : sig * (b - 1.0) / (c.dx * (sig * kap + al * kap * kap));
            if (h) { s.kmx[i] = (R)kap; s.bmx[i] = (R)b; s.amx[i] = (R)a; }
            else { s.kex[i] = (R)kap; s.bex[i] = (R)b; s.aex[i] = (R)a; }
        }
    }
}

// Reference disc rasterisation + coefficient precompute.
// GPU form: FDTD_GPU.cu:718-743 and 1112-1123; CPU form: FDTD_CPU.cxx:936-971,1078-1164,893-905.
template <class R>
static bool ref_materials(Sim<R>& s, const double* p, size_t np) {
    const tomo_sim_config& c = s.c;
    const int nx = s.nx, ny = s.ny, npch = c.numpatches;
    const bool cpu = (c.flags & TOMO_FLAG_CPU_COEFS) != 0;
    const bool eps_sigma = (c.flags & TOMO_FLAG_PARAMS_EPS_SIGMA) != 0;
    const size_t need = (size_t)npch * npch * (eps_sigma ? 2 : 3);
    if (np < need) { g_err = "parameter vector too short"; return false; }
    std::vector<float> image(need);
    for (size_t k = 0; k < need; ++k) image[k] = (float)p[k];
    const double dx = c.dx, dy = c.dy, dt = c.dt, R_ = c.breast_radius, tau_r = c.debye_tau;
    const double Kp = (1 - dt / (2 * tau_r)) / (1 + dt / (2 * tau_r));
    const int N1 = (nx + 1) * (ny + 1);
    s.mat.assign(N1, 0);
    for (auto* v : {&s.Ceze, &s.Cezhy, &s.Cezhx, &s.Cezeic, &s.Cezeip, &s.Cezjp, &s.Beta_p, &s.BpDt, &s.Kpc})
        v->assign(N1, 0);
    // per material id (0 = vacuum, 1 + patch) coefficient table, computed once
    const int nm = 1 + npch * npch;
    for (int j = 0; j < ny + 1; ++j) {
        for (int i = 0; i < nx + 1; ++i) {
            float radius = (float)std::sqrt(std::pow(((float)i - nx / 2) * dx, 2) +
                                            std::pow(((float)j - ny / 2) * dy, 2));
            int m = 0;
            if (radius <= R_) {
                int xc = (int)((i - (nx / 2 - (int)(R_ / dx))) / (2 * R_ / (npch * dx)));
                int yc = (int)((j - (ny / 2 - R_ / dy)) / (2 * R_ / (npch * dy)));
                int patch = xc + npch * yc;
                if (patch < 0 || patch >= npch * npch) { g_err = "patch index out of range"; return false; }
                m = 1 + patch;
            }
            s.mat[s.E(i, j)] = (uint8_t)m;
        }
    }
    (void)nm;
    for (int idx = 0; idx < N1; ++idx) {
        int m = s.mat[idx];
        if (!cpu) {
            float eps_inf = 1, del_eps = 0, sigma = 0;
            if (m) {
                eps_inf = image[m - 1];
                del_eps = image[m - 1 + npch * npch];
                sigma = image[m - 1 + 2 * npch * npch];
            }
            R Bp = (R)(kEPS0 * del_eps * dt / (tau_r * (1 + dt / (2 * tau_r))));
            double den_h = 2 * eps_inf * kEPS0 + sigma * dt + (double)Bp;
            s.Beta_p[idx] = Bp;
            s.Cezhy[idx] = (R)((2 * dt / dx) / den_h);
            s.Cezhx[idx] = s.Cezhy[idx];  // E_field_update is launched with Cezhy twice (FDTD_GPU.cu:947)
            s.Ceze[idx] = (R)((2 * eps_inf * kEPS0 - sigma * dt + (double)Bp) / den_h);
            double den_d = 2 * kEPS0 * eps_inf + sigma * dt + (double)Bp;
            s.Cezeic[idx] = (R)((-2 * kEPS0 * eps_inf + 2 * kEPS0 - sigma * dt - (double)Bp) / den_d);
            s.Cezeip[idx] = (R)((2 * kEPS0 * eps_inf - 2 * kEPS0 - sigma * dt + (double)Bp) / den_d);
            s.Cezjp[idx] = (R)(((1 + Kp) * dt) / den_d);
            s.BpDt[idx] = (R)((double)Bp / dt);
            s.Kpc[idx] = (R)Kp;
            if (del_eps > 0) s.dispersive_any = true;
        } else {
            float eps_r = 1, sigma = 0;
            if (m) {
                eps_r = image[m - 1];
                sigma = image[m - 1 + npch * npch];
            }
            s.Ceze[idx] = (R)((2 * eps_r * kEPS0 - dt * sigma) / (2 * eps_r * kEPS0 + dt * sigma));
            s.Cezhy[idx] = (R)((2 * dt / dx) / (2 * eps_r * kEPS0 + dt * sigma));
            s.Cezhx[idx] = s.Cezhy[idx];  // FDTD_CPU.cxx:807 passes Cezhy for Cezhx
            s.Cezeic[idx] = (R)((2 * (kEPS0 - kEPS0 * eps_r) - sigma * dt) / (2 * kEPS0 * eps_r + sigma * dt));
            s.Cezeip[idx] = (R)(-1 * (2 * (kEPS0 - kEPS0 * eps_r) + sigma * dt) / (2 * kEPS0 * eps_r + sigma * dt));
        }
    }
    return true;
}

template <class R>
static bool phys_materials(Sim<R>& s, const double* p, size_t np) {
    const tomo_sim_config& c = s.c;
    if (np < 8) { g_err = "parameter vector too short"; return false; }
    PhysPhantom ph;
    if (!build_phys_phantom(c, p, np, ph)) return false;
    const int N1 = (s.nx + 1) * (s.ny + 1);
    s.mat.assign(N1, 0);
    for (auto* v : {&s.Ceze, &s.Cezhy, &s.Cezhx, &s.Cezeic, &s.Cezeip, &s.Cezjp, &s.Beta_p, &s.BpDt, &s.Kpc})
        v->assign(N1, 0);
    for (int y = 0; y <= s.ny; ++y)
        for (int x = 0; x <= s.nx; ++x) s.mat[s.E(x, y)] = phys_material(ph, x, y, s.nx, s.ny, c.dx, c.dy);
    const double dt = c.dt;
    for (int idx = 0; idx < N1; ++idx) {
        const Tissue& t = ph.tissue[s.mat[idx]];
        double Bp = 0, Kp = 0;
        if (t.deps > 0 && t.tau > 0) {
            Bp = kEPS0 * t.deps * dt / (t.tau * (1 + dt / (2 * t.tau)));
            Kp = (1 - dt / (2 * t.tau)) / (1 + dt / (2 * t.tau));
            s.dispersive_any = true;
        }
        double den = 2 * t.einf * kEPS0 + t.sigma * dt + Bp;
        s.Ceze[idx] = (R)((2 * t.einf * kEPS0 - t.sigma * dt + Bp) / den);
        s.Cezhy[idx] = (R)(2 * dt / (c.dx * den));
        s.Cezhx[idx] = (R)(2 * dt / (c.dy * den));
        s.Cezjp[idx] = (R)((1 + Kp) * dt / den);
        s.Beta_p[idx] = (R)Bp;
        s.BpDt[idx] = (R)(Bp / dt);
        s.Kpc[idx] = (R)Kp;
    }
    return true;
}

// Physical source: soft Ez point source on an antenna ring, differentiated
// Gaussian s(t) = -((t-t0)/tau) exp(-((t-t0)/tau)^2), t = (n+1) dt.
static void phys_source_cell(const tomo_sim_config& c, int exc, int& sx, int& sy) {
    double th = 2.0 * kPI * exc / (c.n_exc > 0 ? c.n_exc : 1);
    sx = c.nx / 2 + (int)std::lround(c.src_radius * std::cos(th) / c.dx);
    sy = c.ny / 2 + (int)std::lround(c.src_radius * std::sin(th) / c.dy);
}
static double phys_source_value(const tomo_sim_config& c, int n) {
    double t = (n + 1) * c.dt;
    double u = (t - c.pulse_t0) / c.pulse_tau;
    return -u * std::exp(-u * u);
}

template <class R>
static void reset_fields(Sim<R>& s) {
    const int N1 = (s.nx + 1) * (s.ny + 1), NH = s.nx * s.ny;
    s.Ez.assign(N1, 0); s.Jp.assign(N1, 0); s.Hx.assign(NH, 0); s.Hy.assign(NH, 0);
    s.psi_ezx.assign((size_t)(s.ny + 1) * 2 * s.nc, 0);
    s.psi_hyx.assign((size_t)(s.ny + 1) * 2 * s.nc, 0);
    s.psi_ezy.assign((size_t)(s.nx + 1) * 2 * s.nc, 0);
    s.psi_hxy.assign((size_t)(s.nx + 1) * 2 * s.nc, 0);
    s.J.assign((size_t)s.c.n_freq * s.ct.P, 0);
    s.M.assign((size_t)s.c.n_freq * s.ct.P, 0);
    s.n = 0;
}

template <class R>
static bool sim_init(Sim<R>& s, const tomo_sim_config& c, const double* p, size_t np, int exc) {
    s.c = c;
    s.nx = c.nx; s.ny = c.ny; s.nc = c.ncpml; s.exc = exc;
    s.ref = (c.mode == TOMO_MODE_REF_COMPAT);
    if (c.nx < 8 || c.ny < 8 || c.ncpml < 1 || 2 * c.ncpml + 2 > c.nx || c.n_freq < 1 || c.n_angles < 1 ||
        c.ntff_offset < 1 || c.nx - 2 * c.ntff_offset - 2 < 1 || exc < 0 || exc >= c.n_exc) {
        g_err = "invalid configuration";
        return false;
    }
    if (s.ref && c.nx != c.ny) { g_err = "ref-compat mode requires nx == ny"; return false; }
    s.ct = make_contour(c.nx, c.ny, c.ntff_offset, !s.ref);
    s.freq.resize(c.n_freq);
    for (int f = 0; f < c.n_freq; ++f) {
        double df = c.n_freq > 1 ? (c.f_stop - c.f_start) / (c.n_freq - 1) : 0.0;
        s.freq[f] = c.f_start + df * f;
        if (s.ref) s.freq[f] = (double)(float)s.freq[f];  // `float freq` (FDTD_GPU.cu:698, 951)
    }
    if (s.ref) {
        if (!ref_materials(s, p, np)) return false;
        ref_cpml(s);
        s.Phi = (float)(exc * 2 * kPI / c.n_exc);
    } else {
        if (!phys_materials(s, p, np)) return false;
        phys_cpml(s);
        phys_source_cell(c, exc, s.sx, s.sy);
        s.srcw.resize(c.steps);
        for (int n = 0; n < c.steps; ++n) s.srcw[n] = (R)phys_source_value(c, n);
    }
    reset_fields(s);
    return true;
}

// fwf: FDTD_GPU.cu:283-296 (sincosf, centred origin) / FDTD_CPU.cxx:211-225
// (double sin/cos, absolute origin).  The exponent is evaluated in double.
template <class R>
static inline R fwf(const Sim<R>& s, float timestep, float x, float y) {
    const tomo_sim_config& c = s.c;
    const bool cpu = (c.flags & TOMO_FLAG_FWF_ABS_ORIGIN) != 0;
    const double dt = c.dt, dx = c.dx, dy = c.dy, tau = c.pulse_tau;
    if (sizeof(R) == 4) {
        const float l = (float)(-c.breast_radius);  // fwf's `float l` parameter
        float ky, kx;
        if (cpu) { ky = (float)std::sin((double)s.Phi); kx = (float)std::cos((double)s.Phi); }
        else { ky = std::sin(s.Phi); kx = std::cos(s.Phi); }
        float ar = (float)((float)timestep * dt - (float)c.pulse_t0 -
                           (1 / (float)c0()) * (ky * y * dx + kx * x * dy - l));
        return (R)std::exp(-1 * ar * ar / (tau * tau));
    } else {
        const double l = -c.breast_radius;
        double ky = std::sin((double)s.Phi), kx = std::cos((double)s.Phi);
        double ar = (double)timestep * dt - c.pulse_t0 - (1 / c0()) * (ky * y * dx + kx * x * dy - l);
        return (R)std::exp(-ar * ar / (tau * tau));
    }
}

template <class R>
static inline bool isnan_(R v) { return v != v; }

// H_field_update: FDTD_GPU.cu:298-355 / FDTD_CPU.cxx:227-285.
template <class R>
static void ref_H(Sim<R>& s) {
    const int nx = s.nx, ny = s.ny, nc = s.nc;
    const double dx = s.c.dx, dy = s.c.dy, dt = s.c.dt;
    const R Chez = (R)((dt / dx) / kMU0);
    const bool clamp = !(s.c.flags & TOMO_FLAG_NO_NAN_CLAMP);
    for (int y = 0; y < ny; ++y) {
        for (int x = 0; x < nx; ++x) {
            const R ez = s.Ez[s.E(x, y)];
            const R dEx = s.Ez[s.E(x + 1, y)] - ez;
            const R dEy = s.Ez[s.E(x, y + 1)] - ez;
            const R hy0 = s.Hy[s.H(x, y)], hx0 = s.Hx[s.H(x, y)];
            R bHy = hy0 + Chez * dEx;
            R bHx = hx0 - Chez * dEy;
            if (x < nc) {
                int i = nc - 1 - x;
                bHy = hy0 + Chez * dEx / s.kmx[i];
                R& ps = s.psi_hyx[(size_t)y * 2 * nc + x];
                ps = s.bmx[i] * ps + s.amx[i] * dEx;
                bHy = (R)((double)bHy + (double)Chez * dx * (double)ps);
            }
            if (x >= nx - nc) {
                int i = x - nx + nc;
                bHy = hy0 + Chez * dEx / s.kmx[i];
                R& ps = s.psi_hyx[(size_t)y * 2 * nc + nc + i];
                ps = s.bmx[i] * ps + s.amx[i] * dEx;
                bHy = (R)((double)bHy + (double)Chez * dx * (double)ps);
            }
            if (y < nc) {
                int i = nc - 1 - y;
                bHx = hx0 - Chez * dEy / s.kmx[i];
                R& ps = s.psi_hxy[(size_t)y * (nx + 1) + x];
                ps = s.bmx[i] * ps + s.amx[i] * dEy;
                bHx = (R)((double)bHx - (double)Chez * dy * (double)ps);
            }
            if (y >= ny - nc) {
                int i = y - ny + nc;
                bHx = hx0 - Chez * dEy / s.kmx[i];
                R& ps = s.psi_hxy[(size_t)(nc + i) * (nx + 1) + x];
                ps = s.bmx[i] * ps + s.amx[i] * dEy;
                bHx = (R)((double)bHx - (double)Chez * dy * (double)ps);
            }
            if (clamp && isnan_(bHx)) bHx = 0;
            if (clamp && isnan_(bHy)) bHy = 0;
            s.Hx[s.H(x, y)] = bHx;
            s.Hy[s.H(x, y)] = bHy;
        }
    }
}

// E_field_update: FDTD_GPU.cu:358-434 (Debye Jp) / FDTD_CPU.cxx:288-373.
template <class R>
static void ref_E(Sim<R>& s, int it) {
    const tomo_sim_config& c = s.c;
    const int nx = s.nx, ny = s.ny, nc = s.nc;
    const double dx = c.dx, dy = c.dy, dt = c.dt;
    const double Kp = (1 - dt / (2 * c.debye_tau)) / (1 + dt / (2 * c.debye_tau));
    const bool cpu_coefs = (c.flags & TOMO_FLAG_CPU_COEFS) != 0;
    const bool abs_origin = (c.flags & TOMO_FLAG_FWF_ABS_ORIGIN) != 0;
    const bool clamp = !(c.flags & TOMO_FLAG_NO_NAN_CLAMP);
    for (int y = 0; y <= ny; ++y) {
        for (int x = 0; x <= nx; ++x) {
            const int e = s.E(x, y);
            R buf;
            if (x == nx || y == ny || x == 0 || y == 0) {
                buf = 0;
            } else {
                const R ez = s.Ez[e];
                const R dHy = s.Hy[s.H(x, y)] - s.Hy[s.H(x - 1, y)];
                const R dHx = s.Hx[s.H(x, y)] - s.Hx[s.H(x, y - 1)];
                float fx = abs_origin ? (float)x : (float)(x - nx / 2);
                float fy = abs_origin ? (float)y : (float)(y - ny / 2);
                R f1 = 0, f0 = 0;
                if (s.Cezeic[e] != 0 || s.Cezeip[e] != 0 || s.Beta_p[e] != 0) {
                    f1 = fwf(s, (float)((float)it + 0.5), fx, fy);
                    f0 = fwf(s, (float)((float)it - 0.5), fx, fy);
                }
                if (cpu_coefs) {
                    buf = s.Ceze[e] * ez + s.Cezhy[e] * dHy - s.Cezhx[e] * dHx + s.Cezeic[e] * f1 + s.Cezeip[e] * f0;
                } else {
                    buf = s.Ceze[e] * ez + s.Cezhy[e] * dHy - s.Cezhx[e] * dHx + s.Cezeic[e] * f1 +
                          s.Cezeip[e] * f0 - s.Cezjp[e] * s.Jp[e];
                    s.Jp[e] = (R)(Kp * (double)s.Jp[e] +
                                  ((double)s.Beta_p[e] / dt) * (double)(buf - ez + f1 - f0));
                }
                if (x <= nc && x != 0) {
                    int i = nc - x;
                    buf = s.Ceze[e] * ez + s.Cezhy[e] * dHy / s.kex[i] - s.Cezhx[e] * dHx / s.kex[i];
                    R& ps = s.psi_ezx[(size_t)y * 2 * nc + (x - 1)];
                    ps = s.bex[i] * ps + s.aex[i] * dHy;
                    buf = (R)((double)buf + (double)s.Cezhy[e] * dx * (double)ps);
                }
                if (x >= nx - nc && x != nx) {
                    int i = x - nx + nc;
                    buf = s.Ceze[e] * ez + s.Cezhy[e] * dHy / s.kex[i] - s.Cezhx[e] * dHx / s.kex[i];
                    R& ps = s.psi_ezx[(size_t)y * 2 * nc + nc + i];
                    ps = s.bex[i] * ps + s.aex[i] * dHy;
                    buf = (R)((double)buf + (double)s.Cezhy[e] * dx * (double)ps);
                }
                if (y <= nc && y != 0) {
                    int i = nc - y;
                    buf = s.Ceze[e] * ez + s.Cezhy[e] * dHy / s.kex[i] - s.Cezhx[e] * dHx / s.kex[i];
                    R& ps = s.psi_ezy[(size_t)(y - 1) * (nx + 1) + x];
                    ps = s.bex[i] * ps + s.aex[i] * dHx;
                    buf = (R)((double)buf - (double)s.Cezhx[e] * dy * (double)ps);
                }
                if (y >= ny - nc && y != ny) {
                    int i = y - ny + nc;
                    buf = s.Ceze[e] * ez + s.Cezhy[e] * dHy / s.kex[i] - s.Cezhx[e] * dHx / s.kex[i];
                    R& ps = s.psi_ezy[(size_t)(nc + i) * (nx + 1) + x];
                    ps = s.bex[i] * ps + s.aex[i] * dHx;
                    buf = (R)((double)buf - (double)s.Cezhx[e] * dy * (double)ps);
                }
            }
            if (clamp && isnan_(buf)) buf = 0;
            s.Ez[e] = buf;
        }
    }
}

// Physical-mode updates: Yee leapfrog with Roden-Gedney CPML on both axes
// (corners keep both psi terms, each kappa divides its own derivative),
// Debye polarisation current Jp = Kp Jp + (Bp/dt)(E^{n+1} - E^n), soft source.
template <class R>
static void phys_H(Sim<R>& s) {
    const int nx = s.nx, ny = s.ny, nc = s.nc;
    const R Chy = (R)(s.c.dt / (kMU0 * s.c.dx)), Chx = (R)(s.c.dt / (kMU0 * s.c.dy));
    const R dx = (R)s.c.dx, dy = (R)s.c.dy;
    for (int y = 0; y < ny; ++y) {
        for (int x = 0; x < nx; ++x) {
            const R ez = s.Ez[s.E(x, y)];
            const R dEx = s.Ez[s.E(x + 1, y)] - ez;
            const R dEy = s.Ez[s.E(x, y + 1)] - ez;
            R ty = dEx, tx = dEy;
            if (x < nc || x >= nx - nc) {
                int i = x < nc ? nc - 1 - x : x - nx + nc;
                int j = x < nc ? x : nc + i;
                R& ps = s.psi_hyx[(size_t)y * 2 * nc + j];
                ps = s.bmx[i] * ps + s.amx[i] * dEx;
                ty = dEx / s.kmx[i] + dx * ps;
            }
            if (y < nc || y >= ny - nc) {
                int i = y < nc ? nc - 1 - y : y - ny + nc;
                int j = y < nc ? y : nc + i;
                R& ps = s.psi_hxy[(size_t)j * (nx + 1) + x];
                ps = s.bmx[i] * ps + s.amx[i] * dEy;
                tx = dEy / s.kmx[i] + dy * ps;
            }
            s.Hy[s.H(x, y)] += Chy * ty;
            s.Hx[s.H(x, y)] -= Chx * tx;
        }
    }
}

template <class R>
static void phys_E(Sim<R>& s, int it) {
    const int nx = s.nx, ny = s.ny, nc = s.nc;
    const R dx = (R)s.c.dx, dy = (R)s.c.dy;
    for (int y = 1; y < ny; ++y) {
        for (int x = 1; x < nx; ++x) {
            const int e = s.E(x, y);
            const R ez = s.Ez[e];
            const R dHy = s.Hy[s.H(x, y)] - s.Hy[s.H(x - 1, y)];
            const R dHx = s.Hx[s.H(x, y)] - s.Hx[s.H(x, y - 1)];
            R cy_ = dHy, cx_ = dHx;
            if (x <= nc || x >= nx - nc) {
                int i = x <= nc ? nc - x : x - nx + nc;
                int j = x <= nc ? x - 1 : nc + i;
                R& ps = s.psi_ezx[(size_t)y * 2 * nc + j];
                ps = s.bex[i] * ps + s.aex[i] * dHy;
                cy_ = dHy / s.kex[i] + dx * ps;
            }
            if (y <= nc || y >= ny - nc) {
                int i = y <= nc ? nc - y : y - ny + nc;
                int j = y <= nc ? y - 1 : nc + i;
                R& ps = s.psi_ezy[(size_t)j * (nx + 1) + x];
                ps = s.bex[i] * ps + s.aex[i] * dHx;
                cx_ = dHx / s.kex[i] + dy * ps;
            }
            R en = s.Ceze[e] * ez + s.Cezhy[e] * cy_ - s.Cezhx[e] * cx_ - s.Cezjp[e] * s.Jp[e];
            if (x == s.sx && y == s.sy) en += s.srcw[it];
            s.Jp[e] = s.Kpc[e] * s.Jp[e] + s.BpDt[e] * (en - ez);
            s.Ez[e] = en;
        }
    }
}

// Running DFT: calculate_JandM (FDTD_GPU.cu:231-280).  In ref mode with
// R = float the phase argument is formed in float exactly as the reference
// does; the FP64 oracle uses exact phases.  Physical mode samples H at
// (n+1/2) dt and E at (n+1) dt (the times the leapfrog defines them).
template <class R>
static void dft_step(Sim<R>& s, int it) {
    const tomo_sim_config& c = s.c;
    const int P = s.ct.P, nx = s.nx;
    const double dt = c.dt;
    const double tJ = s.ref ? 0.0 : 0.5, tM = s.ref ? 0.5 : 1.0;
    const bool ref_float = s.ref && sizeof(R) == 4;
    for (int f = 0; f < c.n_freq; ++f) {
        std::complex<R> wJ, wM;
        R dtR;
        if (ref_float) {
            float freq = (float)s.freq[f];
            float dtf = (float)dt;
            float twopi = -2.0f * (float)kPI;
            float thJ = ((twopi * freq) * (float)it) * dtf;
            float thM = ((twopi * freq) * (float)((float)it + 0.5)) * dtf;
            wJ = std::complex<R>((R)(float)std::cos((double)thJ), (R)(float)std::sin((double)thJ));
            wM = std::complex<R>((R)(float)std::cos((double)thM), (R)(float)std::sin((double)thM));
            dtR = (R)dtf;
        } else {
            double w = 2.0 * kPI * s.freq[f];
            wJ = std::complex<R>((R)std::cos(-w * (it + tJ) * dt), (R)std::sin(-w * (it + tJ) * dt));
            wM = std::complex<R>((R)std::cos(-w * (it + tM) * dt), (R)std::sin(-w * (it + tM) * dt));
            dtR = (R)dt;
        }
        for (int p = 0; p < P; ++p) {
            const int x = s.ct.x[p], y = s.ct.y[p], g = s.ct.leg[p];
            R h, ez, sg;
            if (g == 0 || g == 2) {  // y faces: Hx and Ez averaged over y, y+1
                ez = (s.Ez[s.E(x, y + 1)] + s.Ez[s.E(x, y)]) / 2;
                h = s.Hx[s.H(x, y)];
                sg = (g == 2) ? -1 : 1;
            } else {                 // x faces: Hy and Ez averaged over x, x+1
                ez = (s.Ez[s.E(x, y)] + s.Ez[s.E(x + 1, y)]) / 2;
                h = s.Hy[s.H(x, y)];
                sg = (g == 3) ? -1 : 1;
            }
            (void)nx;
            R vJ = (sg * h) * dtR, vM = (sg * ez) * dtR;
            std::complex<R>& Jr = s.J[(size_t)f * P + p];
            std::complex<R>& Mr = s.M[(size_t)f * P + p];
            Jr = std::complex<R>(Jr.real() + vJ * wJ.real(), Jr.imag() + vJ * wJ.imag());
            Mr = std::complex<R>(Mr.real() + vM * wM.real(), Mr.imag() + vM * wM.imag());
        }
    }
}

template <class R>
static void sim_step(Sim<R>& s, int nsteps) {
    for (int k = 0; k < nsteps; ++k) {
        if (s.n >= s.c.steps) return;
        if (s.ref) { ref_H(s); ref_E(s, s.n); }
        else { phys_H(s); phys_E(s, s.n); }
        dft_step(s, s.n);
        s.n++;
    }
}

// calc_incident_power (FDTD_common.cxx:48-51).
static double incident_power(const tomo_sim_config& c, double freq) {
    double tau = c.pulse_tau;
    return (0.5 / eta0()) * std::pow(tau * std::sqrt(kPI) * std::exp(-tau * tau * 2 * kPI * freq * 2 * kPI * freq / 4), 2);
}

// Discrete spectrum power of the physical source waveform at f.
static double source_power(const tomo_sim_config& c, double freq) {
    double w = 2.0 * kPI * freq;
    std::complex<double> S = 0;
    for (int n = 0; n < c.steps; ++n) {
        double t = (n + 1) * c.dt;
        S += phys_source_value(c, n) * c.dt * std::exp(std::complex<double>(0.0, -w * t));
    }
    return std::norm(S);
}

// Contour point positions in metres and per-point weights.
//  ref:      N2FPostProcess (FDTD_GPU.cu:553-634): (x [+.5], y [+.5]) - (nx/2 - 1)
//  physical: true Yee face positions relative to the domain centre nx/2.
static void contour_geometry(const tomo_sim_config& c, const Contour& ct, std::vector<double>& px,
                             std::vector<double>& py) {
    px.resize(ct.P); py.resize(ct.P);
    const bool ref = c.mode == TOMO_MODE_REF_COMPAT;
    for (int p = 0; p < ct.P; ++p) {
        double fx = ct.x[p], fy = ct.y[p];
        if (ct.leg[p] == 0 || ct.leg[p] == 2) fy += 0.5; else fx += 0.5;
        if (ref) {
            px[p] = c.dx * (-(double)(float)c.nx / 2 + 1 + fx);
            py[p] = c.dy * (-(double)(float)c.ny / 2 + 1 + fy);
        } else {
            px[p] = c.dx * (fx - c.nx / 2.0);
            py[p] = c.dy * (fy - c.ny / 2.0);
        }
    }
}

// Far-field projection; D and P are [freq][angle].  Accumulates in the
// oracle's R for the DFT inputs but forms every phase in double.
template <class R>
static void sim_farfield(const Sim<R>& s, std::vector<double>& D, std::vector<double>& Pph) {
    const tomo_sim_config& c = s.c;
    const int P = s.ct.P, A = c.n_angles, F = c.n_freq;
    std::vector<double> px, py;
    contour_geometry(c, s.ct, px, py);
    D.assign((size_t)F * A, 0);
    Pph.assign((size_t)F * A, 0);
    const double e0 = eta0();
    for (int f = 0; f < F; ++f) {
        double freq = s.ref ? (double)(float)s.freq[f] : s.freq[f];
        double k = 2 * kPI * freq / c0();
        double norm = s.ref ? 8 * kPI * e0 * incident_power(c, freq) * c.ff_norm
                            : 8 * kPI * e0 * source_power(c, freq);
        for (int a = 0; a < A; ++a) {
            double phi = 2 * kPI / A * a;
            double cph = std::cos(phi), sph = std::sin(phi);
            std::complex<double> N = 0, L = 0;
            for (int p = 0; p < P; ++p) {
                double arg = k * (px[p] * cph + py[p] * sph);
                std::complex<double> w(std::cos(arg), std::sin(arg));
                std::complex<double> Jv((double)s.J[(size_t)f * P + p].real(), (double)s.J[(size_t)f * P + p].imag());
                std::complex<double> Mv((double)s.M[(size_t)f * P + p].real(), (double)s.M[(size_t)f * P + p].imag());
                int g = s.ct.leg[p];
                double del = (g == 0 || g == 2) ? c.dx : c.dy;
                N += -Jv * w * del;
                if (g == 0 || g == 2) L += -sph * Mv * w * c.dx;
                else L += cph * Mv * w * c.dy;
            }
            std::complex<double> Ef = L + e0 * N;
            D[(size_t)f * A + a] = k * k * std::norm(Ef) / norm;
            Pph[(size_t)f * A + a] = std::atan2(Ef.imag(), Ef.real());
        }
    }
}

struct Handle {
    int kind;  // 0 float, 1 double
    Sim<float> sf;
    Sim<double> sd;
};

// Fitness of far-field records D, P (n_exc x n_freq x n_angles each) against
// a target D (and, for TOMO_FLAG_PHASE_FITNESS, a target phase), in FP64:
//   FDTD_GPU.cu:1004-1008   -sum (D - m)^2 / (n_exc n_freq)
//   FDTD_CPU.cxx:841-845    -sum (D - m)^2 / (m^2 n_exc n_freq)   (TOMO_FLAG_REL_FITNESS)
//   FDTD_common.cxx:53-61   -sum (m - D)^2 / (m^2 n_exc)          (+ TOMO_FLAG_FITNESS_PER_EXC)
//   FDTD_GPU.cu:1009-1018   phase term -dphi^2 / (10 pi^2 n_exc n_freq), dphi wrapped to [-pi, pi]
static double fitness_of(const tomo_sim_config& c, const double* D, const double* P, const double* tgt,
                         const float* tphase, size_t L) {
    double fit = 0;
    const double norm = (c.flags & TOMO_FLAG_FITNESS_PER_EXC) ? (double)c.n_exc : (double)c.n_exc * c.n_freq;
    const bool rel = (c.flags & TOMO_FLAG_REL_FITNESS) != 0;
    for (size_t i = 0; i < L; ++i) fit -= (D[i] - tgt[i]) * (D[i] - tgt[i]) / (rel ? tgt[i] * tgt[i] * norm : norm);
    if ((c.flags & TOMO_FLAG_PHASE_FITNESS) && tphase && P) {
        for (size_t i = 0; i < L; ++i) {
            double dphi = P[i] - (double)tphase[i];
            if (dphi > kPI) dphi -= 2 * kPI;
            else if (dphi < -kPI) dphi += 2 * kPI;
            fit -= dphi * dphi / (10 * kPI * kPI * norm);
        }
    }
    return fit;
}

// Far field of one excitation run for c.steps steps.
template <class R>
static bool farfield_one(const tomo_sim_config& c, const double* p, size_t np, int exc, std::vector<double>& D,
                         std::vector<double>& P) {
    Sim<R> s;
    if (!sim_init(s, c, p, np, exc)) return false;
    sim_step(s, c.steps);
    sim_farfield(s, D, P);
    return true;
}

// Full evaluation: every excitation, far field, fitness against target.
template <class R>
static bool eval_full(const tomo_sim_config& c, const double* p, size_t np, const float* target, size_t nt,
                      double* fitness, double* Dout, double* Pout) {
    const int F = c.n_freq, A = c.n_angles;
    const size_t L = (size_t)c.n_exc * F * A;
    std::vector<double> Dall(L), Pall(L);
    for (int e = 0; e < c.n_exc; ++e) {
        Sim<R> s;
        if (!sim_init(s, c, p, np, e)) return false;
        sim_step(s, c.steps);
        std::vector<double> D, Pp;
        sim_farfield(s, D, Pp);
        std::copy(D.begin(), D.end(), Dall.begin() + (size_t)e * F * A);
        std::copy(Pp.begin(), Pp.end(), Pall.begin() + (size_t)e * F * A);
    }
    std::vector<double> tgt(L, 0.0);
    if (c.mode == TOMO_MODE_PHYSICAL && target == nullptr) {
        std::vector<double> tv = target_vector(c);
        for (int e = 0; e < c.n_exc; ++e) {
            Sim<R> s;
            if (!sim_init(s, c, tv.data(), tv.size(), e)) return false;
            sim_step(s, c.steps);
            std::vector<double> D, Pp;
            sim_farfield(s, D, Pp);
            std::copy(D.begin(), D.end(), tgt.begin() + (size_t)e * F * A);
        }
    } else {
        if (target == nullptr || nt < L) { g_err = "target too short"; return false; }
        for (size_t i = 0; i < L; ++i) tgt[i] = target[i];
    }
    const float* tphase = (target != nullptr && nt >= 2 * L) ? target + L : nullptr;
    *fitness = fitness_of(c, Dall.data(), Pall.data(), tgt.data(), tphase, L);
    if (Dout) std::copy(Dall.begin(), Dall.end(), Dout);
    if (Pout) std::copy(Pall.begin(), Pall.end(), Pout);
    return true;
}

}  // namespace orc

using namespace orc;

template <class R>
static void copy_fields(const Sim<R>& s, double* ez, double* hx, double* hy, double* jp) {
    if (ez) for (size_t i = 0; i < s.Ez.size(); ++i) ez[i] = s.Ez[i];
    if (hx) for (size_t i = 0; i < s.Hx.size(); ++i) hx[i] = s.Hx[i];
    if (hy) for (size_t i = 0; i < s.Hy.size(); ++i) hy[i] = s.Hy[i];
    if (jp) for (size_t i = 0; i < s.Jp.size(); ++i) jp[i] = s.Jp[i];
}

// Per-cell coefficients (9 planes: Ceze, Cezhy, Cezhx, Cezeic, Cezeip, Cezjp, Beta_p, BpDt, Kp).
template <class R>
static void copy_mat(const Sim<R>& s, uint8_t* mat, double* coef) {
    size_t N1 = s.mat.size();
    if (mat) std::memcpy(mat, s.mat.data(), N1);
    if (coef) {
        const std::vector<R>* planes[9] = {&s.Ceze, &s.Cezhy, &s.Cezhx, &s.Cezeic, &s.Cezeip,
                                           &s.Cezjp, &s.Beta_p, &s.BpDt, &s.Kpc};
        for (int k = 0; k < 9; ++k)
            for (size_t i = 0; i < N1; ++i) coef[k * N1 + i] = (*planes[k])[i];
    }
}


extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_contour(const tomo_sim_config* c, int* xs, int* ys, int* legs) {
    Contour ct = make_contour(c->nx, c->ny, c->ntff_offset, c->mode == TOMO_MODE_PHYSICAL);
    for (int p = 0; p < ct.P; ++p) { xs[p] = ct.x[p]; ys[p] = ct.y[p]; legs[p] = ct.leg[p]; }
    return ct.P;
}

int orc_sim_create(const tomo_sim_config* c, const double* p, size_t np, int exc, int fp64, void** out) {
    Handle* h = new Handle();
    h->kind = fp64 ? 1 : 0;
    bool ok = fp64 ? sim_init(h->sd, *c, p, np, exc) : sim_init(h->sf, *c, p, np, exc);
    if (!ok) { delete h; return -1; }
    *out = h;
    return 0;
}

void orc_sim_destroy(void* hp) { delete (Handle*)hp; }

int orc_sim_step(void* hp, int n) {
    Handle* h = (Handle*)hp;
    if (h->kind) sim_step(h->sd, n); else sim_step(h->sf, n);
    return 0;
}

// Ez at one node (the value FDTD_CPU.cxx:808 prints for the centre node).
double orc_sim_probe(void* hp, int x, int y) {
    Handle* h = (Handle*)hp;
    return h->kind ? h->sd.Ez[h->sd.E(x, y)] : (double)h->sf.Ez[h->sf.E(x, y)];
}

int orc_sim_fields(void* hp, double* ez, double* hx, double* hy, double* jp) {
    Handle* h = (Handle*)hp;
    if (h->kind) copy_fields(h->sd, ez, hx, hy, jp); else copy_fields(h->sf, ez, hx, hy, jp);
    return 0;
}

int orc_sim_materials(void* hp, uint8_t* mat, double* coef) {
    Handle* h = (Handle*)hp;
    if (h->kind) copy_mat(h->sd, mat, coef); else copy_mat(h->sf, mat, coef);
    return 0;
}

// CPML tables: 6 x ncells (kex, kmx, aex, bex, amx, bmx).
int orc_sim_cpml(void* hp, double* out) {
    Handle* h = (Handle*)hp;
    auto put = [&](auto& s) {
        int nc = s.nc;
        for (int i = 0; i < nc; ++i) {
            out[0 * nc + i] = s.kex[i]; out[1 * nc + i] = s.kmx[i]; out[2 * nc + i] = s.aex[i];
            out[3 * nc + i] = s.bex[i]; out[4 * nc + i] = s.amx[i]; out[5 * nc + i] = s.bmx[i];
        }
    };
    if (h->kind) put(h->sd); else put(h->sf);
    return 0;
}

// J, M as interleaved complex [freq][point].
int orc_sim_dft(void* hp, double* J, double* M) {
    Handle* h = (Handle*)hp;
    auto put = [&](auto& s) {
        for (size_t i = 0; i < s.J.size(); ++i) {
            J[2 * i] = s.J[i].real(); J[2 * i + 1] = s.J[i].imag();
            M[2 * i] = s.M[i].real(); M[2 * i + 1] = s.M[i].imag();
        }
    };
    if (h->kind) put(h->sd); else put(h->sf);
    return 0;
}

int orc_sim_farfield(void* hp, double* D, double* P) {
    Handle* h = (Handle*)hp;
    std::vector<double> d, pp;
    if (h->kind) sim_farfield(h->sd, d, pp); else sim_farfield(h->sf, d, pp);
    std::copy(d.begin(), d.end(), D);
    if (P) std::copy(pp.begin(), pp.end(), P);
    return 0;
}

// Far field D, P ([freq][angle], FP64 results of the R-precision run) of one
// excitation run to c->steps: the unit a test can spread over host threads.
int orc_farfield(const tomo_sim_config* c, const double* p, size_t np, int exc, int fp64, double* D, double* P) {
    std::vector<double> d, pp;
    const bool ok = fp64 ? farfield_one<double>(*c, p, np, exc, d, pp) : farfield_one<float>(*c, p, np, exc, d, pp);
    if (!ok) return -1;
    std::copy(d.begin(), d.end(), D);
    if (P) std::copy(pp.begin(), pp.end(), P);
    return 0;
}

// Fitness of given far-field records (the arithmetic orc_eval applies);
// tphase may be NULL, P too unless TOMO_FLAG_PHASE_FITNESS is set.
int orc_fitness(const tomo_sim_config* c, const double* D, const double* P, const float* target,
                const float* tphase, size_t L, double* fitness) {
    std::vector<double> tgt(target, target + L);
    *fitness = fitness_of(*c, D, P, tgt.data(), tphase, L);
    return 0;
}

// The full evaluation mirrors FDTD_GPU(): every excitation, far field, fitness.
int orc_eval(const tomo_sim_config* c, const double* p, size_t np, const float* target, size_t nt, int fp64,
             double* fitness, double* D, double* P) {
    bool ok = fp64 ? eval_full<double>(*c, p, np, target, nt, fitness, D, P)
                   : eval_full<float>(*c, p, np, target, nt, fitness, D, P);
    return ok ? 0 : -1;
}

// Physical-mode target vector and source cell (for tests of the product's setup).
int orc_target_vector(const tomo_sim_config* c, double* out8) {
    std::vector<double> v = target_vector(*c);
    for (int i = 0; i < 8; ++i) out8[i] = v[i];
    return 0;
}

int orc_source_cell(const tomo_sim_config* c, int exc, int* sx, int* sy) {
    phys_source_cell(*c, exc, *sx, *sy);
    return 0;
}

double orc_incident_power(const tomo_sim_config* c, double f) { return incident_power(*c, f); }

}  // extern "C"
