// setup.h — host-side derivation of everything a simulation needs before its
// first step: presets, material coefficient tables, CPML profiles, phantom
// geometry for the device rasteriser, source/incident-wave constants and
// far-field normalisers.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "tomo_internal.h"

namespace tomo {

extern thread_local std::string g_last_error;
int fail(int code, const std::string& msg);

constexpr double kPI = 3.141592653589793238;
constexpr double kEPS0 = 8.85418e-12;   // FDTD_common.hxx:5
constexpr double kMU0 = kPI * 4e-7;     // FDTD_common.hxx:10
double c0();
double eta0();

// Geometry of the seeded physical phantom, in the exact FP32 form the device
// rasteriser evaluates (all primitives relative to the node (nx/2, ny/2)).
struct PhantomGeom {
    float ro2, ri2;
    float ell[3][6];   // cx, cy, cos, sin, 1/a^2, 1/b^2
    float tum[4][3];   // cx, cy, r^2
    int ntum;
    int pad;
};

// Reference disc rasterisation inputs: per-row inclusive extent of the disc
// and per-axis patch coordinates (getOptimizationCell, FDTD_GPU.cu:544-551).
struct RefRaster {
    std::vector<int> lo, hi;  // per row; hi < lo means empty row
    std::vector<int> xc, yc;  // per column / row patch coordinate
};

struct Tissue {
    double einf, deps, tau, sigma;
};

// Everything derived on the host for one simulation.
struct SimSetup {
    tomo_sim_config cfg{};
    int exc = 0;
    int pitch = 0;
    int P = 0, L1 = 0, L2 = 0, L3 = 0;
    int cx_hi = 0, cy_lo = 0;  // last x of the y faces, first y of the x faces
    std::vector<Tissue> tissues;      // material id -> tissue (physical)
    std::vector<float> image;         // unpacked parameters (ref)
    float coef[TOMO_MAX_MAT * TOMO_NCOEF] = {};
    int nmat = 0;
    float ke[TOMO_MAX_NC], km[TOMO_MAX_NC], ike[TOMO_MAX_NC], ikm[TOMO_MAX_NC];
    float ae[TOMO_MAX_NC], be[TOMO_MAX_NC], am[TOMO_MAX_NC], bm[TOMO_MAX_NC];
    float chx = 0, chy = 0, chx_d = 0, chy_d = 0;
    PhantomGeom geom{};
    RefRaster ref;
    std::vector<int> span;            // lo, hi, off per row
    int jp_len = 0;
    int sx = -1, sy = -1;
    std::vector<float> srcw;
    float inc_dt = 0, inc_c = 0, inc_ax = 0, inc_ay = 0, inc_itau = 0;
    int inc_abs = 0;
    double freq[TOMO_MAX_FREQ] = {};
    double wdt[TOMO_MAX_FREQ] = {};
    double inv_norm[TOMO_MAX_FREQ] = {};
    double tJ = 0, tM = 0;
};

int preset(int id, int grid, tomo_sim_config* out);
int validate(const tomo_sim_config& c);
int build_setup(const tomo_sim_config& c, const double* params, size_t n, int exc, SimSetup& s);
std::vector<double> target_vector(const tomo_sim_config& c);
int contour_len(const tomo_sim_config& c);

}  // namespace tomo
