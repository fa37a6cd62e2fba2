// tomo_internal.h — data layout shared by the host setup, the HIP kernels and
// the runtime of the MI355X FDTD fitness library.
//
// HBM layout of one simulation slot (all rows padded to `pitch` floats):
//   ez[2], hx[2], hy[2]   : (ny+1) x pitch FP32, ping-pong (state n -> n+T)
//   mat                   : (ny+1) x pitch uint8 material id
//   jp[2]                 : compact Debye polarisation current, one contiguous
//                           run per row [span_lo, span_hi) of the dispersive
//                           region (row offsets in span[3*(ny+1)])
//   pex[2], phx[2]        : x-strip CPML psi (E and H), (ny+1) x 2*nc
//   pey[2], phy[2]        : y-strip CPML psi (E and H), 2*nc x pitch
//   samp                  : surface samples of one DFT chunk, 3 x K x P
//                           (H, Ez at the face cell, Ez at its neighbour)
//   accJ, accM            : running DFT, F x P float4 {re, im, comp_re, comp_im}
//                           (Kahan FP32) or F x P double2 (FP64 study mode)
// Only CPML strips and dispersive rows carry psi / Jp state.
#pragma once
#include <stdint.h>

#include "../../include/tomo.h"

#define TOMO_MAX_MAT TOMO_MAX_MATERIALS  // materials per simulation (ref: 1 + numpatches^2, 82 for the CPU path)
#define TOMO_MAX_NC 32    // CPML cells
#define TOMO_MAX_FREQ 32
#define TOMO_NCOEF 8      // Ceze, Cezh, Cezeic, Cezeip, Cezjp, Kp, BpDt, dispersive(0/1)

// Device-resident descriptor of one simulation.
struct DevSim {
    float* ez[2];
    float* hx[2];
    float* hy[2];
    float* jp[2];
    float* pex[2];
    float* phx[2];
    float* pey[2];
    float* phy[2];
    // the same buffers as (base, stride): field planes ez0 ez1 hx0 hx1 hy0 hy1
    // at fbase + k*ps, jp0 jp1 at jbase + k*js, pex0 pex1 phx0 phx1 at
    // xbase + k*xs, pey0 pey1 phy0 phy1 at ybase + k*ys (strides in floats)
    float* fbase;
    float* jbase;
    float* xbase;
    float* ybase;
    int ps, js, xs, ys;
    const uint8_t* mat;
    const int* span;       // lo, hi, off per row
    float* samp;           // 3 x K x P
    void* accJ;            // F x P
    void* accM;
    const float* srcw;     // physical source waveform per step (nullptr in ref mode)
    const float* pos;      // contour point positions in metres: x[P] then y[P]
    int nx, ny, pitch, nc;
    int D, P, L1, L2, L3;  // NTFF contour
    int cx_hi, cy_lo;      // last x of the y faces, first y of the x faces
    int sx, sy;            // physical source cell
    int mode, flags;
    int nmat, nfreq, chunk, steps;
    float coef[TOMO_MAX_MAT * TOMO_NCOEF];
    // CPML 1-D profiles: E (inverse kappa, a, b) and H (inverse kappa, a, b)
    float ike[TOMO_MAX_NC], ae[TOMO_MAX_NC], be[TOMO_MAX_NC];
    float ikm[TOMO_MAX_NC], am[TOMO_MAX_NC], bm[TOMO_MAX_NC];
    float ke[TOMO_MAX_NC], km[TOMO_MAX_NC];  // kappa itself (ref-compat divides)
    float chx, chy;        // H update: dt/(mu0 dy), dt/(mu0 dx)   (ref: Chez for both)
    float chx_d, chy_d;    // H psi scale: Chez*dy, Chez*dx       (physical: Chx*dy, Chy*dx)
    float dxf, dyf;        // E psi scale factors (dx, dy)
    // reference incident plane wave (ref mode): ar = (n+-0.5) dt + inc_c - inc_ax X - inc_ay Y
    float inc_dt, inc_c, inc_ax, inc_ay, inc_itau;
    int inc_abs_origin;
    // running DFT phase: omega_f * dt (FP64) and time offsets for J and M
    double wdt[TOMO_MAX_FREQ];
    double tJ, tM;
    float sgn_dt[4];       // per leg: sign * dt (legs yn, xp, yp, xn)
};

// One simulation's share of an epoch (K = chunk steps, T steps per launch).
struct LaunchItem {
    int sim;       // index into the DevSim table
    int step0;     // first step of this epoch
    int parity0;   // buffer holding the state at step0
    int nsteps;    // steps to do in this epoch (<= chunk)
};

// Tile classes of physical-mode simulations (TileItem::cls).
enum { TILE_CORNER = 0, TILE_ROW = 1, TILE_COL = 2, TILE_INNER = 3 };

// Work unit of the time-blocked update: `nband` (<= waves) adjacent column
// bands over output rows [y0, y1) of one simulation.  Band k owns columns
// [x0 + k V, min(x0 + (k+1) V, xe)) with V = 64 - 2T and reads T ghost
// columns on each side.  Interior tiles (no CPML strip, NTFF contour
// row/column or grid edge within their ghost zone, full-width bands) run the
// lean interior kernel; the rest the edge kernel.
struct TileItem {
    int item;      // index into the LaunchItem array
    int x0;        // first owned column of band 0
    int nband;     // bands in the group
    int y0, y1;    // output rows
    int xe;        // end of the owned columns of the tile (exclusive)
    int cls;       // TILE_* (physical mode)
    int pad;
};


// Tile of the resident update form: one workgroup keeps a band of 64 lanes
// (owned columns [x0, x0 + 64 - 2T), T ghost columns on each side) times the
// owned rows [y0, y1) plus T ghost rows above and below in registers for a
// whole epoch, and exchanges only a T-deep frame with its neighbours between
// time blocks.  nb: the eight neighbouring tiles of the same simulation
// (-1 where the grid ends), whose completed-block counters it waits for.
struct ResTile {
    int item;      // index into the LaunchItem array
    int x0;        // first owned column
    int y0, y1;    // owned rows
    int nb[8];
};
