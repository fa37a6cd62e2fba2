// kernels.hip — gfx950 kernels of the 2D TMz FDTD fitness evaluation.
//
//   k_tblock    time-blocked E/H update (replaces H_field_update +
//               E_field_update + calculate_JandM's sampling, FDTD_GPU.cu:231-434)
//   k_fold      running DFT of one chunk of surface samples (calculate_JandM's
//               accumulation, FDTD_GPU.cu:231-280), Kahan FP32 or FP64
//   k_project   far-field projection (N2FPostProcess, FDTD_GPU.cu:553-634)
//   k_fitness   FP64 fitness reduction (FDTD_GPU.cu:1004-1018)
//   k_raster_*  material map (FDTD_GPU.cu:718-743 / the seeded phantom)
//
// k_tblock layout: one wave64 owns a band of 64 consecutive Ez columns and
// sweeps the rows of its tile bottom to top.  The wave keeps T time levels in
// flight in registers: stage t receives row y+1 at level n+t and emits row y
// at level n+t+1 to stage t+1, so each row is read from HBM once and written
// once per T steps.  Lateral neighbours (Ez(x+1), Hy(x-1)) come from the
// adjacent lane; every stage invalidates one lane at each band edge and one
// row at each tile edge, so bands overlap by 2T columns and tiles read T ghost
// rows above and below: stage t is exact on lanes [t+1, 63-t) and the final
// stage on [T, 64-T).  State is ping-ponged between two buffers per launch so
// that overlapping ghost reads never see a neighbour's update.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace tomo {
namespace {

constexpr int kWave = 64;

struct Cell {
    float ez, hx, hy, jp, pex, phx, pey, phy;
    int mat;
};

__device__ __forceinline__ Cell zero_cell() {
    Cell c;
    c.ez = c.hx = c.hy = c.jp = c.pex = c.phx = c.pey = c.phy = 0.f;
    c.mat = 0;
    return c;
}

// Scalar (wave-uniform) view of a DevSim, held in SGPRs.
struct SimView {
    float *ez, *hx, *hy, *jp, *pex, *phx, *pey, *phy;   // buffers read (parity par)
    float *oez, *ohx, *ohy, *ojp, *opex, *ophx, *opey, *ophy;  // buffers written
    const uint8_t* mat;
    const int* span;
    float* samp;
    const float* srcw;
    int nx, ny, pitch, nc, D, P, L1, L2, L3, cx_hi, cy_lo, sx, sy, flags, K;
    float chx, chy, chx_d, chy_d, dxf, dyf;
    float inc_dt, inc_c, inc_ax, inc_ay, inc_itau;
    int inc_abs;
};

__device__ __forceinline__ void load_cell(const SimView& S, const float* __restrict__ lcoef, int r, int x,
                                          bool xin, Cell& c) {
    c = zero_cell();
    if (r < 0 || r > S.ny || !xin) return;
    const size_t base = (size_t)r * S.pitch + x;
    c.ez = S.ez[base];
    c.mat = S.mat[base];
    if (x < S.nx && r < S.ny) {
        c.hx = S.hx[base];
        c.hy = S.hy[base];
    }
    const int lo = S.span[3 * r], hi = S.span[3 * r + 1];
    if (x >= lo && x < hi && lcoef[c.mat * TOMO_NCOEF + 7] != 0.f) c.jp = S.jp[S.span[3 * r + 2] + x - lo];
    const int nc = S.nc, nx = S.nx, ny = S.ny, w = 2 * nc;
    if (x >= 1 && x <= nc) c.pex = S.pex[r * w + x - 1];
    else if (x >= nx - nc && x <= nx - 1) c.pex = S.pex[r * w + nc + x - (nx - nc)];
    if (x < nc) c.phx = S.phx[r * w + x];
    else if (x >= nx - nc && x < nx) c.phx = S.phx[r * w + nc + x - (nx - nc)];
    if (r >= 1 && r <= nc) c.pey = S.pey[(size_t)(r - 1) * S.pitch + x];
    else if (r >= ny - nc && r <= ny - 1) c.pey = S.pey[(size_t)(nc + r - (ny - nc)) * S.pitch + x];
    if (r < nc) c.phy = S.phy[(size_t)r * S.pitch + x];
    else if (r >= ny - nc && r < ny) c.phy = S.phy[(size_t)(nc + r - (ny - nc)) * S.pitch + x];
}

__device__ __forceinline__ void store_cell(const SimView& S, const float* __restrict__ lcoef, int r, int x,
                                           const Cell& c) {
    const size_t base = (size_t)r * S.pitch + x;
    S.oez[base] = c.ez;
    if (x < S.nx && r < S.ny) {
        S.ohx[base] = c.hx;
        S.ohy[base] = c.hy;
    }
    const int lo = S.span[3 * r], hi = S.span[3 * r + 1];
    if (x >= lo && x < hi && lcoef[c.mat * TOMO_NCOEF + 7] != 0.f) S.ojp[S.span[3 * r + 2] + x - lo] = c.jp;
    const int nc = S.nc, nx = S.nx, ny = S.ny, w = 2 * nc;
    if (x >= 1 && x <= nc) S.opex[r * w + x - 1] = c.pex;
    else if (x >= nx - nc && x <= nx - 1) S.opex[r * w + nc + x - (nx - nc)] = c.pex;
    if (x < nc) S.ophx[r * w + x] = c.phx;
    else if (x >= nx - nc && x < nx) S.ophx[r * w + nc + x - (nx - nc)] = c.phx;
    if (r >= 1 && r <= nc) S.opey[(size_t)(r - 1) * S.pitch + x] = c.pey;
    else if (r >= ny - nc && r <= ny - 1) S.opey[(size_t)(nc + r - (ny - nc)) * S.pitch + x] = c.pey;
    if (r < nc) S.ophy[(size_t)r * S.pitch + x] = c.phy;
    else if (r >= ny - nc && r < ny) S.ophy[(size_t)(nc + r - (ny - nc)) * S.pitch + x] = c.phy;
}

// LDS CPML table layout: [0] ke [1] km [2] ike [3] ikm [4] ae [5] be [6] am [7] bm, each TOMO_MAX_NC.
#define CP(k, i) lcp[(k) * TOMO_MAX_NC + (i)]

// One leapfrog step of one row: c holds row y at level n (fields Ez^n,
// H^{n-1/2}); ez_up = Ez^n(x, y+1); hx_prev = Hx^{n+1/2}(x, y-1) on entry and
// Hx^{n+1/2}(x, y) on exit.  On exit c holds row y at level n+1.
template <int MODE>
__device__ __forceinline__ void stage_update(const SimView& S, const float* __restrict__ lcoef,
                                             const float* __restrict__ lcp, Cell& c, float ez_up, float& hx_prev,
                                             int x, int y, int n) {
    const int nx = S.nx, ny = S.ny, nc = S.nc;
    // ---------------- H half-step (H_field_update, FDTD_GPU.cu:298-355)
    const float ez_r = __shfl_down(c.ez, 1);
    const float dEx = ez_r - c.ez, dEy = ez_up - c.ez;
    float hy, hx;
    // H nodes exist for 0 <= x < nx, 0 <= y < ny; the CPML branches run only there
    const bool hx_in = (x >= 0 && x < nx), hy_in = (y >= 0 && y < ny);
    if (MODE == TOMO_MODE_REF_COMPAT) {
        const float chez = S.chx;
        hy = c.hy + chez * dEx;
        hx = c.hx - chez * dEy;
        if (hx_in && x < nc) {
            const int i = nc - 1 - x;
            hy = c.hy + chez * dEx / CP(1, i);
            c.phx = CP(7, i) * c.phx + CP(6, i) * dEx;
            hy = hy + S.chy_d * c.phx;
        }
        if (hx_in && x >= nx - nc) {
            const int i = x - nx + nc;
            hy = c.hy + chez * dEx / CP(1, i);
            c.phx = CP(7, i) * c.phx + CP(6, i) * dEx;
            hy = hy + S.chy_d * c.phx;
        }
        if (hy_in && y < nc) {
            const int i = nc - 1 - y;
            hx = c.hx - chez * dEy / CP(1, i);
            c.phy = CP(7, i) * c.phy + CP(6, i) * dEy;
            hx = hx - S.chx_d * c.phy;
        }
        if (hy_in && y >= ny - nc) {
            const int i = y - ny + nc;
            hx = c.hx - chez * dEy / CP(1, i);
            c.phy = CP(7, i) * c.phy + CP(6, i) * dEy;
            hx = hx - S.chx_d * c.phy;
        }
        if (!(S.flags & TOMO_FLAG_NO_NAN_CLAMP)) {
            if (__builtin_isnan(hx)) hx = 0.f;
            if (__builtin_isnan(hy)) hy = 0.f;
        }
    } else {
        float ty = dEx, tx = dEy;
        if (hx_in && (x < nc || x >= nx - nc)) {
            const int i = x < nc ? nc - 1 - x : x - nx + nc;
            c.phx = CP(7, i) * c.phx + CP(6, i) * dEx;
            ty = dEx * CP(3, i) + S.dxf * c.phx;
        }
        if (hy_in && (y < nc || y >= ny - nc)) {
            const int i = y < nc ? nc - 1 - y : y - ny + nc;
            c.phy = CP(7, i) * c.phy + CP(6, i) * dEy;
            tx = dEy * CP(3, i) + S.dyf * c.phy;
        }
        hy = c.hy + S.chy * ty;
        hx = c.hx - S.chx * tx;
    }
    if (!hx_in || !hy_in) {
        hx = 0.f;
        hy = 0.f;
    }
    // ---------------- E half-step (E_field_update, FDTD_GPU.cu:358-434)
    const float hy_l = __shfl_up(hy, 1);
    const float dHy = hy - hy_l, dHx = hx - hx_prev;
    hx_prev = hx;
    c.hx = hx;
    c.hy = hy;
    if (x <= 0 || x >= nx || y <= 0 || y >= ny) {
        c.ez = 0.f;
        return;
    }
    const float* q = lcoef + c.mat * TOMO_NCOEF;
    const float ez = c.ez;
    if (MODE == TOMO_MODE_REF_COMPAT) {
        float f1 = 0.f, f0 = 0.f;
        if (q[2] != 0.f || q[3] != 0.f || q[6] != 0.f) {
            const float X = S.inc_abs ? (float)x : (float)(x - nx / 2);
            const float Y = S.inc_abs ? (float)y : (float)(y - ny / 2);
            const float sp = S.inc_c - S.inc_ay * Y - S.inc_ax * X;
            const float a1 = (((float)n + 0.5f) * S.inc_dt + sp) * S.inc_itau;
            const float a0 = (((float)n - 0.5f) * S.inc_dt + sp) * S.inc_itau;
            f1 = expf(-a1 * a1);
            f0 = expf(-a0 * a0);
        }
        float buf = q[0] * ez + q[1] * dHy - q[1] * dHx + q[2] * f1 + q[3] * f0 - q[4] * c.jp;
        if (!(S.flags & TOMO_FLAG_CPU_COEFS)) c.jp = q[5] * c.jp + q[6] * (buf - ez + f1 - f0);
        if (x <= nc) {
            const int i = nc - x;
            buf = q[0] * ez + q[1] * dHy / CP(0, i) - q[1] * dHx / CP(0, i);
            c.pex = CP(5, i) * c.pex + CP(4, i) * dHy;
            buf = buf + q[1] * S.dxf * c.pex;
        }
        if (x >= nx - nc) {
            const int i = x - nx + nc;
            buf = q[0] * ez + q[1] * dHy / CP(0, i) - q[1] * dHx / CP(0, i);
            c.pex = CP(5, i) * c.pex + CP(4, i) * dHy;
            buf = buf + q[1] * S.dxf * c.pex;
        }
        if (y <= nc) {
            const int i = nc - y;
            buf = q[0] * ez + q[1] * dHy / CP(0, i) - q[1] * dHx / CP(0, i);
            c.pey = CP(5, i) * c.pey + CP(4, i) * dHx;
            buf = buf - q[1] * S.dyf * c.pey;
        }
        if (y >= ny - nc) {
            const int i = y - ny + nc;
            buf = q[0] * ez + q[1] * dHy / CP(0, i) - q[1] * dHx / CP(0, i);
            c.pey = CP(5, i) * c.pey + CP(4, i) * dHx;
            buf = buf - q[1] * S.dyf * c.pey;
        }
        if (!(S.flags & TOMO_FLAG_NO_NAN_CLAMP) && __builtin_isnan(buf)) buf = 0.f;
        c.ez = buf;
    } else {
        float cy = dHy, cx = dHx;
        if (x <= nc || x >= nx - nc) {
            const int i = x <= nc ? nc - x : x - nx + nc;
            c.pex = CP(5, i) * c.pex + CP(4, i) * dHy;
            cy = dHy * CP(2, i) + S.dxf * c.pex;
        }
        if (y <= nc || y >= ny - nc) {
            const int i = y <= nc ? nc - y : y - ny + nc;
            c.pey = CP(5, i) * c.pey + CP(4, i) * dHx;
            cx = dHx * CP(2, i) + S.dyf * c.pey;
        }
        float en = q[0] * ez + q[1] * cy - q[1] * cx - q[4] * c.jp;
        if (x == S.sx && y == S.sy) en += S.srcw[n];
        c.jp = q[5] * c.jp + q[6] * (en - ez);
        c.ez = en;
    }
}

// Surface samples of the NTFF contour (contour map of FDTD_common.cxx:4-46):
// plane 0 = H on the face cell, 1 = Ez on the face cell, 2 = Ez on the cell
// across the face.  The running DFT averages planes 1 and 2.
__device__ __forceinline__ void write_samples(const SimView& S, const Cell& c, int x, int y, int k) {
    const int D = S.D, nx = S.nx, ny = S.ny, P = S.P;
    float* h_ = S.samp + (size_t)k * P;
    float* e0 = S.samp + (size_t)(S.K + k) * P;
    float* e1 = S.samp + (size_t)(2 * S.K + k) * P;
    const bool xr = (x >= D + 1 && x <= S.cx_hi);
    if (y == D && xr) {
        const int p = x - D - 1;
        h_[p] = c.hx;
        e0[p] = c.ez;
    }
    if (y == ny - D - 1 && xr) {
        const int p = S.L2 + (S.cx_hi - x);
        h_[p] = c.hx;
        e0[p] = c.ez;
    }
    if (y == D + 1 && xr) e1[x - D - 1] = c.ez;
    if (y == ny - D && xr) e1[S.L2 + (S.cx_hi - x)] = c.ez;
    if (y >= S.cy_lo && y <= ny - D - 1) {
        if (x == nx - D - 1) {
            const int p = S.L1 + y - S.cy_lo;
            h_[p] = c.hy;
            e0[p] = c.ez;
        }
        if (x == nx - D) e1[S.L1 + y - S.cy_lo] = c.ez;
        if (x == D) {
            const int p = S.L3 + (ny - D - 1 - y);
            h_[p] = c.hy;
            e0[p] = c.ez;
        }
        if (x == D + 1) e1[S.L3 + (ny - D - 1 - y)] = c.ez;
    }
}

template <int T, int MODE>
__global__ __launch_bounds__(512) void k_tblock(const TileItem* __restrict__ tiles, int ntiles,
                                                const LaunchItem* __restrict__ items,
                                                const DevSim* __restrict__ sims, int ordinal, int seg_rows) {
    __shared__ float lcoef[TOMO_MAX_MAT * TOMO_NCOEF];
    __shared__ float lcp[8 * TOMO_MAX_NC];
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    constexpr int V = kWave - 2 * T;
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const TileItem tile = tiles[ti];
        const LaunchItem it = items[tile.item];
        const int nst = min(T, it.nsteps - ordinal * T);
        if (nst <= 0) continue;  // uniform over the workgroup
        const DevSim* Sp = sims + it.sim;
        __syncthreads();
        for (int i = threadIdx.x; i < Sp->nmat * TOMO_NCOEF; i += blockDim.x) lcoef[i] = Sp->coef[i];
        for (int i = threadIdx.x; i < Sp->nc; i += blockDim.x) {
            CP(0, i) = Sp->ke[i]; CP(1, i) = Sp->km[i]; CP(2, i) = Sp->ike[i]; CP(3, i) = Sp->ikm[i];
            CP(4, i) = Sp->ae[i]; CP(5, i) = Sp->be[i]; CP(6, i) = Sp->am[i]; CP(7, i) = Sp->bm[i];
        }
        __syncthreads();
        const int par = it.parity0 ^ (ordinal & 1);
        SimView S;
        S.ez = Sp->ez[par]; S.hx = Sp->hx[par]; S.hy = Sp->hy[par]; S.jp = Sp->jp[par];
        S.pex = Sp->pex[par]; S.phx = Sp->phx[par]; S.pey = Sp->pey[par]; S.phy = Sp->phy[par];
        S.oez = Sp->ez[par ^ 1]; S.ohx = Sp->hx[par ^ 1]; S.ohy = Sp->hy[par ^ 1]; S.ojp = Sp->jp[par ^ 1];
        S.opex = Sp->pex[par ^ 1]; S.ophx = Sp->phx[par ^ 1]; S.opey = Sp->pey[par ^ 1]; S.ophy = Sp->phy[par ^ 1];
        S.mat = Sp->mat; S.span = Sp->span; S.samp = Sp->samp; S.srcw = Sp->srcw;
        S.nx = Sp->nx; S.ny = Sp->ny; S.pitch = Sp->pitch; S.nc = Sp->nc; S.D = Sp->D; S.P = Sp->P;
        S.L1 = Sp->L1; S.L2 = Sp->L2; S.L3 = Sp->L3; S.cx_hi = Sp->cx_hi; S.cy_lo = Sp->cy_lo; S.sx = Sp->sx; S.sy = Sp->sy; S.flags = Sp->flags;
        S.K = Sp->chunk;
        S.chx = Sp->chx; S.chy = Sp->chy; S.chx_d = Sp->chx_d; S.chy_d = Sp->chy_d; S.dxf = Sp->dxf;
        S.dyf = Sp->dyf; S.inc_dt = Sp->inc_dt; S.inc_c = Sp->inc_c; S.inc_ax = Sp->inc_ax;
        S.inc_ay = Sp->inc_ay; S.inc_itau = Sp->inc_itau; S.inc_abs = Sp->inc_abs_origin;
        const int nbands = (S.nx + 1 + V - 1) / V;
        const int band = tile.band0 + wv;
        if (band < nbands) {
            const int x = band * V - T + lane;
            const bool xin = (x >= 0 && x <= S.nx);
            const bool xown = xin && lane >= T && lane < kWave - T;
            const int y0 = tile.seg * seg_rows;
            const int y1 = min(y0 + seg_rows, S.ny + 1);
            const int step0 = it.step0 + ordinal * T;
            const int k0 = ordinal * T;
            Cell cur[T];
            float hxp[T];
#pragma unroll
            for (int t = 0; t < T; ++t) {
                cur[t] = zero_cell();
                hxp[t] = 0.f;
            }
            Cell nxt;
            load_cell(S, lcoef, y0 - T, x, xin, nxt);
            for (int r = y0 - T; r < y1 + T; ++r) {
                Cell in = nxt;
                if (r + 1 < y1 + T) load_cell(S, lcoef, r + 1, x, xin, nxt);
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    const int y = r - 1 - t;
                    Cell out = cur[t];
                    if (t < nst) {
                        stage_update<MODE>(S, lcoef, lcp, out, in.ez, hxp[t], x, y, step0 + t);
                        if (xown && y >= y0 && y < y1) write_samples(S, out, x, y, k0 + t);
                    }
                    cur[t] = in;
                    in = out;
                }
                const int yo = r - T;
                if (xown && yo >= y0 && yo < y1) store_cell(S, lcoef, yo, x, in);
            }
        }
    }
}
#undef CP

// ---------------------------------------------------------------- running DFT fold
// Phase e^{-j w (step + t0) dt}: reseeded in FP64 at the chunk start, advanced
// by an FP32 rotation inside the chunk (<= 64 steps), summed in FP32 over the
// chunk and added to the long-running accumulator with Kahan compensation.
template <bool FP64>
__global__ __launch_bounds__(256) void k_fold(const LaunchItem* __restrict__ items,
                                              const DevSim* __restrict__ sims) {
    const LaunchItem it = items[blockIdx.y];
    const DevSim* S = sims + it.sim;
    const int P = S->P, F = S->nfreq, K = S->chunk;
    const int nk = min(it.nsteps, K);
    if (nk <= 0 || ((S->flags & TOMO_FLAG_DFT_FP64) != 0) != FP64) return;
    __shared__ float2 wJ[FP64 ? 1 : TOMO_MAX_FREQ * 64];
    __shared__ float2 wM[FP64 ? 1 : TOMO_MAX_FREQ * 64];
    __shared__ double2 dJ[FP64 ? TOMO_MAX_FREQ * 64 : 1];
    __shared__ double2 dRot[FP64 ? TOMO_MAX_FREQ : 1];
    if (!FP64) {
        if (threadIdx.x < 2 * F) {
            const int f = threadIdx.x >> 1, which = threadIdx.x & 1;
            const double w = S->wdt[f];
            const double th0 = -w * ((double)it.step0 + (which ? S->tM : S->tJ));
            double s0, c0, sr, cr;
            sincos(th0, &s0, &c0);
            sincos(-w, &sr, &cr);
            float zr = (float)c0, zi = (float)s0;
            const float rr = (float)cr, ri = (float)sr;
            float2* dst = which ? wM : wJ;
            for (int k = 0; k < nk; ++k) {
                dst[f * 64 + k] = make_float2(zr, zi);
                const float nr = zr * rr - zi * ri;
                const float ni = zr * ri + zi * rr;
                zr = nr;
                zi = ni;
            }
        }
    } else {
        // FP64 study mode: exact phases; the M phase is the J phase rotated by
        // -w (tM - tJ).
        for (int i = threadIdx.x; i < F * nk; i += blockDim.x) {
            const int f = i / nk, k = i % nk;
            double s, c;
            sincos(-S->wdt[f] * ((double)it.step0 + k + S->tJ), &s, &c);
            dJ[f * 64 + k] = make_double2(c, s);
        }
        if (threadIdx.x < F) {
            double s, c;
            sincos(-S->wdt[threadIdx.x] * (S->tM - S->tJ), &s, &c);
            dRot[threadIdx.x] = make_double2(c, s);
        }
    }
    __syncthreads();
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int leg = p < S->L1 ? 0 : (p < S->L2 ? 1 : (p < S->L3 ? 2 : 3));
    const float sdt = S->sgn_dt[leg];
    const float* hs = S->samp + p;
    const float* e0 = S->samp + (size_t)K * P + p;
    const float* e1 = S->samp + (size_t)2 * K * P + p;
    for (int f = 0; f < F; ++f) {
        if (!FP64) {
            float jr = 0.f, ji = 0.f, mr = 0.f, mi = 0.f;
            for (int k = 0; k < nk; ++k) {
                const float h = hs[(size_t)k * P];
                const float e = (e0[(size_t)k * P] + e1[(size_t)k * P]) / 2;
                const float2 a = wJ[f * 64 + k], b = wM[f * 64 + k];
                jr += h * a.x;
                ji += h * a.y;
                mr += e * b.x;
                mi += e * b.y;
            }
            float4* aJ = (float4*)S->accJ + (size_t)f * P + p;
            float4* aM = (float4*)S->accM + (size_t)f * P + p;
            float4 vJ = *aJ, vM = *aM;
            // Kahan: value (x, y), compensation (z, w)
            float yv, tv;
            yv = sdt * jr - vJ.z; tv = vJ.x + yv; vJ.z = (tv - vJ.x) - yv; vJ.x = tv;
            yv = sdt * ji - vJ.w; tv = vJ.y + yv; vJ.w = (tv - vJ.y) - yv; vJ.y = tv;
            yv = sdt * mr - vM.z; tv = vM.x + yv; vM.z = (tv - vM.x) - yv; vM.x = tv;
            yv = sdt * mi - vM.w; tv = vM.y + yv; vM.w = (tv - vM.y) - yv; vM.y = tv;
            *aJ = vJ;
            *aM = vM;
        } else {
            double jr = 0, ji = 0, mr = 0, mi = 0;
            for (int k = 0; k < nk; ++k) {
                const double h = hs[(size_t)k * P];
                const double e = 0.5 * ((double)e0[(size_t)k * P] + (double)e1[(size_t)k * P]);
                const double2 a = dJ[f * 64 + k], r = dRot[f];
                const double2 b = make_double2(a.x * r.x - a.y * r.y, a.x * r.y + a.y * r.x);
                jr += h * a.x;
                ji += h * a.y;
                mr += e * b.x;
                mi += e * b.y;
            }
            double2* aJ = (double2*)S->accJ + (size_t)f * P + p;
            double2* aM = (double2*)S->accM + (size_t)f * P + p;
            double2 vJ = *aJ, vM = *aM;
            vJ.x += sdt * jr; vJ.y += sdt * ji; vM.x += sdt * mr; vM.y += sdt * mi;
            *aJ = vJ;
            *aM = vM;
        }
    }
}

// ---------------------------------------------------------------- far-field projection
// N(phi) = sum_p -del_p J_p w,  L(phi) = -sin(phi) sum_{y faces} dx M_p w + cos(phi) sum_{x faces} dy M_p w,
// w = exp(j k (x_p cos phi + y_p sin phi)),  D = k^2 |L + eta0 N|^2 / norm,  P = arg(L + eta0 N).
constexpr int kProjChunk = 256;

__global__ __launch_bounds__(128) void k_project(const ProjItem* __restrict__ items,
                                                 const DevSim* __restrict__ sims, int A, float eta0,
                                                 float* __restrict__ D, float* __restrict__ Pph) {
    const ProjItem it = items[blockIdx.y];
    const DevSim* S = sims + it.sim;
    const int P = S->P;
    const bool fp64 = (S->flags & TOMO_FLAG_DFT_FP64) != 0;
    __shared__ float sx[kProjChunk], sy[kProjChunk], jr[kProjChunk], ji[kProjChunk], mr[kProjChunk],
        mi[kProjChunk], wt[kProjChunk];
    __shared__ int sleg[kProjChunk];
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    float cph = 0.f, sph = 0.f;
    if (a < A) sincosf(6.283185307179586f * (float)a / (float)A, &sph, &cph);
    float Nr = 0.f, Ni = 0.f, Lyr = 0.f, Lyi = 0.f, Lxr = 0.f, Lxi = 0.f;
    const float k = it.k;
    for (int p0 = 0; p0 < P; p0 += kProjChunk) {
        __syncthreads();
        for (int i = threadIdx.x; i < kProjChunk; i += blockDim.x) {
            const int p = p0 + i;
            if (p < P) {
                sx[i] = S->pos[p];
                sy[i] = S->pos[P + p];
                if (fp64) {
                    const double2 vj = ((const double2*)S->accJ)[(size_t)it.f * P + p];
                    const double2 vm = ((const double2*)S->accM)[(size_t)it.f * P + p];
                    jr[i] = (float)vj.x; ji[i] = (float)vj.y; mr[i] = (float)vm.x; mi[i] = (float)vm.y;
                } else {
                    const float4 vj = ((const float4*)S->accJ)[(size_t)it.f * P + p];
                    const float4 vm = ((const float4*)S->accM)[(size_t)it.f * P + p];
                    jr[i] = vj.x; ji[i] = vj.y; mr[i] = vm.x; mi[i] = vm.y;
                }
                const int leg = p < S->L1 ? 0 : (p < S->L2 ? 1 : (p < S->L3 ? 2 : 3));
                sleg[i] = (leg == 0 || leg == 2) ? 0 : 1;
                wt[i] = (leg == 0 || leg == 2) ? S->dxf : S->dyf;
            } else {
                sx[i] = sy[i] = jr[i] = ji[i] = mr[i] = mi[i] = wt[i] = 0.f;
                sleg[i] = 0;
            }
        }
        __syncthreads();
        const int n = min(kProjChunk, P - p0);
        for (int i = 0; i < n; ++i) {
            float s, c;
            sincosf(k * (sx[i] * cph + sy[i] * sph), &s, &c);
            const float del = wt[i];
            // N += -del * J * w
            Nr -= del * (jr[i] * c - ji[i] * s);
            Ni -= del * (jr[i] * s + ji[i] * c);
            const float tr = del * (mr[i] * c - mi[i] * s), tiv = del * (mr[i] * s + mi[i] * c);
            if (sleg[i] == 0) { Lyr += tr; Lyi += tiv; }
            else { Lxr += tr; Lxi += tiv; }
        }
    }
    if (a < A) {
        const float Lr = -sph * Lyr + cph * Lxr, Li = -sph * Lyi + cph * Lxi;
        const float Er = Lr + eta0 * Nr, Ei = Li + eta0 * Ni;
        D[it.out + a] = k * k * (Er * Er + Ei * Ei) * it.inv_norm;
        if (Pph) Pph[it.out + a] = atan2f(Ei, Er);
    }
}

// ---------------------------------------------------------------- fitness (FP64)
__global__ __launch_bounds__(256) void k_fitness(const float* __restrict__ D, const float* __restrict__ Pph,
                                                 const float* __restrict__ tgt, const float* __restrict__ tph,
                                                 int L, double norm, int phase_term, double* __restrict__ fit) {
    const int e = blockIdx.x;
    const bool rel = (phase_term & 2) != 0;
    double acc = 0.0;
    for (int i = threadIdx.x; i < L; i += blockDim.x) {
        const double m = (double)tgt[i];
        const double d = (double)D[(size_t)e * L + i] - m;
        acc += rel ? d * d / (m * m * norm) : d * d / norm;
        if ((phase_term & 1) && Pph && tph) {
            double dp = (double)Pph[(size_t)e * L + i] - (double)tph[i];
            const double pi = 3.141592653589793238;
            if (dp > pi) dp -= 2 * pi;
            else if (dp < -pi) dp += 2 * pi;
            acc += dp * dp / (10 * pi * pi * norm);
        }
    }
    for (int off = kWave / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    __shared__ double part[256 / kWave];
    if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < (int)(blockDim.x / kWave); ++w) s += part[w];
        fit[e] = -s;
    }
}

// ---------------------------------------------------------------- rasterisers
__global__ void k_raster_ref(uint8_t* __restrict__ mat, int pitch, int nx, int ny, const int* __restrict__ lo,
                             const int* __restrict__ hi, const int* __restrict__ xc, const int* __restrict__ yc,
                             int npatch) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= pitch || y > ny) return;
    uint8_t m = 0;
    if (x <= nx && x >= lo[y] && x <= hi[y]) m = (uint8_t)(1 + xc[x] + npatch * yc[y]);
    mat[(size_t)y * pitch + x] = m;
}

__device__ __forceinline__ bool in_circle(float X, float Y, float cx, float cy, float r2) {
#pragma clang fp contract(off)
    const float dX = X - cx, dY = Y - cy;
    const float s = dX * dX + dY * dY;
    return s <= r2;
}

__device__ __forceinline__ bool in_ellipse(float X, float Y, const float* e) {
#pragma clang fp contract(off)
    const float dX = X - e[0], dY = Y - e[1];
    const float u = dX * e[2] + dY * e[3];
    const float v = dY * e[2] - dX * e[3];
    const float q = u * u * e[4] + v * v * e[5];
    return q <= 1.0f;
}

__global__ void k_raster_phys(uint8_t* __restrict__ mat, int pitch, int nx, int ny, float dxf, float dyf,
                              PhantomGeom g) {
#pragma clang fp contract(off)
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= pitch || y > ny) return;
    uint8_t m = 0;
    if (x <= nx) {
        const float X = (float)(x - nx / 2) * dxf;
        const float Y = (float)(y - ny / 2) * dyf;
        if (in_circle(X, Y, 0.f, 0.f, g.ro2)) {
            m = 1;
            if (in_circle(X, Y, 0.f, 0.f, g.ri2)) {
                m = 2;
                for (int e = 0; e < 3; ++e)
                    if (in_ellipse(X, Y, g.ell[e])) m = 3;
                for (int k = 0; k < g.ntum; ++k)
                    if (in_circle(X, Y, g.tum[k][0], g.tum[k][1], g.tum[k][2])) m = (uint8_t)(4 + k);
            }
        }
    }
    mat[(size_t)y * pitch + x] = m;
}

__global__ void k_zero(uint4* __restrict__ p, size_t n16) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0, 0, 0, 0);
}

}  // namespace

bool tblock_supported(int T) { return T == 1 || T == 2 || T == 4 || T == 6 || T == 8; }

hipError_t launch_tblock(int T, int mode, const TileItem* tiles, int ntiles, const LaunchItem* items,
                         const DevSim* sims, int ordinal, int seg_rows, int waves, int grid, hipStream_t st) {
    dim3 g(grid), b(waves * kWave);
#define TB(TT)                                                                                                 \
    if (T == TT) {                                                                                             \
        if (mode == TOMO_MODE_REF_COMPAT)                                                                      \
            hipLaunchKernelGGL((k_tblock<TT, TOMO_MODE_REF_COMPAT>), g, b, 0, st, tiles, ntiles, items, sims,   \
                               ordinal, seg_rows);                                                             \
        else                                                                                                   \
            hipLaunchKernelGGL((k_tblock<TT, TOMO_MODE_PHYSICAL>), g, b, 0, st, tiles, ntiles, items, sims,     \
                               ordinal, seg_rows);                                                             \
        return hipGetLastError();                                                                              \
    }
    TB(1) TB(2) TB(4) TB(6) TB(8)
#undef TB
    return hipErrorInvalidValue;
}

hipError_t launch_fold(bool fp64, const LaunchItem* items, int nitems, const DevSim* sims, int max_p,
                       hipStream_t st) {
    dim3 g((max_p + 255) / 256, nitems), b(256);
    if (fp64) hipLaunchKernelGGL((k_fold<true>), g, b, 0, st, items, sims);
    else hipLaunchKernelGGL((k_fold<false>), g, b, 0, st, items, sims);
    return hipGetLastError();
}

hipError_t launch_project(const ProjItem* items, int nitems, const DevSim* sims, const float* const* pos,
                          int n_angles, float eta0, float* D, float* Pph, hipStream_t st) {
    (void)pos;
    dim3 g((n_angles + 127) / 128, nitems), b(128);
    hipLaunchKernelGGL(k_project, g, b, 0, st, items, sims, n_angles, eta0, D, Pph);
    return hipGetLastError();
}

hipError_t launch_fitness(const float* D, const float* Pph, const float* tgt, const float* tph, int L, int n_eval,
                          double norm, int phase_term, double* fit, hipStream_t st) {
    hipLaunchKernelGGL(k_fitness, dim3(n_eval), dim3(256), 0, st, D, Pph, tgt, tph, L, norm, phase_term, fit);
    return hipGetLastError();
}

hipError_t launch_raster_ref(uint8_t* mat, int pitch, int nx, int ny, const int* lo, const int* hi, const int* xc,
                             const int* yc, int npatch, hipStream_t st) {
    dim3 g((pitch + 255) / 256, ny + 1), b(256);
    hipLaunchKernelGGL(k_raster_ref, g, b, 0, st, mat, pitch, nx, ny, lo, hi, xc, yc, npatch);
    return hipGetLastError();
}

hipError_t launch_raster_phys(uint8_t* mat, int pitch, int nx, int ny, float dxf, float dyf, const PhantomGeom& g_,
                              hipStream_t st) {
    dim3 g((pitch + 255) / 256, ny + 1), b(256);
    hipLaunchKernelGGL(k_raster_phys, g, b, 0, st, mat, pitch, nx, ny, dxf, dyf, g_);
    return hipGetLastError();
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t st) {
    if (bytes == 0) return hipSuccess;
    if (((uintptr_t)p & 15) || (bytes & 15)) return hipErrorInvalidValue;
    const size_t n16 = bytes / 16;
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, st, (uint4*)p, n16);
    return hipGetLastError();
}

}  // namespace tomo
