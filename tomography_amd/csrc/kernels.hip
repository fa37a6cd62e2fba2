// kernels.hip — gfx950 kernels of the 2D TMz FDTD fitness evaluation.
//
//   k_tblock    time-blocked E/H update (replaces H_field_update +
//               E_field_update + calculate_JandM's sampling, FDTD_GPU.cu:231-434)
//   k_fold      running DFT of one chunk of surface samples (calculate_JandM's
//               accumulation, FDTD_GPU.cu:231-280), Kahan FP32 or FP64
//   k_project_mfma  far-field projection on the matrix cores
//               (N2FPostProcess, FDTD_GPU.cu:553-634)
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
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.h"

namespace tomo {
namespace {

constexpr int kWave = 64;
// Rows per prefetch group (row_stream) of each update kernel.  The interior
// kernel is light enough to hold four rows in flight; the edge kernels
// already run near their register budget, and their throughput in large
// batches depends on occupancy more than on prefetch depth.
constexpr int kGroupInner = 4, kGroupRow = 1, kGroupCol = 1, kGroupEdge = 1, kGroupRef = 2;
// Prefetch groups of the fused kernel (one or a few simulations, latency-bound;
// on 1024^2 with 10-row interior tiles one interior row per group measured
// 1.1% faster than two, and two 0.8% faster than four with 14-row tiles).
constexpr int kFusedGroupInner = 1, kFusedGroupRow = 1, kFusedGroupCol = 1, kFusedGroupEdge = 1;
// Row windows: the per-class kernels read the rows' dispersive spans from a
// per-lane window (SpanWin, 3% faster on a 1024 x 512^2 batch) and so do the
// corner and row-edge stages their y-strip CPML coefficients (YWin; plain
// 1024^2 186 -> 198 Gcell-steps/s, the batch 485 -> 507).  The fused kernel
// keeps scalar span loads: with its 10-row interior tiles the span window
// measured the same (211.8 against 212.6 Gcell-steps/s).
// Window flags of the row-sweep bodies: kWinSpan reads the rows' dispersive
// spans from a SpanWin, kWinY the y-strip CPML coefficients from a YWin.
constexpr int kWinSpan = 1, kWinY = 2;
constexpr int kStageWin = kWinSpan | kWinY, kFusedWin = kWinY;
// s_waitcnt vmcnt(0) with expcnt and lgkmcnt left at their maxima (gfx9 encoding).
constexpr int kWaitVm0 = 0x0F70;

// Streams rows [y0 - T, y1 + T) of one band through a T-stage register
// pipeline: step(r, row r at level n) returns row r - T at level n + T, which
// store() writes when it is an owned output row.  Rows are requested in
// groups of G.  At a group boundary one vmcnt(0) retires this group's loads
// and the previous boundary's stores, both issued a whole group of compute
// earlier; then the previous group's results are stored and the next group's
// rows requested.  (Vector loads and stores share one counter that may retire
// them out of order, so a wait on any load issued before a store drains the
// counter: one wait per group instead of one per row keeps G rows in flight.)
template <int T, int G, class C, class Load, class Store, class Step>
__device__ __forceinline__ void row_stream(int y0, int y1, bool own, Load& load, Store& store, Step& step) {
    const int rb = y0 - T, re = y1 + T;
    C A[G], R[G];
#pragma unroll
    for (int i = 0; i < G; ++i)
        if (rb + i < re) load(rb + i, A[i]);
    for (int g0 = rb; g0 < re; g0 += G) {
        __builtin_amdgcn_s_waitcnt(kWaitVm0);
        C X[G];
#pragma unroll
        for (int i = 0; i < G; ++i) X[i] = A[i];
        if (g0 > rb) {
#pragma unroll
            for (int i = 0; i < G; ++i) {
                const int yo = g0 - G + i - T;
                if (yo >= y0 && yo < y1 && own) store(yo, R[i]);
            }
        }
#pragma unroll
        for (int i = 0; i < G; ++i)
            if (g0 + G + i < re) load(g0 + G + i, A[i]);
#pragma unroll
        for (int i = 0; i < G; ++i) {
            const int r = g0 + i;
            if (r >= re) break;
            R[i] = step(r, X[i]);
        }
    }
    // the last group (rows of a partial group at or past re were not computed)
    const int gl = rb + ((re - rb - 1) / G) * G;
#pragma unroll
    for (int i = 0; i < G; ++i) {
        const int yo = gl + i - T;
        if (gl + i < re && yo >= y0 && yo < y1 && own) store(yo, R[i]);
    }
}

// The same rows, with each row's loads issued P rows before the row is
// computed and no explicit wait: the compiler waits for a row's loads before
// its stage chain uses them.  While only loads are in flight those waits count
// down in issue order and keep P rows in flight; once a store is in flight
// too, a wait drains the counter.  A tile whose output rows (the last
// y1 - y0 of the streamed rows) number at most P issues every load before its
// first store, so its rows' memory latencies overlap instead of adding up:
// the form for short boundary tiles, whose rows each wait a whole memory
// latency in the grouped sweep.
template <int T, int P, class C, class Load, class Store, class Step>
__device__ __forceinline__ void row_ring(int y0, int y1, bool own, Load& load, Store& store, Step& step) {
    const int rb = y0 - T, re = y1 + T;
    C ring[P];
#pragma unroll
    for (int i = 0; i < P; ++i)
        if (rb + i < re) load(rb + i, ring[i]);
    for (int r0 = rb; r0 < re; r0 += P) {
#pragma unroll
        for (int i = 0; i < P; ++i) {
            const int r = r0 + i;
            if (r >= re) break;
            const C X = ring[i];
            if (r + P < re) load(r + P, ring[i]);
            const C R = step(r, X);
            const int yo = r - T;
            if (yo >= y0 && yo < y1 && own) store(yo, R);
        }
    }
}

// Row sweep of a tile: grouped (G > 0 rows per wait) or a load ring of -G rows.
template <int T, int G, class C, class Load, class Store, class Step>
__device__ __forceinline__ void row_sweep(int y0, int y1, bool own, Load& load, Store& store, Step& step) {
    if constexpr (G > 0) row_stream<T, G, C>(y0, y1, own, load, store, step);
    else row_ring<T, -G, C>(y0, y1, own, load, store, step);
}

// Global-address-space views: loads and stores through these compile to
// global_* instructions (not flat_*), so LDS waits never wait on HBM.
#define GAS __attribute__((address_space(1)))
typedef GAS float gfloat;
typedef GAS const float gcfloat;
typedef GAS const uint8_t gcu8;
typedef GAS const int gcint;
// Constant-address-space views of data that is read-only for a whole launch
// (row spans, the source waveform): wave-uniform loads through these become
// scalar loads (s_load_*), which are tracked by lgkmcnt and so never force a
// wait on the vector loads prefetched rows ahead.
#define CAS __attribute__((address_space(4)))
typedef CAS const int kcint;
typedef CAS const float kcfloat;

// Lane shifts on the VALU (DPP wave_shl:1 / wave_shr:1): next(v)[i] = v[i+1],
// prev(v)[i] = v[i-1].  The edge lane keeps a stale value; it is a ghost lane.
__device__ __forceinline__ float lane_next(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xF, 0xF, false));
}
__device__ __forceinline__ float lane_prev(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, false));
}

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

// Wave-uniform view of one simulation for one launch (kept in SGPRs).
struct SimView {
    gfloat* f;     // field planes: ez0 ez1 hx0 hx1 hy0 hy1, stride ps
    gfloat* jp;    // jp0 jp1, stride js
    gfloat* xq;    // pex0 pex1 phx0 phx1, stride xs
    gfloat* yq;    // pey0 pey1 phy0 phy1, stride ys
    gcu8* mat;
    kcint* span;
    gfloat* samp;
    kcfloat* srcw;
    int ps, js, xs, ys, par;
    int nx, ny, pitch, nc, D, P, L1, L2, L3, cx_hi, cy_lo, sx, sy, flags, K;
    float chx, chy, chx_d, chy_d, dxf, dyf;
    float inc_dt, inc_c, inc_ax, inc_ay, inc_itau;
    int inc_abs;
};

// Wave-uniform facts about a band of 64 columns [x0, x0 + 63].
struct WaveFlags {
    bool xs;     // holds CPML x-strip cells
    bool ccol;   // holds an NTFF contour column or its neighbour
    bool src;    // holds the source column
};

__device__ __forceinline__ bool row_ystrip_e(const SimView& S, int r) {
    return (r >= 1 && r <= S.nc) || (r >= S.ny - S.nc && r <= S.ny - 1);
}
__device__ __forceinline__ bool row_ystrip_h(const SimView& S, int r) {
    return (r >= 0 && r < S.nc) || (r >= S.ny - S.nc && r < S.ny);
}

// Dispersive-region span (first column, end column, offset in the compact
// jp array) of one row.
struct Span {
    int lo, hi, off;
};

// The spans of the rows a wave streams, one row per lane: lane i holds row
// rb + i.  The window is one vector load per tile; the row loop then reads a
// row's span with v_readlane, where a scalar load per row would make every row
// wait for memory (a scalar load's wait also waits for every LDS read in
// flight).  A tile streams at most 64 rows (its output rows plus T ghost rows
// on each side: the tile builders keep y1 - y0 <= 64 - 2T).
struct SpanWin {
    int lo, hi, off;
    int rb;
};

__device__ __forceinline__ SpanWin span_window(const int* __restrict__ span, int rb, int ny) {
    SpanWin w;
    const int r = rb + (int)(threadIdx.x & (kWave - 1));
    w.lo = w.hi = w.off = 0;
    if (r >= 0 && r <= ny) {
        const gcint* p = (const gcint*)span + 3 * r;
        w.lo = p[0];
        w.hi = p[1];
        w.off = p[2];
    }
    w.rb = rb;
    return w;
}

// Span of row r (wave-uniform, 0 <= r <= ny, 0 <= r - rb < 64).
__device__ __forceinline__ Span span_row(const SpanWin& w, int r) {
    const int i = r - w.rb;
    Span s;
    s.lo = __builtin_amdgcn_readlane(w.lo, i);
    s.hi = __builtin_amdgcn_readlane(w.hi, i);
    s.off = __builtin_amdgcn_readlane(w.off, i);
    return s;
}

// The y-strip CPML coefficients of the rows a wave streams, one row per lane
// (lane i: row rb + i): ike, ae, be of the E row index and ikm, am, bm of the
// H row index (identity off the strips), read once per tile from the LDS
// table t[k * TOMO_MAX_NC + i] (k = 0..5: ike, ikm, ae, be, am, bm).  A stage
// takes its row's six with v_readlane, where it would otherwise make six LDS
// reads at a row-dependent address and wait for them.
struct YWin {
    float ke, ae, be, kh, ah, bh;
    int rb;
};

__device__ __forceinline__ YWin ywin_load(const float* t, int rb, int ny, int nc) {
    constexpr int ID = TOMO_MAX_NC - 1;
    const int y = rb + (int)(threadIdx.x & (kWave - 1));
    const int iye = (y >= 1 && y <= nc) ? nc - y : ((y >= ny - nc && y <= ny - 1) ? y - ny + nc : ID);
    const int iyh = (y >= 0 && y < nc) ? nc - 1 - y : ((y >= ny - nc && y < ny) ? y - ny + nc : ID);
    YWin w;
    w.ke = t[iye];
    w.kh = t[TOMO_MAX_NC + iyh];
    w.ae = t[2 * TOMO_MAX_NC + iye];
    w.be = t[3 * TOMO_MAX_NC + iye];
    w.ah = t[4 * TOMO_MAX_NC + iyh];
    w.bh = t[5 * TOMO_MAX_NC + iyh];
    w.rb = rb;
    return w;
}

struct YCoef {
    float ke, ae, be, kh, ah, bh;
};

__device__ __forceinline__ float lane_read(float v, int i) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), i));
}

// Coefficients of row y.  Rows below the window are the sweep's warm-up rows
// (stage t of the first t rows), whose results never reach an output row; the
// lane index wraps for them.
__device__ __forceinline__ YCoef ycoef(const YWin& w, int y) {
    const int i = (y - w.rb) & (kWave - 1);
    return YCoef{lane_read(w.ke, i), lane_read(w.ae, i), lane_read(w.be, i),
                 lane_read(w.kh, i), lane_read(w.ah, i), lane_read(w.bh, i)};
}

// Element access of load_cell / store_cell.  COH: relaxed atomics at agent
// scope (coherent across the XCDs' L2 caches without a cache flush), for data
// that another workgroup of the same launch wrote or will read.
template <bool COH>
__device__ __forceinline__ float ld_(gfloat* p) {
    if constexpr (COH) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *p;
}
template <bool COH>
__device__ __forceinline__ void st_(gfloat* p, float v) {
    if constexpr (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

template <bool COH = false>
__device__ __forceinline__ void load_cell(const SimView& S, const float4* __restrict__ lc4, int r, int x, bool xin,
                                          const WaveFlags& wf, Cell& c, const SpanWin* sw = nullptr) {
    c = zero_cell();
    // The material id is read unconditionally from the clamped node: a byte
    // loaded under a branch is zero-extended right after the branch, which
    // would force a wait on every load in flight.  Outside the grid the id
    // only selects coefficients of a cell whose result is discarded.
    c.mat = S.mat[min(max(r, 0), S.ny) * S.pitch + min(max(x, 0), S.nx)];
    if (r < 0 || r > S.ny) return;  // wave-uniform
    if (!xin) return;
    const int base = r * S.pitch + x;
    const int in_f = S.par * S.ps;
    c.ez = ld_<COH>(S.f + in_f + base);
    if (x < S.nx && r < S.ny) {
        c.hx = ld_<COH>(S.f + in_f + 2 * S.ps + base);
        c.hy = ld_<COH>(S.f + in_f + 4 * S.ps + base);
    }
    if (sw) {
        const Span sp = span_row(*sw, r);
        if (x >= sp.lo && x < sp.hi) c.jp = ld_<COH>(S.jp + S.par * S.js + sp.off + x - sp.lo);
    } else {
        const int lo = S.span[3 * r], hi = S.span[3 * r + 1];
        if (hi > lo && x >= lo && x < hi)
            c.jp = ld_<COH>(S.jp + S.par * S.js + S.span[3 * r + 2] + x - lo);
    }
    const int nc = S.nc, nx = S.nx, ny = S.ny, w = 2 * nc;
    if (wf.xs) {
        gfloat* pe = S.xq + S.par * S.xs;
        gfloat* ph = S.xq + (2 + S.par) * S.xs;
        if (x >= 1 && x <= nc) c.pex = ld_<COH>(pe + r * w + x - 1);
        else if (x >= nx - nc && x <= nx - 1) c.pex = ld_<COH>(pe + r * w + nc + x - (nx - nc));
        if (x < nc) c.phx = ld_<COH>(ph + r * w + x);
        else if (x >= nx - nc && x < nx) c.phx = ld_<COH>(ph + r * w + nc + x - (nx - nc));
    }
    if (row_ystrip_e(S, r)) {
        const int j = r <= nc ? r - 1 : nc + r - (ny - nc);
        c.pey = ld_<COH>(S.yq + S.par * S.ys + j * S.pitch + x);
    }
    if (row_ystrip_h(S, r)) {
        const int j = r < nc ? r : nc + r - (ny - nc);
        c.phy = ld_<COH>(S.yq + (2 + S.par) * S.ys + j * S.pitch + x);
    }
}

template <bool COH = false>
__device__ __forceinline__ void store_cell(const SimView& S, const float4* __restrict__ lc4, int r, int x,
                                           const WaveFlags& wf, const Cell& c, const SpanWin* sw = nullptr) {
    const int base = r * S.pitch + x;
    const int out_f = (S.par ^ 1) * S.ps;
    st_<COH>(S.f + out_f + base, c.ez);
    if (x < S.nx && r < S.ny) {
        st_<COH>(S.f + out_f + 2 * S.ps + base, c.hx);
        st_<COH>(S.f + out_f + 4 * S.ps + base, c.hy);
    }
    if (sw) {
        const Span sp = span_row(*sw, r);
        if (x >= sp.lo && x < sp.hi) st_<COH>(S.jp + (S.par ^ 1) * S.js + sp.off + x - sp.lo, c.jp);
    } else {
        const int lo = S.span[3 * r], hi = S.span[3 * r + 1];
        if (hi > lo && x >= lo && x < hi)
            st_<COH>(S.jp + (S.par ^ 1) * S.js + S.span[3 * r + 2] + x - lo, c.jp);
    }
    const int nc = S.nc, nx = S.nx, ny = S.ny, w = 2 * nc;
    if (wf.xs) {
        gfloat* pe = S.xq + (S.par ^ 1) * S.xs;
        gfloat* ph = S.xq + (2 + (S.par ^ 1)) * S.xs;
        if (x >= 1 && x <= nc) st_<COH>(pe + r * w + x - 1, c.pex);
        else if (x >= nx - nc && x <= nx - 1) st_<COH>(pe + r * w + nc + x - (nx - nc), c.pex);
        if (x < nc) st_<COH>(ph + r * w + x, c.phx);
        else if (x >= nx - nc && x < nx) st_<COH>(ph + r * w + nc + x - (nx - nc), c.phx);
    }
    if (row_ystrip_e(S, r)) {
        const int j = r <= nc ? r - 1 : nc + r - (ny - nc);
        st_<COH>(S.yq + (S.par ^ 1) * S.ys + j * S.pitch + x, c.pey);
    }
    if (row_ystrip_h(S, r)) {
        const int j = r < nc ? r : nc + r - (ny - nc);
        st_<COH>(S.yq + (2 + (S.par ^ 1)) * S.ys + j * S.pitch + x, c.phy);
    }
}

// LDS CPML table layout: [0] ke [1] km [2] ike [3] ikm [4] ae [5] be [6] am [7] bm, each TOMO_MAX_NC.
#define CP(k, i) lcp[(k) * TOMO_MAX_NC + (i)]

// Packed y-strip tables of the physical-mode row and corner tiles: E entries
// {1/kappa_e, a_e, b_e, exists} at [0, kYTab) and H entries {1/kappa_m, a_m,
// b_m, exists} at [kYTab, 2 kYTab).  Entry i < nc is strip cell i; kYID is
// the identity of rows outside the strips and kYIDZ the identity of rows
// without nodes of that field (PEC rows, rows outside the grid).  One 16-byte
// LDS read per field and stage, and the node masks arrive as data, so the
// stage is branch-free.
constexpr int kYID = TOMO_MAX_NC - 1, kYIDZ = TOMO_MAX_NC, kYTab = TOMO_MAX_NC + 1;

// Stage form of the row and corner tiles.  Deep time blocks (T = 8, the
// single-simulation schedule) leave few waves per SIMD, so a tile's serial
// latency sets the launch time: their stages use the packed tables and
// arithmetic node masks, and write NTFF samples after the stage loop, so each
// row step is one basic block whose LDS reads overlap across stages (corner
// and row-edge tiles alone: 58 -> 42 us per 1024^2 launch).  Shallower
// blocks (the batched schedule, throughput-bound) keep the select form,
// which issues fewer instructions.
// kStageBranchFree: the form of the standalone row and corner kernels;
// kFusedBranchFree: the form inside the fused kernel.
template <int T>
constexpr bool kStageBranchFree = T >= 8;
template <int T>
constexpr bool kFusedBranchFree = T >= 8;
// Minimum resident workgroups per CU of the fused kernel, which caps its
// VGPRs: two for the branch-free corner stage (three would spill it to
// scratch), three otherwise.
template <int T>
constexpr int kFusedWgPerCu = kFusedBranchFree<T> ? 2 : 3;

__device__ __forceinline__ void fill_ytab(const DevSim* __restrict__ Sp, float4* __restrict__ ly) {
    for (int i = threadIdx.x; i < kYTab; i += blockDim.x) {
        const float ok = i == kYIDZ ? 0.f : 1.f;
        const bool id = i >= Sp->nc;
        ly[i] = id ? make_float4(1.f, 0.f, 0.f, ok) : make_float4(Sp->ike[i], Sp->ae[i], Sp->be[i], 1.f);
        ly[kYTab + i] = id ? make_float4(1.f, 0.f, 0.f, ok) : make_float4(Sp->ikm[i], Sp->am[i], Sp->bm[i], 1.f);
    }
}

// Table index of a row at distance d from the outermost node row of its
// field (d < 0: no node): strip cell nc-1-d inside the strip, else kYID.
__device__ __forceinline__ int ytab_idx(int d, int nc) {
    const int e = nc - 1 - d;
    return e < 0 ? kYID : (e >= nc ? kYIDZ : e);
}
// Ez nodes are rows 1..ny-1 and H nodes rows 0..ny-1; the two strips of a
// field never overlap (ny >= 2 nc + 4), so the nearer edge decides.
__device__ __forceinline__ int ytab_e(int y, int ny, int nc) { return ytab_idx(min(y - 1, ny - 1 - y), nc); }
__device__ __forceinline__ int ytab_h(int y, int ny, int nc) { return ytab_idx(min(y, ny - 1 - y), nc); }

// H half-step of stage_update: returns Hx, Hy at n + 1/2 (c's H psi advanced).
template <int MODE>
__device__ __forceinline__ void stage_h(const SimView& S, const float* __restrict__ lcp, const WaveFlags& wf,
                                        Cell& c, float ez_up, int x, int y, float& hx_out, float& hy_out) {
    const int nx = S.nx, ny = S.ny, nc = S.nc;
    // ---------------- H half-step (H_field_update, FDTD_GPU.cu:298-355)
    const float ez_r = lane_next(c.ez);
    const float dEx = ez_r - c.ez, dEy = ez_up - c.ez;
    float hy, hx;
    // H nodes exist for 0 <= x < nx, 0 <= y < ny; the CPML branches run only there
    const bool hx_in = (x >= 0 && x < nx);
    const bool hy_in = (y >= 0 && y < ny);        // wave-uniform
    const bool ys_h = hy_in && (y < nc || y >= ny - nc);  // wave-uniform
    if (MODE == TOMO_MODE_REF_COMPAT) {
        const float chez = S.chx;
        hy = c.hy + chez * dEx;
        hx = c.hx - chez * dEy;
        if (wf.xs) {
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
        }
        if (ys_h && y < nc) {
            const int i = nc - 1 - y;
            hx = c.hx - chez * dEy / CP(1, i);
            c.phy = CP(7, i) * c.phy + CP(6, i) * dEy;
            hx = hx - S.chx_d * c.phy;
        }
        if (ys_h && y >= ny - nc) {
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
        if (wf.xs && hx_in && (x < nc || x >= nx - nc)) {
            const int i = x < nc ? nc - 1 - x : x - nx + nc;
            c.phx = CP(7, i) * c.phx + CP(6, i) * dEx;
            ty = dEx * CP(3, i) + S.dxf * c.phx;
        }
        if (ys_h) {
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
    hx_out = hx;
    hy_out = hy;
}

// E half-step of stage_update given the row's new H; hx_prev holds Hx of the
// row below on entry and this row's Hx on exit.
template <int MODE>
__device__ __forceinline__ void stage_e(const SimView& S, const float4* __restrict__ lc4,
                                        const float* __restrict__ lcp, const WaveFlags& wf, Cell& c, float hx,
                                        float hy, float& hx_prev, int x, int y, int n) {
    const int nx = S.nx, ny = S.ny, nc = S.nc;
    // ---------------- E half-step (E_field_update, FDTD_GPU.cu:358-434)
    const float hy_l = lane_prev(hy);
    const float dHy = hy - hy_l, dHx = hx - hx_prev;
    hx_prev = hx;
    c.hx = hx;
    c.hy = hy;
    if (y <= 0 || y >= ny) {  // wave-uniform: PEC rows and rows outside the grid
        c.ez = 0.f;
        return;
    }
    if (x <= 0 || x >= nx) {
        c.ez = 0.f;
        return;
    }
    const float4 qa = lc4[2 * c.mat], qb = lc4[2 * c.mat + 1];  // Ceze Cezh Cezeic Cezeip | Cezjp Kp BpDt disp
    const float ez = c.ez;
    const bool ys_e = (y <= nc || y >= ny - nc);  // wave-uniform
    if (MODE == TOMO_MODE_REF_COMPAT) {
        float f1 = 0.f, f0 = 0.f;
        if (qa.z != 0.f || qa.w != 0.f || qb.z != 0.f) {
            const float X = S.inc_abs ? (float)x : (float)(x - nx / 2);
            const float Y = S.inc_abs ? (float)y : (float)(y - ny / 2);
            const float sp = S.inc_c - S.inc_ay * Y - S.inc_ax * X;
            const float a1 = (((float)n + 0.5f) * S.inc_dt + sp) * S.inc_itau;
            const float a0 = (((float)n - 0.5f) * S.inc_dt + sp) * S.inc_itau;
            f1 = expf(-a1 * a1);
            f0 = expf(-a0 * a0);
        }
        float buf = qa.x * ez + qa.y * dHy - qa.y * dHx + qa.z * f1 + qa.w * f0 - qb.x * c.jp;
        if (!(S.flags & TOMO_FLAG_CPU_COEFS)) c.jp = qb.y * c.jp + qb.z * (buf - ez + f1 - f0);
        if (wf.xs) {
            if (x <= nc) {
                const int i = nc - x;
                buf = qa.x * ez + qa.y * dHy / CP(0, i) - qa.y * dHx / CP(0, i);
                c.pex = CP(5, i) * c.pex + CP(4, i) * dHy;
                buf = buf + qa.y * S.dxf * c.pex;
            }
            if (x >= nx - nc) {
                const int i = x - nx + nc;
                buf = qa.x * ez + qa.y * dHy / CP(0, i) - qa.y * dHx / CP(0, i);
                c.pex = CP(5, i) * c.pex + CP(4, i) * dHy;
                buf = buf + qa.y * S.dxf * c.pex;
            }
        }
        if (ys_e && y <= nc) {
            const int i = nc - y;
            buf = qa.x * ez + qa.y * dHy / CP(0, i) - qa.y * dHx / CP(0, i);
            c.pey = CP(5, i) * c.pey + CP(4, i) * dHx;
            buf = buf - qa.y * S.dyf * c.pey;
        }
        if (ys_e && y >= ny - nc) {
            const int i = y - ny + nc;
            buf = qa.x * ez + qa.y * dHy / CP(0, i) - qa.y * dHx / CP(0, i);
            c.pey = CP(5, i) * c.pey + CP(4, i) * dHx;
            buf = buf - qa.y * S.dyf * c.pey;
        }
        if (!(S.flags & TOMO_FLAG_NO_NAN_CLAMP) && __builtin_isnan(buf)) buf = 0.f;
        c.ez = buf;
    } else {
        float cy = dHy, cx = dHx;
        if (wf.xs && (x <= nc || x >= nx - nc)) {
            const int i = x <= nc ? nc - x : x - nx + nc;
            c.pex = CP(5, i) * c.pex + CP(4, i) * dHy;
            cy = dHy * CP(2, i) + S.dxf * c.pex;
        }
        if (ys_e) {
            const int i = y <= nc ? nc - y : y - ny + nc;
            c.pey = CP(5, i) * c.pey + CP(4, i) * dHx;
            cx = dHx * CP(2, i) + S.dyf * c.pey;
        }
        float en = qa.x * ez + qa.y * cy - qa.y * cx - qb.x * c.jp;
        if (wf.src && y == S.sy && x == S.sx) en += S.srcw[n];
        c.jp = qb.y * c.jp + qb.z * (en - ez);
        c.ez = en;
    }
}

// One leapfrog step of one row: c holds row y at level n (fields Ez^n,
// H^{n-1/2}); ez_up = Ez^n(x, y+1); hx_prev = Hx^{n+1/2}(x, y-1) on entry and
// Hx^{n+1/2}(x, y) on exit.  On exit c holds row y at level n+1.
template <int MODE>
__device__ __forceinline__ void stage_update(const SimView& S, const float4* __restrict__ lc4,
                                             const float* __restrict__ lcp, const WaveFlags& wf, Cell& c,
                                             float ez_up, float& hx_prev, int x, int y, int n) {
    float hx, hy;
    stage_h<MODE>(S, lcp, wf, c, ez_up, x, y, hx, hy);
    stage_e<MODE>(S, lc4, lcp, wf, c, hx, hy, hx_prev, x, y, n);
}

// Surface samples of the NTFF contour (contour map of FDTD_common.cxx:4-46):
// plane 0 = H on the face cell, 1 = Ez on the face cell, 2 = Ez on the cell
// across the face.  The running DFT averages planes 1 and 2.
__device__ __forceinline__ void write_samples(const SimView& S, const WaveFlags& wf, const Cell& c, int x, int y,
                                              int k) {
    const int D = S.D, nx = S.nx, ny = S.ny, P = S.P;
    const bool yrow = (y == D || y == D + 1 || y == ny - D - 1 || y == ny - D);  // wave-uniform
    const bool xrange = (y >= S.cy_lo && y <= ny - D - 1);                      // wave-uniform
    if (!yrow && !(wf.ccol && xrange)) return;
    gfloat* h_ = S.samp + k * P;
    gfloat* e0 = S.samp + (S.K + k) * P;
    gfloat* e1 = S.samp + (2 * S.K + k) * P;
    if (yrow) {
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
    }
    if (wf.ccol && xrange) {
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
                                                const DevSim* __restrict__ sims, int ordinal) {
    __shared__ float4 lc4[TOMO_MAX_MAT * 2];
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
        {
            float* lcf = (float*)lc4;
            for (int i = threadIdx.x; i < Sp->nmat * TOMO_NCOEF; i += blockDim.x) lcf[i] = Sp->coef[i];
        }
        for (int i = threadIdx.x; i < Sp->nc; i += blockDim.x) {
            CP(0, i) = Sp->ke[i]; CP(1, i) = Sp->km[i]; CP(2, i) = Sp->ike[i]; CP(3, i) = Sp->ikm[i];
            CP(4, i) = Sp->ae[i]; CP(5, i) = Sp->be[i]; CP(6, i) = Sp->am[i]; CP(7, i) = Sp->bm[i];
        }
        __syncthreads();
        SimView S;
        S.f = (gfloat*)Sp->fbase; S.jp = (gfloat*)Sp->jbase; S.xq = (gfloat*)Sp->xbase; S.yq = (gfloat*)Sp->ybase;
        S.mat = (gcu8*)Sp->mat; S.span = (kcint*)Sp->span; S.samp = (gfloat*)Sp->samp; S.srcw = (kcfloat*)Sp->srcw;
        S.ps = Sp->ps; S.js = Sp->js; S.xs = Sp->xs; S.ys = Sp->ys;
        S.par = it.parity0 ^ (ordinal & 1);
        S.nx = Sp->nx; S.ny = Sp->ny; S.pitch = Sp->pitch; S.nc = Sp->nc; S.D = Sp->D; S.P = Sp->P;
        S.L1 = Sp->L1; S.L2 = Sp->L2; S.L3 = Sp->L3; S.cx_hi = Sp->cx_hi; S.cy_lo = Sp->cy_lo;
        S.sx = Sp->sx; S.sy = Sp->sy; S.flags = Sp->flags; S.K = Sp->chunk;
        S.chx = Sp->chx; S.chy = Sp->chy; S.chx_d = Sp->chx_d; S.chy_d = Sp->chy_d; S.dxf = Sp->dxf;
        S.dyf = Sp->dyf; S.inc_dt = Sp->inc_dt; S.inc_c = Sp->inc_c; S.inc_ax = Sp->inc_ax;
        S.inc_ay = Sp->inc_ay; S.inc_itau = Sp->inc_itau; S.inc_abs = Sp->inc_abs_origin;
        if (wv < tile.nband) {
            const int bx = tile.x0 + wv * V;
            const int x0 = bx - T;
            const int x = x0 + lane;
            const bool xin = (x >= 0 && x <= S.nx);
            const bool xown = xin && lane >= T && x < min(bx + V, tile.xe);
            WaveFlags wf;
            wf.xs = (x0 <= S.nc) || (x0 + kWave - 1 >= S.nx - S.nc);
            wf.ccol = (x0 <= S.D + 1 && x0 + kWave - 1 >= S.D) ||
                      (x0 <= S.nx - S.D && x0 + kWave - 1 >= S.nx - S.D - 1);
            wf.src = (S.sx >= x0 && S.sx <= x0 + kWave - 1);
            const int y0 = tile.y0;
            const int y1 = tile.y1;
            const int step0 = it.step0 + ordinal * T;
            const int k0 = ordinal * T;
            Cell cur[T];
            float hxp[T];
#pragma unroll
            for (int t = 0; t < T; ++t) {
                cur[t] = zero_cell();
                hxp[t] = 0.f;
            }
            auto load = [&](int r, Cell& c) { load_cell(S, lc4, r, x, xin, wf, c); };
            auto store = [&](int yo, const Cell& c) {
                store_cell(S, lc4, yo, x, wf, c);
            };
            auto step = [&](int r, Cell in) {
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    const int y = r - 1 - t;
                    Cell out = cur[t];
                    if (t < nst) {
                        stage_update<MODE>(S, lc4, lcp, wf, out, in.ez, hxp[t], x, y, step0 + t);
                        if (y >= y0 && y < y1 && xown) write_samples(S, wf, out, x, y, k0 + t);
                    }
                    cur[t] = in;
                    in = out;
                }
                return in;
            };
            row_stream<T, kGroupRef, Cell>(y0, y1, xown, load, store, step);
        }
    }
}
// Edge tiles of physical-mode simulations: the CPML recursions are applied
// to every cell through coefficient tables whose identity entry (1/kappa = 1,
// a = b = 0) covers the cells outside the strips, so strip cells and ordinary
// cells take the same instruction stream; PEC and out-of-grid nodes are
// lane/row selects.  x-strip coefficients are lane constants held in
// registers, y-strip coefficients are wave-uniform per row.  Ordinary cells
// compute exactly what k_tinner computes (multiplications by 1 and additions
// of 0 are exact).
template <int T, bool BF>
__device__ __forceinline__ void fill_edge(const DevSim* __restrict__ Sp, float4* __restrict__ lc4, float* __restrict__ lcp,
                                          float4* __restrict__ ly) {
    if constexpr (BF) fill_ytab(Sp, ly);
    {
        float* lcf = (float*)lc4;
        for (int i = threadIdx.x; i < Sp->nmat * TOMO_NCOEF; i += blockDim.x) lcf[i] = Sp->coef[i];
    }
    for (int i = threadIdx.x; i < TOMO_MAX_NC; i += blockDim.x) {
        const bool id = (i >= Sp->nc);
        CP(2, i) = id ? 1.f : Sp->ike[i];
        CP(3, i) = id ? 1.f : Sp->ikm[i];
        CP(4, i) = id ? 0.f : Sp->ae[i];
        CP(5, i) = id ? 0.f : Sp->be[i];
        CP(6, i) = id ? 0.f : Sp->am[i];
        CP(7, i) = id ? 0.f : Sp->bm[i];
    }
}

template <int T, int G, bool BF, int WIN>
__device__ __forceinline__ void body_edge(const TileItem& tile, const LaunchItem& it, int nst,
                                         const DevSim* __restrict__ Sp, int ordinal,
                                         const float4* __restrict__ lc4, const float* __restrict__ lcp,
                                         const float4* __restrict__ ly) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    constexpr int V = kWave - 2 * T;
    constexpr int ID = TOMO_MAX_NC - 1;
    SimView S;
    S.f = (gfloat*)Sp->fbase; S.jp = (gfloat*)Sp->jbase; S.xq = (gfloat*)Sp->xbase; S.yq = (gfloat*)Sp->ybase;
    S.mat = (gcu8*)Sp->mat; S.span = (kcint*)Sp->span; S.samp = (gfloat*)Sp->samp; S.srcw = (kcfloat*)Sp->srcw;
    S.ps = Sp->ps; S.js = Sp->js; S.xs = Sp->xs; S.ys = Sp->ys;
    S.par = it.parity0 ^ (ordinal & 1);
    S.nx = Sp->nx; S.ny = Sp->ny; S.pitch = Sp->pitch; S.nc = Sp->nc; S.D = Sp->D; S.P = Sp->P;
    S.L1 = Sp->L1; S.L2 = Sp->L2; S.L3 = Sp->L3; S.cx_hi = Sp->cx_hi; S.cy_lo = Sp->cy_lo;
    S.sx = Sp->sx; S.sy = Sp->sy; S.flags = Sp->flags; S.K = Sp->chunk;
    S.chx = Sp->chx; S.chy = Sp->chy; S.dxf = Sp->dxf; S.dyf = Sp->dyf;
    const int nx = S.nx, ny = S.ny, nc = S.nc;
    const int bx = tile.x0 + wv * V;
    const int x0 = bx - T;
    const int x = x0 + lane;
    const bool xin = (x >= 0 && x <= nx);
    const bool xown = xin && lane >= T && x < min(bx + V, tile.xe);
    WaveFlags wf;
    wf.xs = (x0 <= nc) || (x0 + kWave - 1 >= nx - nc);
    wf.ccol = (x0 <= S.D + 1 && x0 + kWave - 1 >= S.D) || (x0 <= nx - S.D && x0 + kWave - 1 >= nx - S.D - 1);
    wf.src = (S.sx >= x0 && S.sx <= x0 + kWave - 1);
    // lane constants: x-strip coefficients and existence masks
    const int ixe = (x >= 1 && x <= nc) ? nc - x : ((x >= nx - nc && x <= nx - 1) ? x - nx + nc : ID);
    const int ixh = (x >= 0 && x < nc) ? nc - 1 - x : ((x >= nx - nc && x < nx) ? x - nx + nc : ID);
    const float kxe = CP(2, ixe), axe = CP(4, ixe), bxe = CP(5, ixe);
    const float kxh = CP(3, ixh), axh = CP(6, ixh), bxh = CP(7, ixh);
    const bool hx_ok = (x >= 0 && x < nx), ex_ok = (x >= 1 && x <= nx - 1);
    const float hxf = hx_ok ? 1.f : 0.f, exf = ex_ok ? 1.f : 0.f;
    const int y0 = tile.y0, y1 = tile.y1;
    const int step0 = it.step0 + ordinal * T;
    const int k0 = ordinal * T;
    Cell cur[T];
    float hxp[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        cur[t] = zero_cell();
        hxp[t] = 0.f;
    }
    const SpanWin sw = (WIN & kWinSpan) ? span_window(Sp->span, y0 - T, ny) : SpanWin{};
    const YWin yw = (!BF && (WIN & kWinY)) ? ywin_load(lcp + 2 * TOMO_MAX_NC, y0 - T, ny, nc) : YWin{};
    auto load = [&](int r, Cell& c) { load_cell(S, lc4, r, x, xin, wf, c, (WIN & kWinSpan) ? &sw : nullptr); };
    auto store = [&](int yo, const Cell& c) { store_cell(S, lc4, yo, x, wf, c, (WIN & kWinSpan) ? &sw : nullptr); };
    // FAST: every stage active (nst == T) and no source cell in reach of the tile,
    // so the stage loop carries no per-stage branch.
    auto stepg = [&](auto fast, int r, Cell in) {
        constexpr bool FAST = decltype(fast)::value;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int y = r - 1 - t;
            Cell c = cur[t];
            if constexpr (!BF) {
                if (FAST || t < nst) {
                    // wave-uniform row coefficients
                    YCoef yc;
                    if constexpr ((WIN & kWinY) != 0) {
                        yc = ycoef(yw, y);
                    } else {
                        const int iye = (y >= 1 && y <= nc) ? nc - y : ((y >= ny - nc && y <= ny - 1) ? y - ny + nc : ID);
                        const int iyh = (y >= 0 && y < nc) ? nc - 1 - y : ((y >= ny - nc && y < ny) ? y - ny + nc : ID);
                        yc = YCoef{CP(2, iye), CP(4, iye), CP(5, iye), CP(3, iyh), CP(6, iyh), CP(7, iyh)};
                    }
                    const bool hy_ok = (y >= 0 && y < ny), ey_ok = (y >= 1 && y <= ny - 1);
                    // H half-step with the CPML recursions
                    const float dEx = lane_next(c.ez) - c.ez;
                    const float dEy = in.ez - c.ez;
                    c.phx = bxh * c.phx + axh * dEx;
                    c.phy = yc.bh * c.phy + yc.ah * dEy;
                    const float ty = dEx * kxh + S.dxf * c.phx;
                    const float tx = dEy * yc.kh + S.dyf * c.phy;
                    const bool hok = hx_ok && hy_ok;
                    const float hy = hok ? c.hy + S.chy * ty : 0.f;
                    const float hx = hok ? c.hx - S.chx * tx : 0.f;
                    // E half-step with the CPML recursions and the Debye current
                    const float dHy = hy - lane_prev(hy);
                    const float dHx = hx - hxp[t];
                    hxp[t] = hx;
                    c.pex = bxe * c.pex + axe * dHy;
                    c.pey = yc.be * c.pey + yc.ae * dHx;
                    const float cy = dHy * kxe + S.dxf * c.pex;
                    const float cx = dHx * yc.ke + S.dyf * c.pey;
                    const float4 qa = lc4[2 * c.mat], qb = lc4[2 * c.mat + 1];
                    float en = qa.x * c.ez + qa.y * cy - qa.y * cx - qb.x * c.jp;
                    if (!FAST && wf.src && y == S.sy && x == S.sx) en += S.srcw[step0 + t];
                    const bool eok = ex_ok && ey_ok;
                    c.jp = eok ? qb.y * c.jp + qb.z * (en - c.ez) : c.jp;
                    c.ez = eok ? en : 0.f;
                    c.hx = hx;
                    c.hy = hy;
                    if (y >= y0 && y < y1 && xown) write_samples(S, wf, c, x, y, k0 + t);
                }
            } else if (FAST || t < nst) {
                // row coefficients and node masks (packed y tables)
                const float4 ch = ly[kYTab + ytab_h(y, ny, nc)], ce = ly[ytab_e(y, ny, nc)];
                // H half-step with the CPML recursions
                const float dEx = lane_next(c.ez) - c.ez;
                const float dEy = in.ez - c.ez;
                c.phx = bxh * c.phx + axh * dEx;
                c.phy = ch.z * c.phy + ch.y * dEy;
                const float ty = dEx * kxh + S.dxf * c.phx;
                const float tx = dEy * ch.x + S.dyf * c.phy;
                // node masks as factors (0 or 1): exact for the nodes that
                // exist, and no uniform branch splits the stage
                const float hm = hxf * ch.w, em = exf * ce.w;
                const float hy = (c.hy + S.chy * ty) * hm;
                const float hx = (c.hx - S.chx * tx) * hm;
                // E half-step with the CPML recursions and the Debye current
                const float dHy = hy - lane_prev(hy);
                const float dHx = hx - hxp[t];
                hxp[t] = hx;
                c.pex = bxe * c.pex + axe * dHy;
                c.pey = ce.z * c.pey + ce.y * dHx;
                const float cy = dHy * kxe + S.dxf * c.pex;
                const float cx = dHx * ce.x + S.dyf * c.pey;
                // without an Ez node: Ez stays 0 and Jp keeps its value
                const float4 qa = lc4[2 * c.mat], qb = lc4[2 * c.mat + 1];
                const float ce_ = qa.x * em, ch_ = qa.y * em, cj_ = qb.x * em;
                const float kp_ = qb.y * em + (1.f - em), bp_ = qb.z * em;
                float en = ce_ * c.ez + ch_ * cy - ch_ * cx - cj_ * c.jp;
                if (!FAST && wf.src && y == S.sy && x == S.sx) en += S.srcw[step0 + t] * em;
                c.jp = kp_ * c.jp + bp_ * (en - c.ez);
                c.ez = en;
                c.hx = hx;
                c.hy = hy;
            }
            cur[t] = in;
            in = c;
        }
        // branch-free form: NTFF samples of every stage after the stage
        // arithmetic; stage t's output row is now cur[t + 1], the last stage's `in`
        if constexpr (BF) {
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int y = r - 1 - t;
                if ((FAST || t < nst) && y >= y0 && y < y1 && xown)
                    write_samples(S, wf, t + 1 < T ? cur[t + 1] : in, x, y, k0 + t);
            }
        }
        return in;
    };
    auto step_fast = [&](int r, Cell in) { return stepg(std::true_type{}, r, in); };
    auto step_gen = [&](int r, Cell in) { return stepg(std::false_type{}, r, in); };
    const bool src_near = wf.src && S.sy >= y0 - 2 * T && S.sy < y1 + T;
    if (nst == T && !src_near) row_sweep<T, G, Cell>(y0, y1, xown, load, store, step_fast);
    else row_sweep<T, G, Cell>(y0, y1, xown, load, store, step_gen);
}

template <int T>
__global__ __launch_bounds__(256) void k_tedge(const TileItem* __restrict__ tiles, int ntiles,
                                               const LaunchItem* __restrict__ items,
                                               const DevSim* __restrict__ sims, int ordinal) {
    __shared__ float4 lc4[TOMO_MAX_MAT * 2];
    constexpr bool BF = kStageBranchFree<T>;
    __shared__ float lcp[8 * TOMO_MAX_NC];
    __shared__ float4 ly[BF ? 2 * kYTab : 1];
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const TileItem tile = tiles[ti];
        const LaunchItem it = items[tile.item];
        const int nst = min(T, it.nsteps - ordinal * T);
        if (nst <= 0) continue;
        const DevSim* Sp = sims + it.sim;
        __syncthreads();
        fill_edge<T, BF>(Sp, lc4, lcp, ly);
        __syncthreads();
        if ((int)(threadIdx.x / kWave) >= tile.nband) continue;  // after the last barrier of this tile
        body_edge<T, kGroupEdge, BF, kStageWin>(tile, it, nst, Sp, ordinal, lc4, lcp, ly);
    }
}


#undef CP

// ---------------------------------------------------------------- resident update
// One workgroup of kResWaves waves per tile (ResTile), all tiles of the launch
// co-resident (cooperative launch).  Wave w keeps region rows
// [y0 - T + w RW, y0 - T + (w + 1) RW) of a 64-lane band in registers for
// the whole epoch, and every step advances them bottom to top with the same
// stage arithmetic as the general edge kernel (select form), so every cell
// gets bit-identical values in every launch form.  Per step the waves of a
// workgroup couple through LDS: each publishes its bottom row's Ez^n, then
// the H half-step of its top row (which needs the next wave's Ez^n), then all
// rows run; two barriers per step.  Per time block of T steps, ghost lanes
// and rows (stale after T steps, like a band's ghost zone) are reloaded from
// the global planes, where the neighbouring tiles stored their T-deep frame
// at the end of the previous block and published it through their
// completed-block counter.  Frame stores, ghost loads and the counters are
// relaxed atomics at agent scope, coherent across the XCDs' L2 caches, and
// every frame store has completed (vmcnt 0, then a barrier) before the
// counter is written, so no cache writeback or invalidation is needed.  Waits are bounded: past `spin_limit` polls, or once
// any workgroup has given up, a workgroup raises the error word and returns,
// so every wave exits.
constexpr int kResWaves = 8;

// Lane constants of a band: x-strip CPML coefficients (identity off the strips)
// and the node masks of the column.
struct ResLane {
    float kxe, axe, bxe, kxh, axh, bxh;
    bool hx_ok, ex_ok;
};

#define CP(k, i) lcp[(k) * TOMO_MAX_NC + (i)]
// H half-step of one row (the select-form stage of body_edge): returns Hx, Hy
// at n + 1/2 and advances the row's H psi.
__device__ __forceinline__ void res_h(const SimView& S, const float* __restrict__ lcp, const ResLane& L, Cell& c,
                                      float ez_up, int y, float& hx, float& hy) {
    constexpr int ID = TOMO_MAX_NC - 1;
    const int nc = S.nc, ny = S.ny;
    const int iyh = (y >= 0 && y < nc) ? nc - 1 - y : ((y >= ny - nc && y < ny) ? y - ny + nc : ID);
    const bool hy_ok = (y >= 0 && y < ny);
    const float dEx = lane_next(c.ez) - c.ez;
    const float dEy = ez_up - c.ez;
    c.phx = L.bxh * c.phx + L.axh * dEx;
    c.phy = CP(7, iyh) * c.phy + CP(6, iyh) * dEy;
    const float ty = dEx * L.kxh + S.dxf * c.phx;
    const float tx = dEy * CP(3, iyh) + S.dyf * c.phy;
    const bool hok = L.hx_ok && hy_ok;
    hy = hok ? c.hy + S.chy * ty : 0.f;
    hx = hok ? c.hx - S.chx * tx : 0.f;
}

// E half-step of one row given its new H (and Hx of the row below in hxp,
// which receives this row's Hx).
__device__ __forceinline__ void res_e(const SimView& S, const float4* __restrict__ lc4, const float* __restrict__ lcp,
                                      const ResLane& L, Cell& c, float hx, float hy, float& hxp, int y, int x, int n,
                                      bool wsrc) {
    constexpr int ID = TOMO_MAX_NC - 1;
    const int nc = S.nc, ny = S.ny;
    const int iye = (y >= 1 && y <= nc) ? nc - y : ((y >= ny - nc && y <= ny - 1) ? y - ny + nc : ID);
    const bool ey_ok = (y >= 1 && y <= ny - 1);
    const float dHy = hy - lane_prev(hy);
    const float dHx = hx - hxp;
    hxp = hx;
    c.pex = L.bxe * c.pex + L.axe * dHy;
    c.pey = CP(5, iye) * c.pey + CP(4, iye) * dHx;
    const float cy = dHy * L.kxe + S.dxf * c.pex;
    const float cx = dHx * CP(2, iye) + S.dyf * c.pey;
    const float4 qa = lc4[2 * c.mat], qb = lc4[2 * c.mat + 1];
    float en = qa.x * c.ez + qa.y * cy - qa.y * cx - qb.x * c.jp;
    if (wsrc && y == S.sy && x == S.sx) en += S.srcw[n];
    const bool eok = L.ex_ok && ey_ok;
    c.jp = eok ? qb.y * c.jp + qb.z * (en - c.ez) : c.jp;
    c.ez = eok ? en : 0.f;
    c.hx = hx;
    c.hy = hy;
}

// Row state the resident kernel keeps in registers; the CPML psi of waves
// that need it live in LDS.
struct RCell {
    float ez, hx, hy, jp;
    int mat;
};

template <int T, int RW, int MODE>
__global__ __launch_bounds__(kResWaves * kWave, 1) void k_tres(const ResTile* __restrict__ tiles,
                                                              const LaunchItem* __restrict__ items,
                                                              const DevSim* __restrict__ sims, int* done,
                                                              int* err, int spin_limit) {
    __shared__ float4 lc4[TOMO_MAX_MAT * 2];
    __shared__ float lcp[8 * TOMO_MAX_NC];
    __shared__ float ezb[kResWaves][kWave];
    __shared__ float hxt[kResWaves][kWave];
    __shared__ float4 lpsi[kResWaves][RW][kWave];  // {pex, phx, pey, phy} of general waves
    __shared__ int s_abort;
    constexpr int ID = TOMO_MAX_NC - 1;
    const int tid = blockIdx.x;
    const ResTile tile = tiles[tid];
    const LaunchItem it = items[tile.item];
    const DevSim* Sp = sims + it.sim;
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);  // wave-uniform
    if constexpr (MODE == TOMO_MODE_PHYSICAL) {
        fill_edge<T, false>(Sp, lc4, lcp, nullptr);
    } else {
        // the reference's tables as they are (its stage divides by kappa)
        float* lcf = (float*)lc4;
        for (int i = threadIdx.x; i < Sp->nmat * TOMO_NCOEF; i += blockDim.x) lcf[i] = Sp->coef[i];
        for (int i = threadIdx.x; i < Sp->nc; i += blockDim.x) {
            CP(0, i) = Sp->ke[i]; CP(1, i) = Sp->km[i]; CP(2, i) = Sp->ike[i]; CP(3, i) = Sp->ikm[i];
            CP(4, i) = Sp->ae[i]; CP(5, i) = Sp->be[i]; CP(6, i) = Sp->am[i]; CP(7, i) = Sp->bm[i];
        }
    }
    if (threadIdx.x == 0) s_abort = 0;
    __syncthreads();
    SimView S;
    S.f = (gfloat*)Sp->fbase; S.jp = (gfloat*)Sp->jbase; S.xq = (gfloat*)Sp->xbase; S.yq = (gfloat*)Sp->ybase;
    S.mat = (gcu8*)Sp->mat; S.span = (kcint*)Sp->span; S.samp = (gfloat*)Sp->samp; S.srcw = (kcfloat*)Sp->srcw;
    S.ps = Sp->ps; S.js = Sp->js; S.xs = Sp->xs; S.ys = Sp->ys;
    S.nx = Sp->nx; S.ny = Sp->ny; S.pitch = Sp->pitch; S.nc = Sp->nc; S.D = Sp->D; S.P = Sp->P;
    S.L1 = Sp->L1; S.L2 = Sp->L2; S.L3 = Sp->L3; S.cx_hi = Sp->cx_hi; S.cy_lo = Sp->cy_lo;
    S.sx = Sp->sx; S.sy = Sp->sy; S.flags = Sp->flags; S.K = Sp->chunk;
    S.chx = Sp->chx; S.chy = Sp->chy; S.chx_d = Sp->chx_d; S.chy_d = Sp->chy_d; S.dxf = Sp->dxf; S.dyf = Sp->dyf;
    S.inc_dt = Sp->inc_dt; S.inc_c = Sp->inc_c; S.inc_ax = Sp->inc_ax; S.inc_ay = Sp->inc_ay;
    S.inc_itau = Sp->inc_itau; S.inc_abs = Sp->inc_abs_origin;
    const int nx = S.nx, ny = S.ny, nc = S.nc, D = S.D;
    const int x0 = tile.x0 - T;  // first lane's column
    const int x = x0 + lane;
    const bool xin = (x >= 0 && x <= nx);
    const bool xown = xin && lane >= T && lane < kWave - T;
    const bool gl = lane < T || lane >= kWave - T;          // ghost lane
    const bool fl = lane < 2 * T || lane >= kWave - 2 * T;  // frame lane (or ghost)
    WaveFlags wf;
    wf.xs = (x0 <= nc) || (x0 + kWave - 1 >= nx - nc);
    wf.ccol = (x0 <= D + 1 && x0 + kWave - 1 >= D) || (x0 <= nx - D && x0 + kWave - 1 >= nx - D - 1);
    wf.src = (S.sx >= x0 && S.sx <= x0 + kWave - 1);
    ResLane L{};
    if constexpr (MODE == TOMO_MODE_PHYSICAL) {
        const int ixe = (x >= 1 && x <= nc) ? nc - x : ((x >= nx - nc && x <= nx - 1) ? x - nx + nc : ID);
        const int ixh = (x >= 0 && x < nc) ? nc - 1 - x : ((x >= nx - nc && x < nx) ? x - nx + nc : ID);
        L.kxe = CP(2, ixe); L.axe = CP(4, ixe); L.bxe = CP(5, ixe);
        L.kxh = CP(3, ixh); L.axh = CP(6, ixh); L.bxh = CP(7, ixh);
        L.hx_ok = (x >= 0 && x < nx);
        L.ex_ok = (x >= 1 && x <= nx - 1);
    }
    const int y0 = tile.y0, y1 = tile.y1;
    const int yb = y0 - T + wv * RW;  // this wave's first row
    // Clean waves (every row and lane clear of CPML strips, contour rows and
    // columns, PEC and the grid's end, like an interior tile) take the bare
    // Yee/Debye stage; the others the general one with psi in LDS.
    const int cbot = max(nc, D + 1) + 1, ctop = min(ny - nc, ny - D - 1) - 1;
    const int cl = max(nc, D + 1) + 1, cr = min(nx - nc, nx - D - 1) - 1;
    const bool clean = MODE == TOMO_MODE_PHYSICAL && yb >= cbot && yb + RW - 1 <= ctop && x0 >= cl &&
                       x0 + kWave - 1 <= cr;
    const int nblocks = (it.nsteps + T - 1) / T;
    RCell rows[RW];
    float4* psi = lpsi[wv][0] + lane;  // row i at psi[i * kWave]
    auto put = [&](int i, const Cell& c) {
        rows[i].ez = c.ez; rows[i].hx = c.hx; rows[i].hy = c.hy; rows[i].jp = c.jp; rows[i].mat = c.mat;
        if (!clean) psi[i * kWave] = make_float4(c.pex, c.phx, c.pey, c.phy);
    };
    auto get = [&](int i) {
        Cell c;
        c.ez = rows[i].ez; c.hx = rows[i].hx; c.hy = rows[i].hy; c.jp = rows[i].jp; c.mat = rows[i].mat;
        const float4 q = clean ? make_float4(0.f, 0.f, 0.f, 0.f) : psi[i * kWave];
        c.pex = q.x; c.phx = q.y; c.pey = q.z; c.phy = q.w;
        return c;
    };
    S.par = it.parity0;
#pragma unroll
    for (int i = 0; i < RW; ++i) {
        // an opaque row index keeps the compiler from hoisting every row's
        // addresses out of the block loop (which would exhaust the registers)
        int y = __builtin_amdgcn_readfirstlane(yb + i);
        asm volatile("" : "+s"(y));
        Cell c;
        load_cell(S, lc4, y, x, xin, wf, c);
        put(i, c);
    }
    for (int b = 0; b < nblocks; ++b) {
        S.par = it.parity0 ^ (b & 1);
        if (b > 0) {
            // wait until every neighbour has stored its frame of block b - 1
            if (threadIdx.x == 0) {
                int spins = 0, ok = 1;
                for (int k = 0; k < 8 && ok; ++k) {
                    const int nbt = tile.nb[k];
                    if (nbt < 0) continue;
                    while (__hip_atomic_load(done + nbt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < b) {
                        if (++spins > spin_limit ||
                            __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                            ok = 0;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(2);
                    }
                }
                if (!ok) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    s_abort = 1;
                }
            }
            __syncthreads();
            if (s_abort) return;  // uniform over the workgroup
            // ghost rows and ghost lanes from the neighbours' frames (coherent
            // loads: the frames were written by other workgroups of this launch)
#pragma unroll
            for (int i = 0; i < RW; ++i) {
                int y = __builtin_amdgcn_readfirstlane(yb + i);
                asm volatile("" : "+s"(y));
                if (y < y0 || y >= y1 || gl) {
                    Cell c;
                    load_cell<true>(S, lc4, y, x, xin, wf, c);
                    put(i, c);
                }
            }
        }
        const int step0 = it.step0 + b * T, k0 = b * T;
        const int nst = min(T, it.nsteps - b * T);
        for (int t = 0; t < nst; ++t) {
            const int n = step0 + t;
            ezb[wv][lane] = rows[0].ez;
            __syncthreads();
            const float ez_top = wv + 1 < kResWaves ? ezb[wv + 1][lane] : 0.f;
            // H half-step of the top row first: the wave above needs its Hx
            float htx, hty;
            if (clean) {
                const RCell& tp = rows[RW - 1];
                const float dEx = lane_next(tp.ez) - tp.ez;
                const float dEy = ez_top - tp.ez;
                hty = tp.hy + S.chy * dEx;
                htx = tp.hx - S.chx * dEy;
            } else {
                int yt = __builtin_amdgcn_readfirstlane(yb + RW - 1);
                asm volatile("" : "+s"(yt));
                Cell tc = get(RW - 1);
                if constexpr (MODE == TOMO_MODE_PHYSICAL) res_h(S, lcp, L, tc, ez_top, yt, htx, hty);
                else stage_h<MODE>(S, lcp, wf, tc, ez_top, x, yt, htx, hty);
                psi[(RW - 1) * kWave] = make_float4(tc.pex, tc.phx, tc.pey, tc.phy);
            }
            hxt[wv][lane] = htx;
            __syncthreads();
            float hxp = wv > 0 ? hxt[wv - 1][lane] : 0.f;
            if (clean) {
                const float chx = S.chx, chy = S.chy;
#pragma unroll
                for (int i = 0; i < RW; ++i) {
                    RCell& c = rows[i];
                    float hx = htx, hy = hty;
                    if (i + 1 < RW) {
                        const float dEx = lane_next(c.ez) - c.ez;
                        const float dEy = rows[i + 1].ez - c.ez;
                        hy = c.hy + chy * dEx;
                        hx = c.hx - chx * dEy;
                    }
                    const float dHy = hy - lane_prev(hy);
                    const float dHx = hx - hxp;
                    hxp = hx;
                    // opaque per step: the coefficient reads stay in the step
                    // loop instead of occupying registers for the whole kernel
                    int m = c.mat;
                    asm volatile("" : "+v"(m));
                    const float4 qa = lc4[2 * m], qb = lc4[2 * m + 1];
                    float en = qa.x * c.ez + qa.y * dHy - qa.y * dHx - qb.x * c.jp;
                    if (wf.src && yb + i == S.sy && x == S.sx) en += S.srcw[n];
                    c.jp = qb.y * c.jp + qb.z * (en - c.ez);
                    c.ez = en;
                    c.hx = hx;
                    c.hy = hy;
                }
            } else {
#pragma unroll
                for (int i = 0; i < RW; ++i) {
                    // opaque per step: the row's CPML and material reads stay in
                    // the step loop instead of occupying registers for the kernel
                    int y = __builtin_amdgcn_readfirstlane(yb + i);
                    asm volatile("" : "+s"(y));
                    Cell c = get(i);
                    asm volatile("" : "+v"(c.mat));
                    float hx = htx, hy = hty;
                    if constexpr (MODE == TOMO_MODE_PHYSICAL) {
                        if (i + 1 < RW) res_h(S, lcp, L, c, rows[i + 1].ez, y, hx, hy);
                        res_e(S, lc4, lcp, L, c, hx, hy, hxp, y, x, n, wf.src);
                    } else {
                        if (i + 1 < RW) stage_h<MODE>(S, lcp, wf, c, rows[i + 1].ez, x, y, hx, hy);
                        stage_e<MODE>(S, lc4, lcp, wf, c, hx, hy, hxp, x, y, n);
                    }
                    put(i, c);
                    if (y >= y0 && y < y1 && xown) write_samples(S, wf, c, x, y, k0 + t);
                }
            }
        }
        // the frame (or, after the last block, every owned cell) to plane par ^ 1
        const bool last = b + 1 == nblocks;
        if (last) {
            // the state for later launches: plain stores (the launch boundary orders them)
#pragma unroll
            for (int i = 0; i < RW; ++i) {
                int y = __builtin_amdgcn_readfirstlane(yb + i);
                asm volatile("" : "+s"(y));
                if (y >= y0 && y < y1 && xown) store_cell(S, lc4, y, x, wf, get(i));
            }
        } else {
            // the frame, with coherent stores; once they have all completed
            // (vmcnt 0 in every wave, then the barrier) the counter publishes it
#pragma unroll
            for (int i = 0; i < RW; ++i) {
                int y = __builtin_amdgcn_readfirstlane(yb + i);
                asm volatile("" : "+s"(y));
                if (y >= y0 && y < y1 && xown && (fl || y < y0 + T || y >= y1 - T))
                    store_cell<true>(S, lc4, y, x, wf, get(i));
            }
            __builtin_amdgcn_s_waitcnt(kWaitVm0);
            __syncthreads();
            if (threadIdx.x == 0) __hip_atomic_store(done + tid, b + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ---------------------------------------------------------------- register-tile update
// Tile launch form (TOMO_FORM_TILE): one workgroup of kTileWaves waves per
// tile and time block.  The workgroup loads a 64-lane band (owned columns
// [x0, x0 + 64 - 2T), T ghost lanes on each side) times the rows
// [y0 - T, y1 + T) into registers, RW rows per wave, all at once, advances
// them T steps with the resident kernel's stage arithmetic (so every form
// gives bit-identical fields), and stores the owned rows.  Per step the waves
// couple through LDS with two barriers (each publishes its bottom row's Ez^n,
// then its top row's Hx^{n+1/2}); inside a wave the RW rows of a half-step are
// independent, so the stage chains of a wave overlap.  Unlike the row sweeps
// no row waits on the one before it, and unlike the resident form there is no
// exchange between tiles: ghost rows and lanes (stale after T steps) are
// reloaded with the whole tile by the next launch.  CPML psi live in
// registers with the row.
constexpr int kTileWaves = kTileWaveCount;

template <int T, int RW>
__global__ __launch_bounds__(kTileWaves * kWave) __attribute__((amdgpu_waves_per_eu(4))) void k_twave(const TileItem* __restrict__ tiles, int ntiles,
                                                             const LaunchItem* __restrict__ items,
                                                             const DevSim* __restrict__ sims, int ordinal) {
    __shared__ float4 lc4[TOMO_MAX_MAT * 2];
    __shared__ float lcp[8 * TOMO_MAX_NC];
    __shared__ float ezb[kTileWaves][kWave];
    __shared__ float hxt[kTileWaves][kWave];
    __shared__ float4 lpsi[kTileWaves][RW][kWave];  // {pex, phx, pey, phy} of general waves
    constexpr int ID = TOMO_MAX_NC - 1;
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);  // wave-uniform
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const TileItem tile = tiles[ti];
        const LaunchItem it = items[tile.item];
        const int nst = min(T, it.nsteps - ordinal * T);
        if (nst <= 0) continue;  // uniform over the workgroup
        const DevSim* Sp = sims + it.sim;
        __syncthreads();  // the previous tile's tables are no longer read
        fill_edge<T, false>(Sp, lc4, lcp, nullptr);
        __syncthreads();
        SimView S;
        S.f = (gfloat*)Sp->fbase; S.jp = (gfloat*)Sp->jbase; S.xq = (gfloat*)Sp->xbase; S.yq = (gfloat*)Sp->ybase;
        S.mat = (gcu8*)Sp->mat; S.span = (kcint*)Sp->span; S.samp = (gfloat*)Sp->samp; S.srcw = (kcfloat*)Sp->srcw;
        S.ps = Sp->ps; S.js = Sp->js; S.xs = Sp->xs; S.ys = Sp->ys;
        S.par = it.parity0 ^ (ordinal & 1);
        S.nx = Sp->nx; S.ny = Sp->ny; S.pitch = Sp->pitch; S.nc = Sp->nc; S.D = Sp->D; S.P = Sp->P;
        S.L1 = Sp->L1; S.L2 = Sp->L2; S.L3 = Sp->L3; S.cx_hi = Sp->cx_hi; S.cy_lo = Sp->cy_lo;
        S.sx = Sp->sx; S.sy = Sp->sy; S.flags = Sp->flags; S.K = Sp->chunk;
        S.chx = Sp->chx; S.chy = Sp->chy; S.chx_d = Sp->chx_d; S.chy_d = Sp->chy_d; S.dxf = Sp->dxf;
        S.dyf = Sp->dyf;
        const int nx = S.nx, ny = S.ny, nc = S.nc, D = S.D;
        const int x0 = tile.x0 - T;  // first lane's column
        const int x = x0 + lane;
        const bool xin = (x >= 0 && x <= nx);
        const bool xown = xin && lane >= T && x < min(tile.x0 + kWave - 2 * T, tile.xe);
        WaveFlags wf;
        wf.xs = (x0 <= nc) || (x0 + kWave - 1 >= nx - nc);
        wf.ccol = (x0 <= D + 1 && x0 + kWave - 1 >= D) || (x0 <= nx - D && x0 + kWave - 1 >= nx - D - 1);
        wf.src = (S.sx >= x0 && S.sx <= x0 + kWave - 1);
        ResLane L;
        {
            const int ixe = (x >= 1 && x <= nc) ? nc - x : ((x >= nx - nc && x <= nx - 1) ? x - nx + nc : ID);
            const int ixh = (x >= 0 && x < nc) ? nc - 1 - x : ((x >= nx - nc && x < nx) ? x - nx + nc : ID);
            L.kxe = CP(2, ixe); L.axe = CP(4, ixe); L.bxe = CP(5, ixe);
            L.kxh = CP(3, ixh); L.axh = CP(6, ixh); L.bxh = CP(7, ixh);
            L.hx_ok = (x >= 0 && x < nx);
            L.ex_ok = (x >= 1 && x <= nx - 1);
        }
        const int y0 = tile.y0, y1 = tile.y1;
        const int yb = y0 - T + wv * RW;  // this wave's first row
        // Clean waves (every row and lane clear of CPML strips, contour rows and
        // columns, PEC and the grid's end) take the bare Yee/Debye stage.
        const int cbot = max(nc, D + 1) + 1, ctop = min(ny - nc, ny - D - 1) - 1;
        const int cl = max(nc, D + 1) + 1, cr = min(nx - nc, nx - D - 1) - 1;
        const bool clean = yb >= cbot && yb + RW - 1 <= ctop && x0 >= cl && x0 + kWave - 1 <= cr;
        RCell rows[RW];
        float4* psi = lpsi[wv][0] + lane;  // row i at psi[i * kWave]
        auto put = [&](int i, const Cell& c) {
            rows[i].ez = c.ez; rows[i].hx = c.hx; rows[i].hy = c.hy; rows[i].jp = c.jp; rows[i].mat = c.mat;
            if (!clean) psi[i * kWave] = make_float4(c.pex, c.phx, c.pey, c.phy);
        };
        auto get = [&](int i) {
            Cell c;
            c.ez = rows[i].ez; c.hx = rows[i].hx; c.hy = rows[i].hy; c.jp = rows[i].jp; c.mat = rows[i].mat;
            const float4 q = clean ? make_float4(0.f, 0.f, 0.f, 0.f) : psi[i * kWave];
            c.pex = q.x; c.phx = q.y; c.pey = q.z; c.phy = q.w;
            return c;
        };
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            Cell c;
            load_cell(S, lc4, yb + i, x, xin, wf, c);
            put(i, c);
        }
        const int step0 = it.step0 + ordinal * T, k0 = ordinal * T;
        for (int t = 0; t < nst; ++t) {
            const int n = step0 + t;
            ezb[wv][lane] = rows[0].ez;
            __syncthreads();
            const float ez_top = wv + 1 < kTileWaves ? ezb[wv + 1][lane] : 0.f;
            // H half-step of the top row first: the wave above needs its Hx
            float htx, hty;
            if (clean) {
                const RCell& tp = rows[RW - 1];
                const float dEx = lane_next(tp.ez) - tp.ez;
                const float dEy = ez_top - tp.ez;
                hty = tp.hy + S.chy * dEx;
                htx = tp.hx - S.chx * dEy;
            } else {
                Cell tc = get(RW - 1);
                res_h(S, lcp, L, tc, ez_top, yb + RW - 1, htx, hty);
                psi[(RW - 1) * kWave] = make_float4(tc.pex, tc.phx, tc.pey, tc.phy);
            }
            hxt[wv][lane] = htx;
            __syncthreads();
            float hxp = wv > 0 ? hxt[wv - 1][lane] : 0.f;
            if (clean) {
                const float chx = S.chx, chy = S.chy;
#pragma unroll
                for (int i = 0; i < RW; ++i) {
                    RCell& c = rows[i];
                    float hx = htx, hy = hty;
                    if (i + 1 < RW) {
                        const float dEx = lane_next(c.ez) - c.ez;
                        const float dEy = rows[i + 1].ez - c.ez;
                        hy = c.hy + chy * dEx;
                        hx = c.hx - chx * dEy;
                    }
                    const float dHy = hy - lane_prev(hy);
                    const float dHx = hx - hxp;
                    hxp = hx;
                    const float4 qa = lc4[2 * c.mat], qb = lc4[2 * c.mat + 1];
                    float en = qa.x * c.ez + qa.y * dHy - qa.y * dHx - qb.x * c.jp;
                    if (wf.src && yb + i == S.sy && x == S.sx) en += S.srcw[n];
                    c.jp = qb.y * c.jp + qb.z * (en - c.ez);
                    c.ez = en;
                    c.hx = hx;
                    c.hy = hy;
                }
            } else {
#pragma unroll
                for (int i = 0; i < RW; ++i) {
                    const int y = yb + i;
                    Cell c = get(i);
                    float hx = htx, hy = hty;
                    if (i + 1 < RW) res_h(S, lcp, L, c, rows[i + 1].ez, y, hx, hy);
                    res_e(S, lc4, lcp, L, c, hx, hy, hxp, y, x, n, wf.src);
                    put(i, c);
                    if (y >= y0 && y < y1 && xown) write_samples(S, wf, c, x, y, k0 + t);
                }
            }
        }
        // the owned rows to plane par ^ 1
#pragma unroll
        for (int i = 0; i < RW; ++i) {
            const int y = yb + i;
            if (y >= y0 && y < y1 && xown) store_cell(S, lc4, y, x, wf, get(i));
        }
    }
}
#undef CP

// Row-edge tiles of physical-mode simulations: the top and bottom rows
// (y-side CPML strips, y-face NTFF rows, PEC rows, rows outside the grid) over
// clean columns (every lane inside the grid, away from x strips and x-face
// contour columns).  Every row-dependent choice is wave-uniform: the y-strip
// coefficients come from the identity-padded table by a uniform index.
struct YCell {
    float ez, hx, hy, jp, pey, phy;
    int mat;
};

template <int T, bool BF>
__device__ __forceinline__ void fill_row(const DevSim* __restrict__ Sp, float4* __restrict__ lc4, float* __restrict__ lcy,
                                         float4* __restrict__ ly) {
    {
        float* lcf = (float*)lc4;
        for (int i = threadIdx.x; i < Sp->nmat * TOMO_NCOEF; i += blockDim.x) lcf[i] = Sp->coef[i];
    }
    if constexpr (BF) {
        fill_ytab(Sp, ly);
    } else {
        for (int i = threadIdx.x; i < TOMO_MAX_NC; i += blockDim.x) {
            const bool id = (i >= Sp->nc);
            lcy[0 * TOMO_MAX_NC + i] = id ? 1.f : Sp->ike[i];
            lcy[1 * TOMO_MAX_NC + i] = id ? 1.f : Sp->ikm[i];
            lcy[2 * TOMO_MAX_NC + i] = id ? 0.f : Sp->ae[i];
            lcy[3 * TOMO_MAX_NC + i] = id ? 0.f : Sp->be[i];
            lcy[4 * TOMO_MAX_NC + i] = id ? 0.f : Sp->am[i];
            lcy[5 * TOMO_MAX_NC + i] = id ? 0.f : Sp->bm[i];
        }
    }
}

template <int T, int G, bool BF, int WIN>
__device__ __forceinline__ void body_row(const TileItem& tile, const LaunchItem& it, int nst,
                                         const DevSim* __restrict__ Sp, int ordinal,
                                         const float4* __restrict__ lc4, const float* __restrict__ lcy,
                                         const float4* __restrict__ ly) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    constexpr int V = kWave - 2 * T;
    constexpr int ID = TOMO_MAX_NC - 1;
    const int par = it.parity0 ^ (ordinal & 1);
    const int ps = Sp->ps, js = Sp->js, ys = Sp->ys, pitch = Sp->pitch, ny = Sp->ny, nc = Sp->nc;
    gfloat* fin = (gfloat*)Sp->fbase + par * ps;
    gfloat* fout = (gfloat*)Sp->fbase + (par ^ 1) * ps;
    gfloat* jin = (gfloat*)Sp->jbase + par * js;
    gfloat* jout = (gfloat*)Sp->jbase + (par ^ 1) * js;
    gfloat* pe_in = (gfloat*)Sp->ybase + par * ys;
    gfloat* pe_out = (gfloat*)Sp->ybase + (par ^ 1) * ys;
    gfloat* ph_in = (gfloat*)Sp->ybase + (2 + par) * ys;
    gfloat* ph_out = (gfloat*)Sp->ybase + (2 + (par ^ 1)) * ys;
    gcu8* mat = (gcu8*)Sp->mat;
    const float chx = Sp->chx, chy = Sp->chy, dyf = Sp->dyf;
    const int bx = tile.x0 + wv * V;
    const int x0 = bx - T;
    const int x = x0 + lane;
    const bool own = lane >= T && x < min(bx + V, tile.xe);
    const int sx = Sp->sx, sy = Sp->sy;
    const bool wsrc = (sx >= x0 && sx <= x0 + kWave - 1);
    kcfloat* srcw = (kcfloat*)Sp->srcw;
    // y-face NTFF rows (FDTD_common.cxx:4-46 legs yn, yp)
    const int D = Sp->D, P = Sp->P, K = Sp->chunk, L2 = Sp->L2, cx_hi = Sp->cx_hi;
    const bool xr = (x >= D + 1 && x <= cx_hi);
    gfloat* samp = (gfloat*)Sp->samp;
    const int y0 = tile.y0, y1 = tile.y1;
    const int step0 = it.step0 + ordinal * T;
    const int k0 = ordinal * T;
    const SpanWin sw = (WIN & kWinSpan) ? span_window(Sp->span, y0 - T, ny) : SpanWin{};
    const YWin yw = (!BF && (WIN & kWinY)) ? ywin_load(lcy, y0 - T, ny, nc) : YWin{};
    kcint* span = (kcint*)Sp->span;
    auto load = [&](int r, YCell& c) {
        c.ez = c.hx = c.hy = c.jp = c.pey = c.phy = 0.f;
        c.mat = mat[min(max(r, 0), ny) * pitch + x];  // unconditional (see load_cell)
        if (r < 0 || r > ny) return;  // wave-uniform
        const int b = r * pitch + x;
        c.ez = fin[b];
        if (r < ny) {
            c.hx = fin[2 * ps + b];
            c.hy = fin[4 * ps + b];
        }
        if constexpr ((WIN & kWinSpan) != 0) {
            const Span sp = span_row(sw, r);
            if (x >= sp.lo && x < sp.hi) c.jp = jin[sp.off + x - sp.lo];
        } else {
            const int lo = span[3 * r], hi = span[3 * r + 1];
            if (x >= lo && x < hi) c.jp = jin[span[3 * r + 2] + x - lo];
        }
        if (r >= 1 && r <= nc) c.pey = pe_in[(r - 1) * pitch + x];
        else if (r >= ny - nc && r <= ny - 1) c.pey = pe_in[(nc + r - (ny - nc)) * pitch + x];
        if (r < nc) c.phy = ph_in[r * pitch + x];
        else if (r >= ny - nc && r < ny) c.phy = ph_in[(nc + r - (ny - nc)) * pitch + x];
    };
    YCell cur[T];
    float hxp[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        cur[t].ez = cur[t].hx = cur[t].hy = cur[t].jp = cur[t].pey = cur[t].phy = 0.f;
        cur[t].mat = 0;
        hxp[t] = 0.f;
    }
    auto store = [&](int yo, const YCell& c) {
            const int b = yo * pitch + x;
            fout[b] = c.ez;
            if (yo < ny) {
                fout[2 * ps + b] = c.hx;
                fout[4 * ps + b] = c.hy;
            }
            if constexpr ((WIN & kWinSpan) != 0) {
                const Span sp = span_row(sw, yo);
                if (x >= sp.lo && x < sp.hi) jout[sp.off + x - sp.lo] = c.jp;
            } else {
                const int lo = span[3 * yo], hi = span[3 * yo + 1];
                if (x >= lo && x < hi) jout[span[3 * yo + 2] + x - lo] = c.jp;
            }
            if (yo >= 1 && yo <= nc) pe_out[(yo - 1) * pitch + x] = c.pey;
            else if (yo >= ny - nc && yo <= ny - 1) pe_out[(nc + yo - (ny - nc)) * pitch + x] = c.pey;
            if (yo < nc) ph_out[yo * pitch + x] = c.phy;
            else if (yo >= ny - nc && yo < ny) ph_out[(nc + yo - (ny - nc)) * pitch + x] = c.phy;
    };
    // y-face NTFF samples of contour row y at step k of the chunk
    auto row_sample = [&](const YCell& o, int y, int k) {
        if (y == D) {
            samp[k * P + x - D - 1] = o.hx;
            samp[(K + k) * P + x - D - 1] = o.ez;
        } else if (y == D + 1) {
            samp[(2 * K + k) * P + x - D - 1] = o.ez;
        } else if (y == ny - D - 1) {
            samp[k * P + L2 + (cx_hi - x)] = o.hx;
            samp[(K + k) * P + L2 + (cx_hi - x)] = o.ez;
        } else {
            samp[(2 * K + k) * P + L2 + (cx_hi - x)] = o.ez;
        }
    };
    // FAST: every stage active (nst == T) and no source cell in reach of the tile,
    // so the stage loop carries no per-stage branch.
    auto stepg = [&](auto fast, int r, YCell in) {
        constexpr bool FAST = decltype(fast)::value;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int y = r - 1 - t;
            YCell c = cur[t];
            if constexpr (!BF) {
                if (FAST || t < nst) {
                    YCoef yc;
                    if constexpr ((WIN & kWinY) != 0) {
                        yc = ycoef(yw, y);
                    } else {
                        const int iye = (y >= 1 && y <= nc) ? nc - y : ((y >= ny - nc && y <= ny - 1) ? y - ny + nc : ID);
                        const int iyh = (y >= 0 && y < nc) ? nc - 1 - y : ((y >= ny - nc && y < ny) ? y - ny + nc : ID);
                        yc = YCoef{lcy[iye], lcy[2 * TOMO_MAX_NC + iye], lcy[3 * TOMO_MAX_NC + iye],
                                   lcy[TOMO_MAX_NC + iyh], lcy[4 * TOMO_MAX_NC + iyh], lcy[5 * TOMO_MAX_NC + iyh]};
                    }
                    const bool hok = (y >= 0 && y < ny), eok = (y >= 1 && y <= ny - 1);  // wave-uniform
                    const float dEx = lane_next(c.ez) - c.ez;
                    const float dEy = in.ez - c.ez;
                    c.phy = yc.bh * c.phy + yc.ah * dEy;
                    const float tx = dEy * yc.kh + dyf * c.phy;
                    const float hy = hok ? c.hy + chy * dEx : 0.f;
                    const float hx = hok ? c.hx - chx * tx : 0.f;
                    const float dHy = hy - lane_prev(hy);
                    const float dHx = hx - hxp[t];
                    hxp[t] = hx;
                    c.pey = yc.be * c.pey + yc.ae * dHx;
                    const float cx = dHx * yc.ke + dyf * c.pey;
                    const float4 qa = lc4[2 * c.mat], qb = lc4[2 * c.mat + 1];
                    float en = qa.x * c.ez + qa.y * dHy - qa.y * cx - qb.x * c.jp;
                    if (!FAST && wsrc && y == sy && x == sx) en += srcw[step0 + t];
                    c.jp = eok ? qb.y * c.jp + qb.z * (en - c.ez) : c.jp;
                    c.ez = eok ? en : 0.f;
                    c.hx = hx;
                    c.hy = hy;
                    if ((y == D || y == D + 1 || y == ny - D - 1 || y == ny - D) && y >= y0 && y < y1 && own && xr)
                        row_sample(c, y, k0 + t);
                }
            } else if (FAST || t < nst) {
                // row coefficients and node masks (packed y tables)
                const float4 ch = ly[kYTab + ytab_h(y, ny, nc)], ce = ly[ytab_e(y, ny, nc)];
                // node masks as factors (0 or 1): exact for the nodes that
                // exist, and no uniform branch splits the stage
                const float hm = ch.w, em = ce.w;
                const float dEx = lane_next(c.ez) - c.ez;
                const float dEy = in.ez - c.ez;
                c.phy = ch.z * c.phy + ch.y * dEy;
                const float tx = dEy * ch.x + dyf * c.phy;
                const float hy = (c.hy + chy * dEx) * hm;
                const float hx = (c.hx - chx * tx) * hm;
                const float dHy = hy - lane_prev(hy);
                const float dHx = hx - hxp[t];
                hxp[t] = hx;
                c.pey = ce.z * c.pey + ce.y * dHx;
                const float cx = dHx * ce.x + dyf * c.pey;
                // without an Ez node: Ez stays 0 and Jp keeps its value
                const float4 qa = lc4[2 * c.mat], qb = lc4[2 * c.mat + 1];
                const float ce_ = qa.x * em, ch_ = qa.y * em, cj_ = qb.x * em;
                const float kp_ = qb.y * em + (1.f - em), bp_ = qb.z * em;
                float en = ce_ * c.ez + ch_ * dHy - ch_ * cx - cj_ * c.jp;
                if (!FAST && wsrc && y == sy && x == sx) en += srcw[step0 + t] * em;
                c.jp = kp_ * c.jp + bp_ * (en - c.ez);
                c.ez = en;
                c.hx = hx;
                c.hy = hy;
            }
            cur[t] = in;
            in = c;
        }
        // branch-free form: y-face NTFF samples after the stage arithmetic;
        // stage t's output row is now cur[t + 1], the last stage's `in`
        if constexpr (BF) {
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int y = r - 1 - t;
                if ((FAST || t < nst) && (y == D || y == D + 1 || y == ny - D - 1 || y == ny - D) && y >= y0 &&
                    y < y1 && own && xr)
                    row_sample(t + 1 < T ? cur[t + 1] : in, y, k0 + t);
            }
        }
        return in;
    };
    auto step_fast = [&](int r, YCell in) { return stepg(std::true_type{}, r, in); };
    auto step_gen = [&](int r, YCell in) { return stepg(std::false_type{}, r, in); };
    const bool src_near = wsrc && sy >= y0 - 2 * T && sy < y1 + T;
    if (nst == T && !src_near) row_sweep<T, G, YCell>(y0, y1, own, load, store, step_fast);
    else row_sweep<T, G, YCell>(y0, y1, own, load, store, step_gen);

}

template <int T>
__global__ __launch_bounds__(256) void k_trow(const TileItem* __restrict__ tiles, int ntiles,
                                               const LaunchItem* __restrict__ items,
                                               const DevSim* __restrict__ sims, int ordinal) {
    __shared__ float4 lc4[TOMO_MAX_MAT * 2];
    constexpr bool BF = kStageBranchFree<T>;
    __shared__ float lcy[6 * TOMO_MAX_NC];
    __shared__ float4 ly[BF ? 2 * kYTab : 1];
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const TileItem tile = tiles[ti];
        const LaunchItem it = items[tile.item];
        const int nst = min(T, it.nsteps - ordinal * T);
        if (nst <= 0) continue;
        const DevSim* Sp = sims + it.sim;
        __syncthreads();
        fill_row<T, BF>(Sp, lc4, lcy, ly);
        __syncthreads();
        if ((int)(threadIdx.x / kWave) >= tile.nband) continue;  // after the last barrier of this tile
        body_row<T, kGroupRow, BF, kStageWin>(tile, it, nst, Sp, ordinal, lc4, lcy, ly);
    }
}


// Column-edge tiles of physical-mode simulations: bands that touch an x-side
// CPML strip, x-face NTFF column or x edge, over rows whose ghost zone is
// clean (no y strip, contour row or PEC row).  The x-strip coefficients and
// the existence masks are lane constants; every row is an x-face contour row,
// so contour columns sample at every stage.
struct XCell {
    float ez, hx, hy, jp, pex, phx;
    int mat;
};

template <int T>
__device__ __forceinline__ void fill_col(const DevSim* __restrict__ Sp, float4* __restrict__ lc4, float* __restrict__ lcx) {
    {
        float* lcf = (float*)lc4;
        for (int i = threadIdx.x; i < Sp->nmat * TOMO_NCOEF; i += blockDim.x) lcf[i] = Sp->coef[i];
    }
    for (int i = threadIdx.x; i < TOMO_MAX_NC; i += blockDim.x) {
        const bool id = (i >= Sp->nc);
        lcx[0 * TOMO_MAX_NC + i] = id ? 1.f : Sp->ike[i];
        lcx[1 * TOMO_MAX_NC + i] = id ? 1.f : Sp->ikm[i];
        lcx[2 * TOMO_MAX_NC + i] = id ? 0.f : Sp->ae[i];
        lcx[3 * TOMO_MAX_NC + i] = id ? 0.f : Sp->be[i];
        lcx[4 * TOMO_MAX_NC + i] = id ? 0.f : Sp->am[i];
        lcx[5 * TOMO_MAX_NC + i] = id ? 0.f : Sp->bm[i];
    }
}

template <int T, int G, int WIN>
__device__ __forceinline__ void body_col(const TileItem& tile, const LaunchItem& it, int nst,
                                         const DevSim* __restrict__ Sp, int ordinal,
                                         const float4* __restrict__ lc4, const float* __restrict__ lcx) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    constexpr int V = kWave - 2 * T;
    constexpr int ID = TOMO_MAX_NC - 1;
    const int par = it.parity0 ^ (ordinal & 1);
    const int ps = Sp->ps, js = Sp->js, xs = Sp->xs, pitch = Sp->pitch, nx = Sp->nx, nc = Sp->nc;
    gfloat* fin = (gfloat*)Sp->fbase + par * ps;
    gfloat* fout = (gfloat*)Sp->fbase + (par ^ 1) * ps;
    gfloat* jin = (gfloat*)Sp->jbase + par * js;
    gfloat* jout = (gfloat*)Sp->jbase + (par ^ 1) * js;
    gfloat* pe_in = (gfloat*)Sp->xbase + par * xs;
    gfloat* pe_out = (gfloat*)Sp->xbase + (par ^ 1) * xs;
    gfloat* ph_in = (gfloat*)Sp->xbase + (2 + par) * xs;
    gfloat* ph_out = (gfloat*)Sp->xbase + (2 + (par ^ 1)) * xs;
    gcu8* mat = (gcu8*)Sp->mat;
    const float chx = Sp->chx, chy = Sp->chy, dxf = Sp->dxf;
    const int bx = tile.x0 + wv * V;
    const int x0 = bx - T;
    const int x = x0 + lane;
    const bool xin = (x >= 0 && x <= nx);
    const bool own = xin && lane >= T && x < min(bx + V, tile.xe);
    const int sx = Sp->sx, sy = Sp->sy;
    const bool wsrc = (sx >= x0 && sx <= x0 + kWave - 1);
    kcfloat* srcw = (kcfloat*)Sp->srcw;
    // x-strip membership, psi storage column, coefficients, existence masks
    const int ixe = (x >= 1 && x <= nc) ? nc - x : ((x >= nx - nc && x <= nx - 1) ? x - nx + nc : ID);
    const int ixh = (x >= 0 && x < nc) ? nc - 1 - x : ((x >= nx - nc && x < nx) ? x - nx + nc : ID);
    const int je = (x >= 1 && x <= nc) ? x - 1 : ((x >= nx - nc && x <= nx - 1) ? nc + x - (nx - nc) : -1);
    const int jh = (x >= 0 && x < nc) ? x : ((x >= nx - nc && x < nx) ? nc + x - (nx - nc) : -1);
    const float kxe = lcx[ixe], kxh = lcx[TOMO_MAX_NC + ixh];
    const float axe = lcx[2 * TOMO_MAX_NC + ixe], bxe = lcx[3 * TOMO_MAX_NC + ixe];
    const float axh = lcx[4 * TOMO_MAX_NC + ixh], bxh = lcx[5 * TOMO_MAX_NC + ixh];
    const bool hok = (x >= 0 && x < nx), eok = (x >= 1 && x <= nx - 1);
    // x-face NTFF contour columns (FDTD_common.cxx:4-46 legs xp, xn)
    const int D = Sp->D, ny = Sp->ny, P = Sp->P, K = Sp->chunk;
    const bool ccol = (x == nx - D - 1 || x == nx - D || x == D || x == D + 1);
    const bool wcc = (x0 <= D + 1 && x0 + kWave - 1 >= D) || (x0 <= nx - D && x0 + kWave - 1 >= nx - D - 1);
    gfloat* samp = (gfloat*)Sp->samp;
    const int L1 = Sp->L1, L3 = Sp->L3, cy_lo = Sp->cy_lo;
    const int y0 = tile.y0, y1 = tile.y1;
    const int step0 = it.step0 + ordinal * T;
    const int k0 = ordinal * T;
    // Sample slots of a contour lane, as lane constants: the first value
    // (H on the face cells x = nx - D - 1 and x = D, Ez on their outer
    // neighbours) goes to samp[(sk + k) P + sb + sd y] with sk = 0 or 2K; the
    // face cells also store Ez K rows of P further on.  Leg xp runs up the
    // column from cy_lo, leg xn down from ny - D - 1.
    const bool s_face = (x == nx - D - 1 || x == D);
    const bool s_ez = !s_face;
    const int sd = (x >= nx - D - 1) ? 1 : -1;
    const int sb = (x >= nx - D - 1) ? L1 - cy_lo : L3 + ny - D - 1;
    const int s_off = (s_face ? 0 : 2 * K) * P + sb + k0 * P;
    const bool s_lane = ccol && own;
    const int ylo = max(max(y0, cy_lo), 0), yhi = min(y1, ny - D);  // rows of the contour's x legs
    const int w2 = 2 * nc;
    const SpanWin sw = (WIN & kWinSpan) ? span_window(Sp->span, y0 - T, ny) : SpanWin{};
    kcint* span = (kcint*)Sp->span;
    auto load = [&](int r, XCell& c) {
        c.ez = c.hx = c.hy = c.jp = c.pex = c.phx = 0.f;
        c.mat = mat[r * pitch + min(max(x, 0), nx)];  // unconditional (see load_cell)
        if (!xin) return;
        const int b = r * pitch + x;
        c.ez = fin[b];
        if (hok) {
            c.hx = fin[2 * ps + b];
            c.hy = fin[4 * ps + b];
        }
        if constexpr ((WIN & kWinSpan) != 0) {
            const Span sp = span_row(sw, r);
            if (x >= sp.lo && x < sp.hi) c.jp = jin[sp.off + x - sp.lo];
        } else {
            const int lo = span[3 * r], hi = span[3 * r + 1];
            if (x >= lo && x < hi) c.jp = jin[span[3 * r + 2] + x - lo];
        }
        if (je >= 0) c.pex = pe_in[r * w2 + je];
        if (jh >= 0) c.phx = ph_in[r * w2 + jh];
    };
    XCell cur[T];
    float hxp[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        cur[t].ez = cur[t].hx = cur[t].hy = cur[t].jp = cur[t].pex = cur[t].phx = 0.f;
        cur[t].mat = 0;
        hxp[t] = 0.f;
    }
    auto store = [&](int yo, const XCell& c) {
            const int b = yo * pitch + x;
            fout[b] = c.ez;
            if (hok) {
                fout[2 * ps + b] = c.hx;
                fout[4 * ps + b] = c.hy;
            }
            if constexpr ((WIN & kWinSpan) != 0) {
                const Span sp = span_row(sw, yo);
                if (x >= sp.lo && x < sp.hi) jout[sp.off + x - sp.lo] = c.jp;
            } else {
                const int lo = span[3 * yo], hi = span[3 * yo + 1];
                if (x >= lo && x < hi) jout[span[3 * yo + 2] + x - lo] = c.jp;
            }
            if (je >= 0) pe_out[yo * w2 + je] = c.pex;
            if (jh >= 0) ph_out[yo * w2 + jh] = c.phx;
    };
    // FAST: every stage active (nst == T) and no source cell in reach of the tile,
    // so the stage loop carries no per-stage branch.
    auto stepg = [&](auto fast, int r, XCell in) {
        constexpr bool FAST = decltype(fast)::value;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int y = r - 1 - t;
            XCell c = cur[t];
            if (FAST || t < nst) {
                const float dEx = lane_next(c.ez) - c.ez;
                const float dEy = in.ez - c.ez;
                c.phx = bxh * c.phx + axh * dEx;
                const float ty = dEx * kxh + dxf * c.phx;
                const float hy = hok ? c.hy + chy * ty : 0.f;
                const float hx = hok ? c.hx - chx * dEy : 0.f;
                const float dHy = hy - lane_prev(hy);
                const float dHx = hx - hxp[t];
                hxp[t] = hx;
                c.pex = bxe * c.pex + axe * dHy;
                const float cy = dHy * kxe + dxf * c.pex;
                const float4 qa = lc4[2 * c.mat], qb = lc4[2 * c.mat + 1];
                float en = qa.x * c.ez + qa.y * cy - qa.y * dHx - qb.x * c.jp;
                if (!FAST && wsrc && y == sy && x == sx) en += srcw[step0 + t];
                c.jp = eok ? qb.y * c.jp + qb.z * (en - c.ez) : c.jp;
                c.ez = eok ? en : 0.f;
                c.hx = hx;
                c.hy = hy;
            }
            cur[t] = in;
            in = c;
        }
        // x-face NTFF samples after the stage arithmetic (which thus stays one
        // basic block): stage t's output row is now cur[t + 1], the last
        // stage's is `in`.
        if (wcc) {
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int y = r - 1 - t;  // wave-uniform: the row test is a scalar branch
                if ((FAST || t < nst) && y >= ylo && y < yhi && s_lane) {
                    const XCell& o = t + 1 < T ? cur[t + 1] : in;
                    gfloat* dst = samp + (s_off + t * P + sd * y);
                    dst[0] = s_ez ? o.ez : o.hy;
                    if (s_face) dst[(size_t)K * P] = o.ez;
                }
            }
        }
        return in;
    };
    auto step_fast = [&](int r, XCell in) { return stepg(std::true_type{}, r, in); };
    auto step_gen = [&](int r, XCell in) { return stepg(std::false_type{}, r, in); };
    const bool src_near = wsrc && sy >= y0 - 2 * T && sy < y1 + T;
    if (nst == T && !src_near) row_sweep<T, G, XCell>(y0, y1, own, load, store, step_fast);
    else row_sweep<T, G, XCell>(y0, y1, own, load, store, step_gen);

}

template <int T>
__global__ __launch_bounds__(256) void k_tcol(const TileItem* __restrict__ tiles, int ntiles,
                                               const LaunchItem* __restrict__ items,
                                               const DevSim* __restrict__ sims, int ordinal) {
    __shared__ float4 lc4[TOMO_MAX_MAT * 2];
    __shared__ float lcx[6 * TOMO_MAX_NC];
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const TileItem tile = tiles[ti];
        const LaunchItem it = items[tile.item];
        const int nst = min(T, it.nsteps - ordinal * T);
        if (nst <= 0) continue;
        const DevSim* Sp = sims + it.sim;
        __syncthreads();
        fill_col<T>(Sp, lc4, lcx);
        __syncthreads();
        if ((int)(threadIdx.x / kWave) >= tile.nband) continue;  // after the last barrier of this tile
        body_col<T, kGroupCol, kStageWin>(tile, it, nst, Sp, ordinal, lc4, lcx);
    }
}


// Interior tiles of physical-mode simulations: every lane is inside the
// grid, no CPML strip, NTFF contour row/column or PEC edge lies within the
// tile's ghost zone, so each stage is the bare Yee/Debye update (plus the
// point source when the source cell is in the tile).  Two rows are kept in
// flight ahead of the sweep to cover HBM latency.
struct ICell {
    float ez, hx, hy, jp;
    int mat;
};

template <int T>
__device__ __forceinline__ void fill_inner(const DevSim* __restrict__ Sp, float4* __restrict__ lc4) {
    {
        float* lcf = (float*)lc4;
        for (int i = threadIdx.x; i < Sp->nmat * TOMO_NCOEF; i += blockDim.x) lcf[i] = Sp->coef[i];
    }
}

template <int T, int G, int WIN>
__device__ __forceinline__ void body_inner(const TileItem& tile, const LaunchItem& it, int nst,
                                         const DevSim* __restrict__ Sp, int ordinal,
                                         const float4* __restrict__ lc4) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    constexpr int V = kWave - 2 * T;
    const int par = it.parity0 ^ (ordinal & 1);
    const int ps = Sp->ps, js = Sp->js, pitch = Sp->pitch;
    gfloat* fin = (gfloat*)Sp->fbase + par * ps;
    gfloat* fout = (gfloat*)Sp->fbase + (par ^ 1) * ps;
    gfloat* jin = (gfloat*)Sp->jbase + par * js;
    gfloat* jout = (gfloat*)Sp->jbase + (par ^ 1) * js;
    gcu8* mat = (gcu8*)Sp->mat;
    const float chx = Sp->chx, chy = Sp->chy;
    const int x0 = tile.x0 + wv * V - T;
    const int x = x0 + lane;
    const bool own = lane >= T && lane < kWave - T;
    const int sx = Sp->sx, sy = Sp->sy;
    const bool wsrc = (sx >= x0 && sx <= x0 + kWave - 1);
    kcfloat* srcw = (kcfloat*)Sp->srcw;
    const int y0 = tile.y0, y1 = tile.y1;
    const int step0 = it.step0 + ordinal * T;
    // Row addresses are wave-uniform (scalar base + row * pitch); lanes add
    // their 32-bit byte offset, so loads and stores use the scalar-base
    // addressing form instead of per-lane 64-bit address arithmetic.
    const unsigned bx = (unsigned)x * 4u;  // x >= 0 on interior tiles
    typedef GAS char gbyte;
    auto at = [&](gfloat* base, unsigned off) -> gfloat& { return *(gfloat*)((gbyte*)base + off); };
    const SpanWin sw = (WIN & kWinSpan) ? span_window(Sp->span, y0 - T, Sp->ny) : SpanWin{};
    kcint* span = (kcint*)Sp->span;
    auto load = [&](int r, ICell& c) {
        gfloat* row = fin + (size_t)r * pitch;
        c.ez = at(row, bx);
        c.hx = at(row + 2 * (size_t)ps, bx);
        c.hy = at(row + 4 * (size_t)ps, bx);
        c.mat = (mat + (size_t)r * pitch)[(unsigned)x];
        c.jp = 0.f;
        if constexpr ((WIN & kWinSpan) != 0) {
            const Span sp = span_row(sw, r);
            if (x >= sp.lo && x < sp.hi) c.jp = at(jin + sp.off - sp.lo, bx);
        } else {
            const int lo = span[3 * r], hi = span[3 * r + 1];
            if (x >= lo && x < hi) c.jp = at(jin + span[3 * r + 2] - lo, bx);
        }
    };
    auto store = [&](int yo, const ICell& c) {
        gfloat* row = fout + (size_t)yo * pitch;
        at(row, bx) = c.ez;
        at(row + 2 * (size_t)ps, bx) = c.hx;
        at(row + 4 * (size_t)ps, bx) = c.hy;
        if constexpr ((WIN & kWinSpan) != 0) {
            const Span sp = span_row(sw, yo);
            if (x >= sp.lo && x < sp.hi) at(jout + sp.off - sp.lo, bx) = c.jp;
        } else {
            const int lo = span[3 * yo], hi = span[3 * yo + 1];
            if (x >= lo && x < hi) at(jout + span[3 * yo + 2] - lo, bx) = c.jp;
        }
    };
    ICell cur[T];
    float hxp[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        cur[t].ez = cur[t].hx = cur[t].hy = cur[t].jp = 0.f;
        cur[t].mat = 0;
        hxp[t] = 0.f;
    }
    // FAST: every stage active (nst == T) and no source cell in reach of the tile,
    // so the stage loop carries no per-stage branch.
    auto stepg = [&](auto fast, int r, ICell in) {
        constexpr bool FAST = decltype(fast)::value;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int y = r - 1 - t;
            ICell c = cur[t];
            if (FAST || t < nst) {
                // H half-step
                const float dEx = lane_next(c.ez) - c.ez;
                const float dEy = in.ez - c.ez;
                const float hy = c.hy + chy * dEx;
                const float hx = c.hx - chx * dEy;
                // E half-step with the Debye polarisation current
                const float dHy = hy - lane_prev(hy);
                const float dHx = hx - hxp[t];
                hxp[t] = hx;
                const float4 qa = lc4[2 * c.mat], qb = lc4[2 * c.mat + 1];
                float en = qa.x * c.ez + qa.y * dHy - qa.y * dHx - qb.x * c.jp;
                if (!FAST && wsrc && y == sy && x == sx) en += srcw[step0 + t];
                c.jp = qb.y * c.jp + qb.z * (en - c.ez);
                c.ez = en;
                c.hx = hx;
                c.hy = hy;
            }
            cur[t] = in;
            in = c;
        }
        return in;
    };
    auto step_fast = [&](int r, ICell in) { return stepg(std::true_type{}, r, in); };
    auto step_gen = [&](int r, ICell in) { return stepg(std::false_type{}, r, in); };
    const bool src_near = wsrc && sy >= y0 - 2 * T && sy < y1 + T;
    if (nst == T && !src_near) row_sweep<T, G, ICell>(y0, y1, own, load, store, step_fast);
    else row_sweep<T, G, ICell>(y0, y1, own, load, store, step_gen);

}

template <int T>
__global__ __launch_bounds__(256) void k_tinner(const TileItem* __restrict__ tiles, int ntiles,
                                               const LaunchItem* __restrict__ items,
                                               const DevSim* __restrict__ sims, int ordinal) {
    __shared__ float4 lc4[TOMO_MAX_MAT * 2];
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const TileItem tile = tiles[ti];
        const LaunchItem it = items[tile.item];
        const int nst = min(T, it.nsteps - ordinal * T);
        if (nst <= 0) continue;
        const DevSim* Sp = sims + it.sim;
        __syncthreads();
        fill_inner<T>(Sp, lc4);
        __syncthreads();
        if ((int)(threadIdx.x / kWave) >= tile.nband) continue;  // after the last barrier of this tile
        body_inner<T, kGroupInner, kStageWin>(tile, it, nst, Sp, ordinal, lc4);
    }
}


// All four tile classes of physical-mode simulations in one launch (tile
// class in TileItem::cls).  One launch per time block with no cross-stream
// joins: the form for few simulations, where every block's latency is on the
// critical path.  Classes are listed slowest first so their tiles start first.
// Workgroups per CU: kFusedWgPerCu.  Three for the select-form stages: the
// corner body alone needs 160 VGPRs, and four or five workgroups per CU
// (128 or 96 VGPRs) spill and run slower, although at three a 1024^2 grid's
// last interior tiles wait for a second round of workgroup slots.  Two for the
// branch-free T = 8 stages, whose corner body needs the registers.
template <int T>
__global__ __launch_bounds__(256, kFusedWgPerCu<T>) void k_tfused(const TileItem* __restrict__ tiles, int ntiles,
                                                const LaunchItem* __restrict__ items,
                                                const DevSim* __restrict__ sims, int ordinal) {
    __shared__ float4 lc4[TOMO_MAX_MAT * 2];
    constexpr bool BF = kFusedBranchFree<T>;
    __shared__ float ltab[8 * TOMO_MAX_NC];
    __shared__ float4 ly[BF ? 2 * kYTab : 1];
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const TileItem tile = tiles[ti];
        const LaunchItem it = items[tile.item];
        const int nst = min(T, it.nsteps - ordinal * T);
        if (nst <= 0) continue;
        const DevSim* Sp = sims + it.sim;
        const int cls = tile.cls;
        __syncthreads();
        if (cls == TILE_CORNER) fill_edge<T, BF>(Sp, lc4, ltab, ly);
        else if (cls == TILE_ROW) fill_row<T, BF>(Sp, lc4, ltab, ly);
        else if (cls == TILE_COL) fill_col<T>(Sp, lc4, ltab);
        else fill_inner<T>(Sp, lc4);
        __syncthreads();
        if ((int)(threadIdx.x / kWave) >= tile.nband) continue;  // after the last barrier of this tile
        // Boundary tiles set the launch's length (their stages carry the CPML
        // and contour work): their waves issue first on a shared SIMD.
        if (cls != TILE_INNER) __builtin_amdgcn_s_setprio(3);
        if (cls == TILE_CORNER) body_edge<T, kFusedGroupEdge, BF, kFusedWin>(tile, it, nst, Sp, ordinal, lc4, ltab, ly);
        else if (cls == TILE_ROW) body_row<T, kFusedGroupRow, BF, kFusedWin>(tile, it, nst, Sp, ordinal, lc4, ltab, ly);
        else if (cls == TILE_COL) body_col<T, kFusedGroupCol, kFusedWin>(tile, it, nst, Sp, ordinal, lc4, ltab);
        else body_inner<T, kFusedGroupInner, kFusedWin>(tile, it, nst, Sp, ordinal, lc4);
        __builtin_amdgcn_s_setprio(0);
    }
}

// ---------------------------------------------------------------- running DFT fold
// Phase e^{-j w (step + t0) dt}: evaluated in FP64 for every step of the chunk
// (<= 64 steps) into an LDS table and rounded once to FP32; the chunk is summed
// in FP32 and added to the long-running accumulator with Kahan compensation.
// Block (x, item, z) folds contour points [256 x, 256 x + 256) of one item
// for frequencies [z fpb, z fpb + fpb), kFoldGroup at a time.  Few items
// (one simulation) take fpb = kFoldGroup = 1, so the frequencies spread over
// more blocks; many items take fpb = all frequencies with kFoldGroup = 8, so
// each sample is read once per 8 frequencies.  Every frequency's sum runs in
// the same order either way.
template <bool FP64, int kFoldGroup>
__global__ __launch_bounds__(256) void k_fold(const LaunchItem* __restrict__ items,
                                              const DevSim* __restrict__ sims, int fpb) {
    const LaunchItem it = items[blockIdx.y];
    const DevSim* S = sims + it.sim;
    const int P = S->P, K = S->chunk;
    const int fb = blockIdx.z * fpb, F = min(S->nfreq, fb + fpb);
    const int nk = min(it.nsteps, K);
    if (nk <= 0 || fb >= F || ((S->flags & TOMO_FLAG_DFT_FP64) != 0) != FP64) return;
    const int nf = F - fb;
    __shared__ float2 wJ[FP64 ? 1 : TOMO_MAX_FREQ * 64];
    __shared__ float2 wM[FP64 ? 1 : TOMO_MAX_FREQ * 64];
    __shared__ double2 dJ[FP64 ? TOMO_MAX_FREQ * 64 : 1];
    __shared__ double2 dRot[FP64 ? TOMO_MAX_FREQ : 1];
    if (!FP64) {
        // every phase of the chunk evaluated in FP64 and rounded once to FP32
        // (a rotation recurrence would drift by an ulp per step)
        for (int i = threadIdx.x; i < 2 * nf * nk; i += blockDim.x) {
            const int which = i / (nf * nk), f = fb + (i / nk) % nf, k = i % nk;
            double s0, c0;
            sincos(-S->wdt[f] * ((double)it.step0 + k + (which ? S->tM : S->tJ)), &s0, &c0);
            (which ? wM : wJ)[f * 64 + k] = make_float2((float)c0, (float)s0);
        }
    } else {
        // FP64 study mode: exact phases; the M phase is the J phase rotated by
        // -w (tM - tJ).
        for (int i = threadIdx.x; i < nf * nk; i += blockDim.x) {
            const int f = fb + i / nk, k = i % nk;
            double s, c;
            sincos(-S->wdt[f] * ((double)it.step0 + k + S->tJ), &s, &c);
            dJ[f * 64 + k] = make_double2(c, s);
        }
        if (threadIdx.x < nf) {
            double s, c;
            sincos(-S->wdt[fb + threadIdx.x] * (S->tM - S->tJ), &s, &c);
            dRot[fb + threadIdx.x] = make_double2(c, s);
        }
    }
    __syncthreads();
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int leg = p < S->L1 ? 0 : (p < S->L2 ? 1 : (p < S->L3 ? 2 : 3));
    const float sdt = S->sgn_dt[leg];
    // global-address-space views (flat loads would also count against the LDS
    // counter, so every phase-table wait would drain the sample loads)
    gcfloat* hs = (gcfloat*)S->samp + p;
    gcfloat* e0 = (gcfloat*)S->samp + (size_t)K * P + p;
    gcfloat* e1 = (gcfloat*)S->samp + (size_t)2 * K * P + p;
    // Frequencies in groups of kFoldGroup: each sample is read once per group
    // (coalesced over p) and multiplied into every frequency of the group.
    for (int f0 = fb; f0 < F; f0 += kFoldGroup) {
        if (!FP64) {
            float jr[kFoldGroup], ji[kFoldGroup], mr[kFoldGroup], mi[kFoldGroup];
#pragma unroll
            for (int q = 0; q < kFoldGroup; ++q) jr[q] = ji[q] = mr[q] = mi[q] = 0.f;
            // unrolled so that several steps' samples are in flight at once
            // (the accumulation order is unchanged); one frequency per thread
            // leaves the registers for 20 steps' loads
#pragma unroll (kFoldGroup == 1 ? 20 : 8)
            for (int k = 0; k < nk; ++k) {
                const float h = hs[(size_t)k * P];
                const float e = (e0[(size_t)k * P] + e1[(size_t)k * P]) / 2;
#pragma unroll
                for (int q = 0; q < kFoldGroup; ++q) {
                    const float2 a = wJ[(f0 + q) * 64 + k], b = wM[(f0 + q) * 64 + k];
                    jr[q] += h * a.x;
                    ji[q] += h * a.y;
                    mr[q] += e * b.x;
                    mi[q] += e * b.y;
                }
            }
#pragma unroll
            for (int q = 0; q < kFoldGroup; ++q) {
                if (f0 + q >= F) break;
                float4* aJ = (float4*)S->accJ + (size_t)(f0 + q) * P + p;
                float4* aM = (float4*)S->accM + (size_t)(f0 + q) * P + p;
                float4 vJ = *aJ, vM = *aM;
                // Kahan: value (x, y), compensation (z, w)
                float yv, tv;
                yv = sdt * jr[q] - vJ.z; tv = vJ.x + yv; vJ.z = (tv - vJ.x) - yv; vJ.x = tv;
                yv = sdt * ji[q] - vJ.w; tv = vJ.y + yv; vJ.w = (tv - vJ.y) - yv; vJ.y = tv;
                yv = sdt * mr[q] - vM.z; tv = vM.x + yv; vM.z = (tv - vM.x) - yv; vM.x = tv;
                yv = sdt * mi[q] - vM.w; tv = vM.y + yv; vM.w = (tv - vM.y) - yv; vM.y = tv;
                *aJ = vJ;
                *aM = vM;
            }
        } else {
            double jr[kFoldGroup], ji[kFoldGroup], mr[kFoldGroup], mi[kFoldGroup];
#pragma unroll
            for (int q = 0; q < kFoldGroup; ++q) jr[q] = ji[q] = mr[q] = mi[q] = 0.0;
#pragma unroll 8
            for (int k = 0; k < nk; ++k) {
                const double h = hs[(size_t)k * P];
                const double e = 0.5 * ((double)e0[(size_t)k * P] + (double)e1[(size_t)k * P]);
#pragma unroll
                for (int q = 0; q < kFoldGroup; ++q) {
                    const double2 a = dJ[(f0 + q) * 64 + k], r = dRot[f0 + q];
                    const double2 b = make_double2(a.x * r.x - a.y * r.y, a.x * r.y + a.y * r.x);
                    jr[q] += h * a.x;
                    ji[q] += h * a.y;
                    mr[q] += e * b.x;
                    mi[q] += e * b.y;
                }
            }
#pragma unroll
            for (int q = 0; q < kFoldGroup; ++q) {
                if (f0 + q >= F) break;
                double2* aJ = (double2*)S->accJ + (size_t)(f0 + q) * P + p;
                double2* aM = (double2*)S->accM + (size_t)(f0 + q) * P + p;
                double2 vJ = *aJ, vM = *aM;
                vJ.x += sdt * jr[q]; vJ.y += sdt * ji[q]; vM.x += sdt * mr[q]; vM.y += sdt * mi[q];
                *aJ = vJ;
                *aM = vM;
            }
        }
    }
}

// ---------------------------------------------------------------- far-field projection
// Far-field projection on the matrix cores.  For one frequency the
// projection of every simulation of a group (same grid, contour and
// frequencies) is one complex matrix product
//     C[(s, c), a] = sum_p X[(s, c), p] W[p, a],   W[p, a] = e^{jk(x_p cos phi_a + y_p sin phi_a)},
// with four rows per simulation: c = 0 carries -del_p J_p (-> N), c = 1
// del_p M_p on y faces (-> L_y), c = 2 del_p M_p on x faces (-> L_x), c = 3
// zero.  A workgroup of four waves owns 32 simulations (8 per wave, one
// 32-row tile each) x kProjAngles angles.  The contour is walked in chunks of
// kProjChunk points: the workgroup evaluates the chunk's phase tile W once
// into LDS (double-buffered: the next chunk's tile and current values are
// prepared while this one runs), then every wave runs v_mfma_f32_32x32x2f32 over the chunk, two
// points at a time, four products per complex step (Re = XrWr - XiWi,
// Im = XrWi + XiWr) for each 32-angle tile, reusing its current values X
// across the angle tiles.  Operand layouts of the 32x32x2 f32 product:
// A[i][k] from lane i + 32k, B[k][j] from lane j + 32k, and accumulator
// register r of lane l holds C[8(r/4) + 4(l/32) + r%4][l%32], so a lane ends
// up holding all four rows of simulations 2q + l/32 (q = r/4) at angle l%32
// and forms D and the phase without leaving its registers.
// Few simulations (SPLIT): the four waves share 8 simulations and take one
// 32-angle tile each, so a lone simulation keeps four waves busy instead of
// one.  Every output element sees the same products in the same order in
// both layouts, so the results are identical.
typedef float f32x16 __attribute__((ext_vector_type(16)));
constexpr int kProjAngleTiles = 4;                 // 32-angle tiles per workgroup
constexpr int kProjAngles = 32 * kProjAngleTiles;
constexpr int kProjChunk = 16;                     // contour points per LDS phase tile
constexpr int kProjWaves = 4;                      // waves (8 simulations each) per workgroup

template <bool SPLIT>
__global__ __launch_bounds__(64 * kProjWaves) void k_project_mfma(const int* __restrict__ sim_idx, int n_sims,
                                                                  const DevSim* __restrict__ sims,
                                                                  const float2* __restrict__ kinv, int A,
                                                                  float eta0, float* __restrict__ D,
                                                                  float* __restrict__ Pph) {
    __shared__ float2 wt[2][kProjChunk][kProjAngles];  // phase tiles {cos, sin}, double-buffered
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int a0 = blockIdx.x * kProjAngles, f = blockIdx.z;
    constexpr int AT = SPLIT ? 1 : kProjAngleTiles;     // 32-angle tiles per wave
    const int s0 = SPLIT ? blockIdx.y * 8 : (blockIdx.y * kProjWaves + wv) * 8;  // this wave's first simulation
    const int mt0 = SPLIT ? wv : 0;                      // this wave's first angle tile
    const DevSim* G = sims + sim_idx[0];                 // geometry shared by the group
    const int P = G->P, L1 = G->L1, L2 = G->L2, L3 = G->L3, F = G->nfreq;
    const bool fp64 = (G->flags & TOMO_FLAG_DFT_FP64) != 0;
    const float2 ki = kinv[f];
    const float k = ki.x;
    const float dxf = G->dxf, dyf = G->dyf;
    const float* __restrict__ px = G->pos;
    const float* __restrict__ py = G->pos + P;
    // phase-tile filler: thread i covers angle a0 + i % kProjAngles for points (i / kProjAngles) + 2j
    const int fa = threadIdx.x % kProjAngles, fp = threadIdx.x / kProjAngles;
    float fc = 0.f, fs = 0.f;
    sincosf(6.283185307179586f * (float)(a0 + fa) / (float)A, &fs, &fc);
    // A operand lane: simulation s0 + (lane%32)/4, row kind (lane%32)%4, point p0 + lane/32
    const int kh = lane >> 5, rowk = lane & 3;
    const int sl = s0 + ((lane & 31) >> 2);
    const void* accX = nullptr;  // J for row kind 0, M for kinds 1 and 2
    if (sl < n_sims && rowk != 3) {
        const DevSim* S = sims + sim_idx[sl];
        accX = rowk == 0 ? S->accJ : S->accM;
    }
    f32x16 cre[AT], cim[AT];
#pragma unroll
    for (int m = 0; m < AT; ++m)
#pragma unroll
        for (int r = 0; r < 16; ++r) cre[m][r] = cim[m][r] = 0.f;
    // Chunk c's phase tile and current values are prepared while chunk c - 1
    // runs on the matrix cores: one barrier per chunk, and the current-value
    // loads stay in flight across the products (the scale `del` is applied
    // when they are consumed, so nothing waits on them earlier).
    auto fill = [&](int c0, int b) {
#pragma unroll
        for (int j = 0; j < kProjChunk / 2; ++j) {
            const int pp = fp + 2 * j, p = c0 + pp;
            float sn = 0.f, cs = 0.f;
            if (p < P && a0 + fa < A) sincosf(k * (px[p] * fc + py[p] * fs), &sn, &cs);
            wt[b][pp][fa] = make_float2(cs, sn);
        }
    };
    float rr[kProjChunk / 2], ri[kProjChunk / 2], sc[kProjChunk / 2];
    auto load = [&](int c0) {
#pragma unroll
        for (int j = 0; j < kProjChunk / 2; ++j) {
            const int p = c0 + 2 * j + kh;
            rr[j] = ri[j] = sc[j] = 0.f;
            if (p < P && accX) {
                const bool ysf = p < L1 || (p >= L2 && p < L3);
                if (rowk == 0 || (rowk == 1) == ysf) {
                    sc[j] = (ysf ? dxf : dyf) * (rowk == 0 ? -1.f : 1.f);
                    const size_t o = (size_t)f * P + p;
                    if (fp64) { const double2 v = ((const double2*)accX)[o]; rr[j] = (float)v.x; ri[j] = (float)v.y; }
                    else { const float2 v = ((const float2*)accX)[2 * o]; rr[j] = v.x; ri[j] = v.y; }
                }
            }
        }
    };
    float xr[kProjChunk / 2], xi[kProjChunk / 2];
    fill(0, 0);
    load(0);
#pragma unroll
    for (int j = 0; j < kProjChunk / 2; ++j) { xr[j] = sc[j] * rr[j]; xi[j] = sc[j] * ri[j]; }
    for (int c0 = 0, b = 0; c0 < P; c0 += kProjChunk, b ^= 1) {
        __syncthreads();  // tile b is complete; tile b ^ 1 is no longer read
        const bool more = c0 + kProjChunk < P;
        if (more) {
            load(c0 + kProjChunk);
            fill(c0 + kProjChunk, b ^ 1);
        }
#pragma unroll
        for (int j = 0; j < kProjChunk / 2; ++j) {
#pragma unroll
            for (int m = 0; m < AT; ++m) {
                const float2 w = wt[b][2 * j + kh][32 * (mt0 + m) + (lane & 31)];
                cre[m] = __builtin_amdgcn_mfma_f32_32x32x2f32(xr[j], w.x, cre[m], 0, 0, 0);
                cre[m] = __builtin_amdgcn_mfma_f32_32x32x2f32(-xi[j], w.y, cre[m], 0, 0, 0);
                cim[m] = __builtin_amdgcn_mfma_f32_32x32x2f32(xr[j], w.y, cim[m], 0, 0, 0);
                cim[m] = __builtin_amdgcn_mfma_f32_32x32x2f32(xi[j], w.x, cim[m], 0, 0, 0);
            }
        }
        if (more) {
#pragma unroll
            for (int j = 0; j < kProjChunk / 2; ++j) { xr[j] = sc[j] * rr[j]; xi[j] = sc[j] * ri[j]; }
        }
    }
    const float inv = ki.y;
#pragma unroll
    for (int m = 0; m < AT; ++m) {
        const int ab = a0 + 32 * (mt0 + m) + (lane & 31);
        if (ab >= A) continue;
        float cph, sph;
        sincosf(6.283185307179586f * (float)ab / (float)A, &sph, &cph);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int s = s0 + 2 * q + kh;
            if (s >= n_sims) continue;
            const float Nr = cre[m][4 * q], Ni = cim[m][4 * q];
            const float Lyr = cre[m][4 * q + 1], Lyi = cim[m][4 * q + 1];
            const float Lxr = cre[m][4 * q + 2], Lxi = cim[m][4 * q + 2];
            const float Lr = -sph * Lyr + cph * Lxr, Li = -sph * Lyi + cph * Lxi;
            const float Er = Lr + eta0 * Nr, Ei = Li + eta0 * Ni;
            const size_t o = ((size_t)s * F + f) * A + ab;
            D[o] = k * k * (Er * Er + Ei * Ei) * inv;
            if (Pph) Pph[o] = atan2f(Ei, Er);
        }
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

namespace {
// Events armed for the next timed launch of this thread (see arm_launch_events).
thread_local hipEvent_t t_start = nullptr, t_stop = nullptr;
}  // namespace

void arm_launch_events(hipEvent_t start, hipEvent_t stop) {
    t_start = start;
    t_stop = stop;
}

// A launch whose dispatch packet carries the armed start/stop events (none
// when nothing is armed); the events are consumed by this launch.
#define TIMED_LAUNCH(K, G, B, SHM, ST, ...)                                        \
    do {                                                                           \
        hipEvent_t ea_ = t_start, eb_ = t_stop;                                    \
        t_start = t_stop = nullptr;                                                \
        hipExtLaunchKernelGGL(K, G, B, SHM, ST, ea_, eb_, 0u, __VA_ARGS__);        \
    } while (0)

hipError_t launch_tblock(int T, int mode, const TileItem* tiles, int ntiles, const LaunchItem* items,
                         const DevSim* sims, int ordinal, int waves, int grid, hipStream_t st) {
    dim3 g(grid), b(waves * kWave);
#define TB(TT)                                                                                                 \
    if (T == TT) {                                                                                             \
        if (mode == TOMO_MODE_REF_COMPAT)                                                                      \
            TIMED_LAUNCH((k_tblock<TT, TOMO_MODE_REF_COMPAT>), g, b, 0, st, tiles, ntiles, items, sims,   \
                               ordinal);                                                                       \
        else                                                                                                   \
            TIMED_LAUNCH((k_tblock<TT, TOMO_MODE_PHYSICAL>), g, b, 0, st, tiles, ntiles, items, sims,     \
                               ordinal);                                                                       \
        return hipGetLastError();                                                                              \
    }
    TB(1) TB(2) TB(4) TB(6) TB(8)
#undef TB
    return hipErrorInvalidValue;
}

hipError_t launch_tedge(int T, const TileItem* tiles, int ntiles, const LaunchItem* items, const DevSim* sims,
                        int ordinal, int waves, int grid, hipStream_t st) {
    dim3 g(grid), b(waves * kWave);
#define TE(TT)                                                                                       \
    if (T == TT) {                                                                                   \
        TIMED_LAUNCH((k_tedge<TT>), g, b, 0, st, tiles, ntiles, items, sims, ordinal);         \
        return hipGetLastError();                                                                    \
    }
    TE(1) TE(2) TE(4) TE(6) TE(8)
#undef TE
    return hipErrorInvalidValue;
}

hipError_t launch_trow(int T, const TileItem* tiles, int ntiles, const LaunchItem* items, const DevSim* sims,
                       int ordinal, int waves, int grid, hipStream_t st) {
    dim3 g(grid), b(waves * kWave);
#define TR(TT)                                                                                       \
    if (T == TT) {                                                                                   \
        TIMED_LAUNCH((k_trow<TT>), g, b, 0, st, tiles, ntiles, items, sims, ordinal);          \
        return hipGetLastError();                                                                    \
    }
    TR(1) TR(2) TR(4) TR(6) TR(8)
#undef TR
    return hipErrorInvalidValue;
}

hipError_t launch_tcol(int T, const TileItem* tiles, int ntiles, const LaunchItem* items, const DevSim* sims,
                       int ordinal, int waves, int grid, hipStream_t st) {
    dim3 g(grid), b(waves * kWave);
#define TC(TT)                                                                                       \
    if (T == TT) {                                                                                   \
        TIMED_LAUNCH((k_tcol<TT>), g, b, 0, st, tiles, ntiles, items, sims, ordinal);          \
        return hipGetLastError();                                                                    \
    }
    TC(1) TC(2) TC(4) TC(6) TC(8)
#undef TC
    return hipErrorInvalidValue;
}

hipError_t launch_tinner(int T, const TileItem* tiles, int ntiles, const LaunchItem* items, const DevSim* sims,
                         int ordinal, int waves, int grid, hipStream_t st) {
    dim3 g(grid), b(waves * kWave);
#define TI(TT)                                                                                       \
    if (T == TT) {                                                                                   \
        TIMED_LAUNCH((k_tinner<TT>), g, b, 0, st, tiles, ntiles, items, sims, ordinal);        \
        return hipGetLastError();                                                                    \
    }
    TI(1) TI(2) TI(4) TI(6) TI(8)
#undef TI
    return hipErrorInvalidValue;
}

hipError_t launch_tfused(int T, const TileItem* tiles, int ntiles, const LaunchItem* items, const DevSim* sims,
                         int ordinal, int waves, int grid, hipStream_t st) {
    dim3 g(grid), b(waves * kWave);
#define TF(TT)                                                                                       \
    if (T == TT) {                                                                                   \
        TIMED_LAUNCH((k_tfused<TT>), g, b, 0, st, tiles, ntiles, items, sims, ordinal);        \
        return hipGetLastError();                                                                    \
    }
    TF(1) TF(2) TF(4) TF(6) TF(8)
#undef TF
    return hipErrorInvalidValue;
}

bool twave_supported(int T, int rows_per_wave) {
    return (T == 2 || T == 4 || T == 6 || T == 8) && (rows_per_wave == 4 || rows_per_wave == 5 || rows_per_wave == 6) &&
           kTileWaves * rows_per_wave > 2 * T;
}

hipError_t launch_twave(int T, int rows_per_wave, const TileItem* tiles, int ntiles, const LaunchItem* items,
                        const DevSim* sims, int ordinal, int grid, hipStream_t st) {
    dim3 g(grid), b(kTileWaves * kWave);
#define TW(TT, RR)                                                                                   \
    if (T == TT && rows_per_wave == RR) {                                                            \
        TIMED_LAUNCH((k_twave<TT, RR>), g, b, 0, st, tiles, ntiles, items, sims, ordinal);           \
        return hipGetLastError();                                                                    \
    }
    TW(2, 4) TW(2, 5) TW(2, 6) TW(4, 4) TW(4, 5) TW(4, 6) TW(6, 4) TW(6, 5) TW(6, 6) TW(8, 4) TW(8, 5) TW(8, 6)
#undef TW
    return hipErrorInvalidValue;
}

bool tres_supported(int T, int rows_per_wave) {
    return (T == 2 || T == 4 || T == 8) && (rows_per_wave == 8 || rows_per_wave == 12);
}

template <int T, int RW>
static const void* tres_fn(int mode) {
    return mode == TOMO_MODE_PHYSICAL ? reinterpret_cast<const void*>(&k_tres<T, RW, TOMO_MODE_PHYSICAL>)
                                      : reinterpret_cast<const void*>(&k_tres<T, RW, TOMO_MODE_REF_COMPAT>);
}

static const void* tres_kernel(int T, int RW, int mode) {
#define TR(TT, RR) \
    if (T == TT && RW == RR) return tres_fn<TT, RR>(mode);
    TR(2, 8) TR(2, 12) TR(4, 8) TR(4, 12) TR(8, 8) TR(8, 12)
#undef TR
    return nullptr;
}

int tres_max_tiles(int T, int rows_per_wave, int mode, int cus) {
    const void* f = tres_kernel(T, rows_per_wave, mode);
    if (!f) return 0;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kResWaves * kWave, 0) != hipSuccess) return 0;
    return per_cu * cus;
}

hipError_t launch_tres(int T, int rows_per_wave, int mode, const ResTile* tiles, int ntiles,
                       const LaunchItem* items, const DevSim* sims, int* done, int* err, int spin_limit,
                       hipStream_t st) {
    const void* f = tres_kernel(T, rows_per_wave, mode);
    if (!f) return hipErrorInvalidValue;
    t_start = t_stop = nullptr;  // a cooperative launch carries no events
    void* args[] = {(void*)&tiles, (void*)&items, (void*)&sims, (void*)&done, (void*)&err, (void*)&spin_limit};
    return hipLaunchCooperativeKernel(f, dim3(ntiles), dim3(kResWaves * kWave), args, 0, st);
}

hipError_t launch_project_mfma(const int* sim_idx, int n_sims, int n_freq, const DevSim* sims, const float2* kinv,
                               int n_angles, float eta0, float* D, float* Pph, hipStream_t st) {
    const bool split = n_sims < 8 * kProjWaves;
    const int per_wg = split ? 8 : 8 * kProjWaves;  // simulations per workgroup
    dim3 g((n_angles + kProjAngles - 1) / kProjAngles, (n_sims + per_wg - 1) / per_wg, n_freq), b(64 * kProjWaves);
    if (split) TIMED_LAUNCH(k_project_mfma<true>, g, b, 0, st, sim_idx, n_sims, sims, kinv, n_angles, eta0, D, Pph);
    else TIMED_LAUNCH(k_project_mfma<false>, g, b, 0, st, sim_idx, n_sims, sims, kinv, n_angles, eta0, D, Pph);
    return hipGetLastError();
}

hipError_t launch_fold(bool fp64, const LaunchItem* items, int nitems, const DevSim* sims, int max_p, int max_f,
                       hipStream_t st) {
    const int bx = (max_p + 255) / 256;
    // fewer than about one block per CU: spread the frequencies too
    const bool narrow = bx * nitems < 256;
    const int fpb = narrow ? 1 : max_f;
    dim3 g(bx, nitems, (max_f + fpb - 1) / fpb), b(256);
    if (fp64) {
        if (narrow) TIMED_LAUNCH((k_fold<true, 1>), g, b, 0, st, items, sims, fpb);
        else TIMED_LAUNCH((k_fold<true, 8>), g, b, 0, st, items, sims, fpb);
    } else {
        if (narrow) TIMED_LAUNCH((k_fold<false, 1>), g, b, 0, st, items, sims, fpb);
        else TIMED_LAUNCH((k_fold<false, 8>), g, b, 0, st, items, sims, fpb);
    }
    return hipGetLastError();
}

hipError_t launch_fitness(const float* D, const float* Pph, const float* tgt, const float* tph, int L, int n_eval,
                          double norm, int phase_term, double* fit, hipStream_t st) {
    TIMED_LAUNCH(k_fitness, dim3(n_eval), dim3(256), 0, st, D, Pph, tgt, tph, L, norm, phase_term, fit);
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
