// runtime.cpp — context, resident simulation slots, the epoch runner and the
// C ABI of include/tomo.h.
//
// An epoch advances every active simulation by up to `chunk` steps:
// ceil(chunk / T) launches of the time-blocked update (T steps each, one tile
// list spanning all simulations) followed by one running-DFT fold.  Plain
// evaluation (tomo_eval, the FDTD_GPU() twin at FDTD_GPU.cu:650-1110) runs one
// simulation at a time; batched evaluation and the request loop put every
// resident simulation into the same epochs.  Steady-state epochs of the
// batched path are captured once into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "kernels.h"
#include "setup.h"

using namespace tomo;

namespace {

#define HIPCHK(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(TOMO_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Device buffer freed when it leaves scope (so early error returns leak nothing).
struct DevMem {
    void* p = nullptr;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes); }
    template <class X>
    X* as() const { return (X*)p; }
};

struct Slot {
    SimSetup su;
    DevSim d{};
    char* mem = nullptr;
    size_t bytes = 0;
    int idx = -1;       // DevSim table index
    int done = 0;       // steps completed
    int parity = 0;     // buffer holding the current state
    int target = 0;     // steps to reach in the current run
    long req = -1;      // owning request (serving)
    int exc = 0;
    uint64_t serial = 0; // unique per simulation (table indices are reused)
    // cell accounting of the current tile split (for kernel timing): cells and
    // algorithmic bytes per step of the corner, row-edge, column-edge and
    // interior tile classes
    double cls_cells[4] = {0, 0, 0, 0}, cls_bytes[4] = {0, 0, 0, 0};
    std::vector<float> hpos;  // host copy of the contour positions (source of an async upload)
};

// Bytes one simulation slot occupies on the device.
size_t slot_bytes(const SimSetup& s, int chunk) {
    const tomo_sim_config& c = s.cfg;
    const size_t rows = (size_t)c.ny + 1, pitch = s.pitch, nc = c.ncpml, P = s.P;
    size_t b = 0;
    b += 6 * align256(rows * pitch * 4);                      // ez, hx, hy x2
    b += 2 * align256((size_t)std::max(s.jp_len, 1) * 4);     // jp x2
    b += 4 * align256(rows * 2 * nc * 4);                     // pex, phx x2
    b += 4 * align256(2 * nc * pitch * 4);                    // pey, phy x2
    b += align256(rows * pitch);                              // mat
    b += align256(3 * rows * 4);                              // span
    b += align256(3 * (size_t)chunk * P * 4);                 // samples
    b += 2 * align256((size_t)c.n_freq * P * 16);             // accJ, accM
    b += align256((size_t)std::max(c.steps, 1) * 4);          // source waveform
    b += align256(2 * P * 4);                                 // contour positions
    b += align256((2 * rows + (size_t)c.nx + 1 + rows) * 4);  // ref raster inputs
    return b;
}

}  // namespace

// A tomo_sim handle and its context share this record.  Destroying the context
// first frees the simulation's slot and clears `ctx`, so a later call on the
// handle sees a dead context instead of touching freed memory.
struct SimLink {
    tomo_ctx* ctx = nullptr;
    Slot* s = nullptr;
};

namespace {
// Launch staging of the epoch runner: launch items and the tile list (host
// copies and device copies) and the DevSim table the kernels index.  The
// context's own staging grows on demand (pinned host memory, device memory);
// a fixed staging lives in caller-provided memory and never grows.
struct Staging {
    LaunchItem* h_items = nullptr;
    LaunchItem* d_items = nullptr;
    int items_cap = 0;
    TileItem* h_tiles = nullptr;
    TileItem* d_tiles = nullptr;
    int tiles_cap = 0;
    DevSim* d_sims = nullptr;
    // host staging may be rewritten once this completes (growable staging);
    // a fixed staging waits for its stream instead
    hipEvent_t staged = nullptr;
    bool fixed = false;
    std::vector<uint64_t> tile_key;  // serials of the simulations the uploaded tile list was built for
    int key_T = 0, key_S = 0, key_W = 0, key_Wi = 0, key_ER = 0, key_Si = 0;
    int ntiles = 0, n_edge = 0;  // tiles: n_edge edge tiles first, then interior tiles
    int n_rowedge = 0;           // edge tiles: n_corner general ones, then row-edge, then column-edge
    int n_corner = 0;
    // resident form: tiles, per-tile completed-block counters and the error word
    ResTile* h_res = nullptr;
    ResTile* d_res = nullptr;
    int res_cap = 0, nres = 0;
    int* d_flags = nullptr;  // res_cap counters, then the error word (res_err)
    int res_key_form = 0;    // form the tile lists were built for
};
}  // namespace

struct tomo_request {
    uint64_t ticket = 0;
    tomo_sim_config cfg{};
    std::vector<double> params;
    std::vector<Slot*> slots;
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed, 4 cancelled
    int status = TOMO_OK;
    double t_submit = 0, t_done = 0, fitness = 0;
    int64_t bytes = 0;
};

struct tomo_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    char* scratch = nullptr;  // device address of the allocation-free entry's array on this device
    int cus = 256;
    uint64_t budget = 0, used = 0;
    int T = 4, S = 32, W = 4, K = 32;  // schedule (W: waves per edge workgroup)
    int grid_mult = 4;                 // persistent workgroups per CU
    int form = TOMO_FORM_STREAMS;      // launch form of physical-mode blocks
    int ER = 8;                        // output rows per corner / row-edge tile
    int Si = 0;                        // output rows per interior tile (0: S)
    bool use_graph = false;            // replay regular epochs from a captured hipGraph
    bool res_refused = false;          // the device refused a cooperative launch: resident epochs run fused
    hipGraphExec_t gexec = nullptr;    // the captured epoch and what it was captured for
    std::vector<uint64_t> gkey_serials;
    std::vector<int64_t> gkey;
    // DevSim table
    DevSim* d_sims = nullptr;
    int cap = 0;
    std::vector<int> free_idx;
    // epoch staging (its d_sims is the table above)
    Staging stg;
    uint64_t next_serial = 1;
    int Wi = 4;                  // waves per workgroup of the interior kernel
    // serving
    std::mutex mu;
    uint64_t next_ticket = 1;
    std::deque<uint64_t> queue;
    std::map<uint64_t, std::unique_ptr<tomo_request>> reqs;
    std::vector<Slot*> active;
    // simulations created by tomo_sim_create (their slots are in neither `active` nor a request)
    std::vector<std::shared_ptr<SimLink>> sims;
    // physical-mode targets, keyed by configuration bytes
    std::map<std::string, std::vector<float>> targets;
    // device memory of destroyed slots, kept for reuse by size
    std::map<size_t, std::vector<char*>> cache;
    size_t cached_bytes = 0;
    // side streams: the three edge-tile kernels run beside the interior one;
    // ev_side[k] marks side stream k's block done, ev_bar all of a block done
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev_side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev_bar = nullptr;
    // kernel timing
    bool prof = false;
    tomo_stats stats{};
    std::vector<hipEvent_t> ev_pool;
    // kind: TOMO_KERNEL_*, or -1 = update phase (borrows the events of its first and last launches)
    struct Pending { hipEvent_t a, b; int kind; double cells, bytes; bool owns = true; };
    std::vector<Pending> pending;
};

namespace {

std::string cfg_key(const tomo_sim_config& c) {
    tomo_sim_config k = c;
    k.n_params = 0;
    return std::string((const char*)&k, sizeof(k));
}

int ensure_table(tomo_ctx* c, int need) {
    if (need <= c->cap) return TOMO_OK;
    int ncap = std::max(need, std::max(256, c->cap * 2));
    DevSim* nd = nullptr;
    HIPCHK(hipMalloc(&nd, sizeof(DevSim) * ncap));
    if (c->d_sims) {
        HIPCHK(hipMemcpyAsync(nd, c->d_sims, sizeof(DevSim) * c->cap, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipFree(c->d_sims));
    }
    for (int i = ncap - 1; i >= c->cap; --i) c->free_idx.push_back(i);
    c->d_sims = nd;
    c->cap = ncap;
    return TOMO_OK;
}

int ensure_res_staging(tomo_ctx* c, Staging& st, int nres) {
    if (nres <= st.res_cap) return TOMO_OK;
    if (st.fixed) return fail(TOMO_E_NOMEM, "resident tiles exceed the fixed staging");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (st.h_res) HIPCHK(hipHostFree(st.h_res));
    if (st.d_res) HIPCHK(hipFree(st.d_res));
    if (st.d_flags) HIPCHK(hipFree(st.d_flags));
    const int cap = std::max(nres, 2 * st.res_cap);
    st.h_res = nullptr;
    st.d_res = nullptr;
    st.d_flags = nullptr;
    st.res_cap = 0;
    HIPCHK(hipHostMalloc(&st.h_res, sizeof(ResTile) * cap, hipHostMallocDefault));
    HIPCHK(hipMalloc(&st.d_res, sizeof(ResTile) * cap));
    HIPCHK(hipMalloc(&st.d_flags, sizeof(int) * (cap + 64)));
    st.res_cap = cap;
    st.tile_key.clear();
    return TOMO_OK;
}

int ensure_staging(tomo_ctx* c, Staging& st, int nitems, int ntiles) {
    if (st.fixed) {
        if (nitems > st.items_cap || ntiles > st.tiles_cap)
            return fail(TOMO_E_NOMEM, "launch items or tiles exceed the fixed staging");
        return TOMO_OK;
    }
    if (nitems > st.items_cap) {
        HIPCHK(hipStreamSynchronize(c->stream));
        if (st.h_items) HIPCHK(hipHostFree(st.h_items));
        if (st.d_items) HIPCHK(hipFree(st.d_items));
        const int cap = std::max(nitems, 2 * st.items_cap);
        st.h_items = nullptr;
        st.d_items = nullptr;
        st.items_cap = 0;
        HIPCHK(hipHostMalloc(&st.h_items, sizeof(LaunchItem) * cap, hipHostMallocDefault));
        HIPCHK(hipMalloc(&st.d_items, sizeof(LaunchItem) * cap));
        st.items_cap = cap;
    }
    if (ntiles > st.tiles_cap) {
        HIPCHK(hipStreamSynchronize(c->stream));
        if (st.h_tiles) HIPCHK(hipHostFree(st.h_tiles));
        if (st.d_tiles) HIPCHK(hipFree(st.d_tiles));
        const int cap = std::max(ntiles, 2 * st.tiles_cap);
        st.h_tiles = nullptr;
        st.d_tiles = nullptr;
        st.tiles_cap = 0;
        HIPCHK(hipHostMalloc(&st.h_tiles, sizeof(TileItem) * cap, hipHostMallocDefault));
        HIPCHK(hipMalloc(&st.d_tiles, sizeof(TileItem) * cap));
        st.tiles_cap = cap;
        st.tile_key.clear();
    }
    return TOMO_OK;
}

void release_cache(tomo_ctx* c) {
    (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->cache)
        for (char* p : kv.second) (void)hipFree(p);
    c->cache.clear();
    c->cached_bytes = 0;
}

void slot_destroy(tomo_ctx* c, Slot* s);
int slot_fill(tomo_ctx* c, Slot* s, DevSim* d_entry);

// Build the device-resident simulation for one excitation of one phantom.
int slot_create(tomo_ctx* c, const tomo_sim_config& cfg, const double* p, size_t np, int exc, Slot** out) {
    std::unique_ptr<Slot> s(new Slot());
    int rc = build_setup(cfg, p, np, exc, s->su);
    if (rc) return rc;
    const SimSetup& su = s->su;
    if (c->K % c->T) return fail(TOMO_E_INVALID, "chunk must be a multiple of the fused step count");
    const size_t bytes = slot_bytes(su, c->K);
    if (c->used + bytes > c->budget) return fail(TOMO_E_NOMEM, "HBM budget exceeded");
    auto hit = c->cache.find(bytes);
    if (hit != c->cache.end() && !hit->second.empty()) {
        s->mem = hit->second.back();
        hit->second.pop_back();
        c->cached_bytes -= bytes;
    } else {
        if (hipMalloc(&s->mem, bytes) != hipSuccess) {
            (void)hipGetLastError();
            release_cache(c);
            HIPCHK(hipMalloc(&s->mem, bytes));
        }
    }
    s->bytes = bytes;
    c->used += bytes;
    // every error return from here on hands the memory, the budget and the
    // table index back through slot_destroy
    struct Undo {
        tomo_ctx* c;
        std::unique_ptr<Slot>& s;
        ~Undo() {
            if (s) slot_destroy(c, s.release());
        }
    } undo{c, s};
    rc = ensure_table(c, 1);
    if (rc) return rc;
    if (c->free_idx.empty()) {
        rc = ensure_table(c, c->cap + 1);
        if (rc) return rc;
    }
    s->idx = c->free_idx.back();
    c->free_idx.pop_back();
    s->serial = c->next_serial++;
    rc = slot_fill(c, s.get(), c->d_sims + s->idx);
    if (rc) return rc;
    s->exc = exc;
    s->target = su.cfg.steps;
    *out = s.release();
    return TOMO_OK;
}

// Lay out a slot in s->mem (s->bytes), zero its state and accumulators, upload
// its tables and rasterise its material map, and write its DevSim to d_entry
// (all on the context's stream; the uploads read host memory owned by the
// slot, which must outlive them).
int slot_fill(tomo_ctx* c, Slot* s, DevSim* d_entry) {
    const SimSetup& su = s->su;
    const size_t bytes = s->bytes;
    const tomo_sim_config& cf = su.cfg;
    const size_t rows = (size_t)cf.ny + 1, pitch = su.pitch, nc = cf.ncpml, P = su.P;
    char* m = s->mem;
    auto take = [&](size_t b) { char* r = m; m += align256(b); return r; };
    DevSim& d = s->d;
    for (int k = 0; k < 2; ++k) d.ez[k] = (float*)take(rows * pitch * 4);
    for (int k = 0; k < 2; ++k) d.hx[k] = (float*)take(rows * pitch * 4);
    for (int k = 0; k < 2; ++k) d.hy[k] = (float*)take(rows * pitch * 4);
    for (int k = 0; k < 2; ++k) d.jp[k] = (float*)take((size_t)std::max(su.jp_len, 1) * 4);
    for (int k = 0; k < 2; ++k) d.pex[k] = (float*)take(rows * 2 * nc * 4);
    for (int k = 0; k < 2; ++k) d.phx[k] = (float*)take(rows * 2 * nc * 4);
    for (int k = 0; k < 2; ++k) d.pey[k] = (float*)take(2 * nc * pitch * 4);
    for (int k = 0; k < 2; ++k) d.phy[k] = (float*)take(2 * nc * pitch * 4);
    char* state_end = m;
    d.fbase = d.ez[0];
    d.ps = (int)(d.ez[1] - d.ez[0]);
    d.jbase = d.jp[0];
    d.js = (int)(d.jp[1] - d.jp[0]);
    d.xbase = d.pex[0];
    d.xs = (int)(d.pex[1] - d.pex[0]);
    d.ybase = d.pey[0];
    d.ys = (int)(d.pey[1] - d.pey[0]);
    if (d.hx[0] != d.fbase + 2 * (size_t)d.ps || d.hy[1] != d.fbase + 5 * (size_t)d.ps ||
        d.phx[1] != d.xbase + 3 * (size_t)d.xs || d.phy[1] != d.ybase + 3 * (size_t)d.ys)
        return fail(TOMO_E_INVALID, "internal: slot planes are not equally strided");
    uint8_t* mat = (uint8_t*)take(rows * pitch);
    int* span = (int*)take(3 * rows * 4);
    d.samp = (float*)take(3 * (size_t)c->K * P * 4);
    char* acc0 = m;
    d.accJ = take((size_t)cf.n_freq * P * 16);
    d.accM = take((size_t)cf.n_freq * P * 16);
    char* acc1 = m;
    float* srcw = (float*)take((size_t)std::max(cf.steps, 1) * 4);
    float* pos = (float*)take(2 * P * 4);
    int* rr = (int*)take((2 * rows + (size_t)cf.nx + 1 + rows) * 4);
    if ((size_t)(m - s->mem) > bytes) return fail(TOMO_E_INVALID, "internal: slot layout overflow");
    d.mat = mat;
    d.span = span;
    d.srcw = cf.mode == TOMO_MODE_PHYSICAL ? srcw : nullptr;
    d.pos = pos;
    d.nx = cf.nx; d.ny = cf.ny; d.pitch = su.pitch; d.nc = cf.ncpml;
    d.D = cf.ntff_offset; d.P = su.P; d.L1 = su.L1; d.L2 = su.L2; d.L3 = su.L3;
    d.cx_hi = su.cx_hi; d.cy_lo = su.cy_lo;
    d.sx = su.sx; d.sy = su.sy; d.mode = cf.mode; d.flags = cf.flags;
    d.nmat = su.nmat; d.nfreq = cf.n_freq; d.chunk = c->K; d.steps = cf.steps;
    std::memcpy(d.coef, su.coef, sizeof(d.coef));
    for (int i = 0; i < TOMO_MAX_NC; ++i) {
        d.ike[i] = su.ike[i]; d.ae[i] = su.ae[i]; d.be[i] = su.be[i];
        d.ikm[i] = su.ikm[i]; d.am[i] = su.am[i]; d.bm[i] = su.bm[i];
        d.ke[i] = su.ke[i]; d.km[i] = su.km[i];
    }
    d.chx = su.chx; d.chy = su.chy; d.chx_d = su.chx_d; d.chy_d = su.chy_d;
    d.dxf = (float)cf.dx; d.dyf = (float)cf.dy;
    d.inc_dt = su.inc_dt; d.inc_c = su.inc_c; d.inc_ax = su.inc_ax; d.inc_ay = su.inc_ay;
    d.inc_itau = su.inc_itau; d.inc_abs_origin = su.inc_abs;
    for (int f = 0; f < TOMO_MAX_FREQ; ++f) d.wdt[f] = su.wdt[f];
    d.tJ = su.tJ; d.tM = su.tM;
    const float dtf = (float)cf.dt;
    // calculate_JandM signs (FDTD_GPU.cu:251-276): yn +, xp +, yp -, xn -
    d.sgn_dt[0] = dtf; d.sgn_dt[1] = dtf; d.sgn_dt[2] = -dtf; d.sgn_dt[3] = -dtf;

    // contour positions (N2FPostProcess geometry, FDTD_GPU.cu:584-621)
    std::vector<float>& hpos = s->hpos;
    hpos.resize(2 * P);
    {
        const int D = cf.ntff_offset, nx = cf.nx, ny = cf.ny;
        for (int p = 0; p < (int)P; ++p) {
            int x, y, leg;
            if (p < su.L1) { x = p + D + 1; y = D; leg = 0; }
            else if (p < su.L2) { x = nx - D - 1; y = (p - su.L1) + su.cy_lo; leg = 1; }
            else if (p < su.L3) { x = su.cx_hi - (p - su.L2); y = ny - D - 1; leg = 2; }
            else { x = D; y = ny - D - (p - su.L3) - 1; leg = 3; }
            double fx = x, fy = y;
            if (leg == 0 || leg == 2) fy += 0.5; else fx += 0.5;
            double px, py;
            if (cf.mode == TOMO_MODE_REF_COMPAT) {
                px = cf.dx * (-(double)(float)nx / 2 + 1 + fx);
                py = cf.dy * (-(double)(float)ny / 2 + 1 + fy);
            } else {
                px = cf.dx * (fx - nx / 2.0);
                py = cf.dy * (fy - ny / 2.0);
            }
            hpos[p] = (float)px;
            hpos[P + p] = (float)py;
        }
    }
    hipStream_t st = c->stream;
    // zero the dynamic state and accumulators
    HIPCHK(launch_zero(s->mem, (size_t)(state_end - s->mem), st));
    HIPCHK(launch_zero(acc0, (size_t)(acc1 - acc0), st));
    HIPCHK(hipMemcpyAsync(span, su.span.data(), 3 * rows * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(pos, hpos.data(), 2 * P * 4, hipMemcpyHostToDevice, st));
    if (cf.mode == TOMO_MODE_PHYSICAL) {
        HIPCHK(hipMemcpyAsync(srcw, su.srcw.data(), (size_t)cf.steps * 4, hipMemcpyHostToDevice, st));
        HIPCHK(launch_raster_phys(mat, su.pitch, cf.nx, cf.ny, (float)cf.dx, (float)cf.dy, su.geom, st));
    } else {
        int* lo = rr;
        int* hi = rr + rows;
        int* xc = rr + 2 * rows;
        int* yc = xc + cf.nx + 1;
        HIPCHK(hipMemcpyAsync(lo, su.ref.lo.data(), rows * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(hi, su.ref.hi.data(), rows * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(xc, su.ref.xc.data(), ((size_t)cf.nx + 1) * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(yc, su.ref.yc.data(), rows * 4, hipMemcpyHostToDevice, st));
        HIPCHK(launch_raster_ref(mat, su.pitch, cf.nx, cf.ny, lo, hi, xc, yc, cf.numpatches, st));
    }
    HIPCHK(hipMemcpyAsync(d_entry, &s->d, sizeof(DevSim), hipMemcpyHostToDevice, st));
    // the uploads above read host memory owned by the slot, which outlives them
    // (slot_destroy waits for the stream before releasing it)
    return TOMO_OK;
}

// Return a slot's device memory to the cache (work still queued on the
// stream may use it: it is handed out again only to later stream work).
void slot_destroy(tomo_ctx* c, Slot* s) {
    if (!s) return;
    if (s->mem) {
        c->cache[s->bytes].push_back(s->mem);
        c->cached_bytes += s->bytes;
        c->used -= s->bytes;
        if (c->used + c->cached_bytes > c->budget) release_cache(c);
    }
    if (s->idx >= 0) c->free_idx.push_back(s->idx);
    // the host staging in s (span, waveform, positions) may still be read by
    // queued uploads: wait before releasing it
    (void)hipStreamSynchronize(c->stream);
    delete s;
}

int n_bands(int nx, int T) { return (nx + 1 + (64 - 2 * T) - 1) / (64 - 2 * T); }

// Tiles of one simulation.  Columns and rows whose ghost zone touches a PEC
// edge, a CPML strip or an NTFF contour row/column go to the edge kernel;
// full-width bands of the clean interior go to the interior kernel (physical
// mode only).  Interior bands start exactly where the clean zone starts, so
// the edge frame is one narrow band wide on each side.
// Algorithmic bytes of the cells a tile owns inside [0, nx) x [0, ny), added
// to its class: Ez, Hx, Hy read and written once (24) and the material id (1)
// on every cell, Jp (8) on Debye cells, the psi pair (16) on CPML-strip cells.
void account_tile(Slot* sl, int k, const TileItem& t) {
    const DevSim& d = sl->d;
    const int nx = d.nx, ny = d.ny, nc = d.nc;
    const int x0 = std::max(t.x0, 0), x1 = std::min(t.xe, nx);
    const int y0 = std::max(t.y0, 0), y1 = std::min(t.y1, ny);
    if (x1 <= x0 || y1 <= y0) return;
    const double cells = (double)(x1 - x0) * (y1 - y0);
    double jp = 0;
    for (int y = y0; y < y1; ++y)
        jp += std::max(0, std::min(x1, sl->su.span[3 * y + 1]) - std::max(x0, sl->su.span[3 * y]));
    const double cx = std::max(0, std::min(x1, nx - nc) - std::max(x0, nc));
    const double cy = std::max(0, std::min(y1, ny - nc) - std::max(y0, nc));
    sl->cls_cells[k] += cells;
    sl->cls_bytes[k] += 25.0 * cells + 8.0 * jp + 16.0 * (cells - cx * cy);
}

void build_tiles(Slot* sl, int item, int T, int S, int Si, int W, int Wi, int ER, std::vector<TileItem>& edge,
                 std::vector<TileItem>& rowedge, std::vector<TileItem>& coledge, std::vector<TileItem>& inner) {
    const DevSim& d = sl->d;
    for (int k = 0; k < 4; ++k) sl->cls_cells[k] = sl->cls_bytes[k] = 0;
    const int nx = d.nx, ny = d.ny, nc = d.nc, D = d.D, V = 64 - 2 * T;
    // A tile streams its output rows plus T ghost rows on each side through
    // the row sweep, at most 64 (the sweeps hold the rows' dispersive spans
    // one row per lane): taller segments are split.
    S = std::min(S, kMaxTileRows - 2 * T);
    Si = std::min(Si, kMaxTileRows - 2 * T);
    ER = std::min(ER, kMaxTileRows - 2 * T);
    auto account = [&](int k, const TileItem& t) { account_tile(sl, k, t); };
    // column ranges [a, b) split into bands of V and grouped w bands per tile
    auto add = [&](int k, std::vector<TileItem>& v, int a, int b, int y0, int y1, int w) {
        for (int x = a; x < b; x += w * V) {
            TileItem t{};
            t.item = item;
            t.x0 = x;
            t.xe = std::min(x + w * V, b);
            t.nband = (t.xe - x + V - 1) / V;
            t.y0 = y0;
            t.y1 = y1;
            t.cls = k;  // TILE_CORNER, TILE_ROW, TILE_COL, TILE_INNER
            v.push_back(t);
            account(k, t);
        }
    };
    // clean rows / columns lie strictly between the special ones
    const int ybot = std::max(nc, D + 1) + 1, ytop = std::min(ny - nc, ny - D - 1) - 1;
    const int xbot = std::max(nc, D + 1) + 1, xtop = std::min(nx - nc, nx - D - 1) - 1;
    const int ixl0 = xbot + T;                        // first column whose T-step cone is clean
    const int ixmax = xtop - T + 1;                   // end of such columns
    const int nbi = std::max(0, (ixmax - ixl0) / V);  // full interior bands
    // The full bands sit where both column-edge regions (left of them and
    // right of them) fit in one band each if they can: a column-edge lane
    // costs about three interior ones, and a third band would be mostly idle.
    const int ixl = nbi > 0 ? std::min(std::max(ixl0, nx + 1 - nbi * V - V), ixmax - nbi * V) : ixl0;
    const int ixr = ixl + nbi * V;                    // end of interior owned columns
    const int ia = ybot + T, ib = ytop - T + 1;       // interior output rows [ia, ib)
    const bool split = d.mode == TOMO_MODE_PHYSICAL && nbi > 0 && ib > ia;
    if (!split) {
        for (int a = 0; a < ny + 1; a += S) add(0, edge, 0, nx + 1, a, std::min(a + S, ny + 1), W);
        return;
    }
    // top and bottom rows: corner bands to the general kernel, clean-column
    // bands (full width, lanes inside [xbot, xtop]) to the row-edge kernel
    auto edge_rows_split = [&](int y0, int y1) {
        // corner and row-edge tiles: short row chunks (their code path is the
        // slowest, and a tile's latency grows with its rows)
        const int sc = std::min(S, ER);
        for (int a = y0; a < y1; a += sc) {
            const int a1 = std::min(a + sc, y1);
            add(0, edge, 0, ixl, a, a1, W);
            add(0, edge, ixr, nx + 1, a, a1, W);
        }
        for (int a = y0; a < y1; a += sc) add(1, rowedge, ixl, ixr, a, std::min(a + sc, y1), Wi);
    };
    edge_rows_split(0, ia);
    // column-edge tiles S rows tall, interior tiles Si (their stages are
    // shorter, so taller interior tiles finish with the column-edge ones)
    for (int a = ia; a < ib; a += S) {
        const int a1 = std::min(a + S, ib);
        add(2, coledge, 0, ixl, a, a1, W);
        add(2, coledge, ixr, nx + 1, a, a1, W);
    }
    for (int a = ia; a < ib; a += Si) add(3, inner, ixl, ixr, a, std::min(a + Si, ib), Wi);
    edge_rows_split(ib, ny + 1);
}

// Tiles of the resident form for one simulation: bands of 64 - 2T owned
// columns, row blocks of at most 8 RW - 2T owned rows (balanced), each with
// its eight neighbours' indices in `out`.
void build_res_tiles(const Slot* sl, int item, int T, int RW, std::vector<ResTile>& out) {
    const DevSim& d = sl->d;
    const int V = 64 - 2 * T, R = kResidentWaves * RW - 2 * T;
    const int nbx = (d.nx + 1 + V - 1) / V, nby = (d.ny + 1 + R - 1) / R;
    const int rb = (d.ny + 1 + nby - 1) / nby;
    const int base = (int)out.size();
    for (int ry = 0; ry < nby; ++ry)
        for (int bx = 0; bx < nbx; ++bx) {
            ResTile t{};
            t.item = item;
            t.x0 = bx * V;
            t.y0 = ry * rb;
            t.y1 = std::min(t.y0 + rb, d.ny + 1);
            int k = 0;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (!dx && !dy) continue;
                    const int X = bx + dx, Y = ry + dy;
                    t.nb[k++] = (X >= 0 && X < nbx && Y >= 0 && Y < nby) ? base + Y * nbx + X : -1;
                }
            out.push_back(t);
        }
}

// Workgroups of a launch are dealt to the eight XCDs round-robin, and a
// kernel's grid is a multiple of 8, so tile p runs on XCD p mod 8.  Order a
// class's tile list (simulation-major, each simulation's tiles row-major) so
// that XCD g walks the g-th contiguous eighth of it: vertically and
// horizontally adjacent tiles, which read each other's halo rows and
// columns, then share one L2.  The last eighth is padded with empty tiles
// (nband 0), which every kernel skips.  The fused launch concatenates the
// classes after ordering each, so every XCD holds an eighth of each class.
void xcd_order(std::vector<TileItem>& v) {
    if (v.size() < 16) return;
    const size_t L = (v.size() + 7) / 8;
    std::vector<TileItem> out(8 * L);
    for (size_t k = 0; k < L; ++k)
        for (size_t g = 0; g < 8; ++g) {
            const size_t src = g * L + k;
            TileItem e = v[src < v.size() ? src : 0];
            if (src >= v.size()) e.nband = 0;
            out[8 * k + g] = e;
        }
    v.swap(out);
}

// Tiles of the tile form: one band of V = 64 - 2T owned columns times up to
// kTileWaveCount * RW - 2T owned rows (rows split evenly), one workgroup each.
void build_wave_tiles(Slot* sl, int item, int T, int RW, std::vector<TileItem>& out) {
    const DevSim& d = sl->d;
    for (int k = 0; k < 4; ++k) sl->cls_cells[k] = sl->cls_bytes[k] = 0;
    const int V = 64 - 2 * T, R = kTileWaveCount * RW - 2 * T;
    const int nbx = (d.nx + 1 + V - 1) / V, nby = (d.ny + 1 + R - 1) / R;
    const int rb = (d.ny + 1 + nby - 1) / nby;
    for (int ry = 0; ry < nby; ++ry)
        for (int bx = 0; bx < nbx; ++bx) {
            TileItem t{};
            t.item = item;
            t.x0 = bx * V;
            t.nband = 1;
            t.y0 = ry * rb;
            t.y1 = std::min(t.y0 + rb, d.ny + 1);
            t.xe = std::min(t.x0 + V, d.nx + 1);
            t.cls = TILE_INNER;
            account_tile(sl, TILE_INNER, t);
            out.push_back(t);
        }
}

// The resident launches' error word: 16-byte aligned past the counters.
int* res_err(const Staging& st) { return st.d_flags + ((st.res_cap + 3) & ~3) + 32; }

// Polls a resident workgroup makes before it gives up on a neighbour (about
// a second of waiting at the kernel's poll interval).
constexpr int kResSpinLimit = 1 << 24;

hipEvent_t pool_event(tomo_ctx* c) {
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    // timing only: no system-scope fence (cache write-back) when the event completes
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return e;
}

// Resolve recorded launch timings (synchronises the streams).
int drain_timings(tomo_ctx* c) {
    if (c->pending.empty()) return TOMO_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 3; ++k) HIPCHK(hipStreamSynchronize(c->side[k]));
    for (auto& p : c->pending) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
        tomo_kernel_stats& k = p.kind < 0 ? c->stats.update : c->stats.kernel[p.kind];
        k.launches++;
        k.ms += ms;
        k.cell_steps += p.cells;
        k.alg_bytes += p.bytes;
    }
    for (auto& p : c->pending)
        if (p.owns) {
            c->ev_pool.push_back(p.a);
            c->ev_pool.push_back(p.b);
        }
    c->pending.clear();
    return TOMO_OK;
}

// Times the one kernel launched between construction and done(): the launch's
// dispatch packet carries the two events, so profiling inserts no marker
// packets between kernels (a marker per launch slowed a 1024^2 evaluation by
// about 12%).
struct LaunchTimer {
    tomo_ctx* c;
    hipEvent_t a = nullptr, b = nullptr;
    explicit LaunchTimer(tomo_ctx* c_) : c(c_) {
        if (c->prof) {
            a = pool_event(c);
            b = pool_event(c);
            if (a && b) arm_launch_events(a, b);
        }
    }
    ~LaunchTimer() { arm_launch_events(nullptr, nullptr); }
    // call only after the launch succeeded: its events are then recorded
    void done(int kind, double cells = 0, double bytes = 0) {
        if (c->prof && a && b) c->pending.push_back({a, b, kind, cells, bytes, true});
    }
};

// Epochs whose launch items one upload carries, and the item count it may hold.
constexpr int kMaxStagedEpochs = 256, kMaxStagedItems = 65536;

// The context's own staging, over its current DevSim table.
Staging& stage(tomo_ctx* c) {
    c->stg.d_sims = c->d_sims;
    return c->stg;
}

// Run epochs until every slot in `sl` reaches its target step, staging the
// launches in `st` (whose d_sims holds the slots' DevSim entries).
int run_epochs(tomo_ctx* c, const std::vector<Slot*>& sl, Staging& st) {
    const int T = c->T, S = c->S, W = c->W, K = c->K;
    if (!tblock_supported(T) || K % T || K > 64 || W < 1 || W > 4 || S < 1)
        return fail(TOMO_E_INVALID, "unsupported schedule");
    for (Slot* s : sl)
        if (s->d.chunk != K) return fail(TOMO_E_STATE, "simulation was created under a different DFT chunk");
    bool res_used = false;  // a resident launch ran: its error word is checked at the end
    for (int mode = 0; mode < 2; ++mode) {
        std::vector<Slot*> ms;
        for (Slot* s : sl)
            if (s->d.mode == mode && s->done < s->target) ms.push_back(s);
        bool fp64_any = false, fp32_any = false;
        for (Slot* s : ms) (s->d.flags & TOMO_FLAG_DFT_FP64) ? fp64_any = true : fp32_any = true;
        while (!ms.empty()) {
            // tile list (rebuilt only when the set of simulations changes)
            std::vector<uint64_t> key;
            for (Slot* s : ms) key.push_back(s->serial);
            std::vector<TileItem> te, tr, tc, tn;
            const bool rebuild = key != st.tile_key || st.key_T != T || st.key_S != S || st.key_W != W ||
                                 st.key_Wi != c->Wi || st.key_ER != c->ER || st.key_Si != c->Si ||
                                 st.res_key_form != c->form;
            // resident form: every tile of the set in one cooperative launch,
            // when they fit on the device at once (else the fused form runs)
            const bool want_res = c->form == TOMO_FORM_RESIDENT && tres_supported(T, S) && !st.fixed;
            // tile form: whole register tiles, one workgroup each, in place of the four classes
            const bool wave = c->form == TOMO_FORM_TILE && mode == TOMO_MODE_PHYSICAL && twave_supported(T, S);
            std::vector<ResTile> trs;
            if (rebuild) {
                for (size_t i = 0; i < ms.size(); ++i) {
                    if (wave) build_wave_tiles(ms[i], (int)i, T, S, tn);
                    else build_tiles(ms[i], (int)i, T, S, c->Si ? c->Si : S, W, c->Wi, c->ER, te, tr, tc, tn);
                }
                if (c->form == TOMO_FORM_STREAMS || c->form == TOMO_FORM_FUSED) {
                    xcd_order(te);
                    xcd_order(tr);
                    xcd_order(tc);
                    xcd_order(tn);
                }
                if (want_res) {
                    for (size_t i = 0; i < ms.size(); ++i) build_res_tiles(ms[i], (int)i, T, S, trs);
                    if ((int)trs.size() > tres_max_tiles(T, S, mode, c->cus)) trs.clear();
                }
            }
            const int ntiles = rebuild ? (int)(te.size() + tr.size() + tc.size() + tn.size()) : st.ntiles;
            // Items of E consecutive epochs in one upload: every simulation of
            // the set is active in each of them (only the last may be partial),
            // so the set, and with it the tile list, stays fixed.  A captured
            // graph reads one epoch's items at a fixed address, so graphs stage one.
            const int nsim = (int)ms.size();
            int E = (c->use_graph && !st.fixed) ? 1 : std::max(1, std::min(kMaxStagedEpochs, kMaxStagedItems / nsim));
            for (Slot* s : ms) E = std::min(E, (s->target - s->done + K - 1) / K);
            int rc = ensure_staging(c, st, E * nsim, ntiles);
            if (rc) return rc;
            // the host staging is rewritten below: wait for its last upload
            if (st.staged) HIPCHK(hipEventSynchronize(st.staged));
            else HIPCHK(hipStreamSynchronize(c->stream));
            int maxP = 0, maxF = 0;
            for (int i = 0; i < nsim; ++i) {
                Slot* s = ms[i];
                for (int e = 0; e < E; ++e) {
                    LaunchItem& it = st.h_items[e * nsim + i];
                    it.sim = s->idx;
                    it.step0 = s->done + e * K;
                    it.parity0 = s->parity ^ ((e * (K / T)) & 1);  // epochs before the last are full
                    it.nsteps = std::min(K, s->target - s->done - e * K);
                }
                maxP = std::max(maxP, s->d.P);
                maxF = std::max(maxF, s->d.nfreq);
            }
            HIPCHK(hipMemcpyAsync(st.d_items, st.h_items, sizeof(LaunchItem) * E * nsim, hipMemcpyHostToDevice,
                                  c->stream));
            if (rebuild) {
                TileItem* h = st.h_tiles;
                h = std::copy(te.begin(), te.end(), h);
                h = std::copy(tr.begin(), tr.end(), h);
                h = std::copy(tc.begin(), tc.end(), h);
                std::copy(tn.begin(), tn.end(), h);
                HIPCHK(hipMemcpyAsync(st.d_tiles, st.h_tiles, sizeof(TileItem) * ntiles, hipMemcpyHostToDevice,
                                      c->stream));
                st.tile_key = key;
                st.key_T = T; st.key_S = S; st.key_W = W; st.key_Wi = c->Wi;
                st.key_ER = c->ER;
                st.key_Si = c->Si;
                st.ntiles = ntiles;
                st.n_edge = (int)(te.size() + tr.size() + tc.size());
                st.n_corner = (int)te.size();
                st.n_rowedge = (int)tr.size();
                st.res_key_form = c->form;
                st.nres = 0;
                if (!trs.empty()) {
                    rc = ensure_res_staging(c, st, (int)trs.size());
                    if (rc) return rc;
                    std::copy(trs.begin(), trs.end(), st.h_res);
                    HIPCHK(hipMemcpyAsync(st.d_res, st.h_res, sizeof(ResTile) * trs.size(), hipMemcpyHostToDevice,
                                          c->stream));
                    st.nres = (int)trs.size();
                    st.tile_key = key;  // ensure_res_staging may have cleared it
                }
            }
            bool resident = want_res && st.nres > 0 && !c->res_refused;
            if (st.staged) HIPCHK(hipEventRecord(st.staged, c->stream));
            const int n_edge = st.n_edge, n_inner = st.ntiles - st.n_edge;
            const int cap = c->cus * c->grid_mult;
            // the staged epoch being launched: its items on the host and the device
            const LaunchItem* h_it = st.h_items;
            LaunchItem* d_it = st.d_items;
            int nl = 0;  // time blocks of that epoch
            // tile classes: corners (general kernel), row-edge, column-edge, interior
            const int n0 = st.n_corner, nr = st.n_rowedge, ncl = n_edge - n0 - nr;
            const int cnt[3] = {n0, nr, ncl};
            int nside = 0;
            for (int k = 0; k < 3; ++k) nside += cnt[k] > 0;
            bool fused = mode == TOMO_MODE_PHYSICAL && !resident && !wave &&
                         (c->form == TOMO_FORM_FUSED || c->form == TOMO_FORM_RESIDENT || c->form == TOMO_FORM_TILE);
            bool multi = mode == TOMO_MODE_PHYSICAL && !fused && !resident && !wave && n_inner > 0 && nside > 0;
            // per-class cell-steps and bytes of block o (kernel timing only)
            auto block_work = [&](int o, int k, double& cells, double& bytes) {
                cells = bytes = 0;
                if (!c->prof) return;
                for (size_t i = 0; i < ms.size(); ++i) {
                    const int nst = std::min(T, h_it[i].nsteps - o * T);
                    if (nst <= 0) continue;
                    cells += nst * ms[i]->cls_cells[k];
                    bytes += nst * ms[i]->cls_bytes[k];
                }
            };
            // the launches of one epoch: ceil(nsteps / T) time blocks, then the fold
            auto launch_epoch = [&]() -> int {
                if (multi) HIPCHK(hipEventRecord(c->ev_bar, c->stream));
                double ucells = 0, ubytes = 0;
                hipEvent_t first = nullptr;  // start of the update phase: its first launch
                if (resident) {
                    // every time block of the epoch in one launch; the counters start at zero
                    for (int o = 0; o < nl; ++o)
                        for (int k = 0; k < 4; ++k) {
                            double cl = 0, by = 0;
                            block_work(o, k, cl, by);
                            ucells += cl;
                            ubytes += by;
                        }
                    if (!res_used) HIPCHK(launch_zero(res_err(st), 16, c->stream));
                    res_used = true;
                    HIPCHK(launch_zero(st.d_flags, (sizeof(int) * st.nres + 15) & ~(size_t)15, c->stream));
                    hipEvent_t a = c->prof ? pool_event(c) : nullptr, b = c->prof ? pool_event(c) : nullptr;
                    if (a && b) HIPCHK(hipEventRecord(a, c->stream));
                    const hipError_t le = launch_tres(T, S, mode, st.d_res, st.nres, d_it, st.d_sims, st.d_flags,
                                                      res_err(st), kResSpinLimit, c->stream);
                    if (le == hipSuccess) {
                        if (a && b) {
                            HIPCHK(hipEventRecord(b, c->stream));
                            c->pending.push_back({a, b, TOMO_KERNEL_RESIDENT, ucells, ubytes, true});
                            first = a;
                        }
                    } else {
                        // The device refused the cooperative launch (one process
                        // at a time may hold a device's cooperative queue), so
                        // nothing was queued: this epoch and the context's later
                        // ones run the fused form, whose results are the same.
                        (void)hipGetLastError();
                        for (hipEvent_t e : {a, b})
                            if (e) c->ev_pool.push_back(e);
                        c->res_refused = true;
                        resident = false;
                        fused = mode == TOMO_MODE_PHYSICAL;
                        multi = false;
                        ucells = ubytes = 0;
                    }
                }
                for (int o = 0; !resident && o < nl; ++o) {
                    double cells[4], bytes[4];
                    for (int k = 0; k < 4; ++k) {
                        block_work(o, k, cells[k], bytes[k]);
                        ucells += cells[k];
                        ubytes += bytes[k];
                    }
                    if (mode != TOMO_MODE_PHYSICAL) {  // (the resident form skipped this loop)
                        LaunchTimer lt(c);
                        HIPCHK(launch_tblock(T, mode, st.d_tiles, n_edge, d_it, st.d_sims, o, W,
                                             std::min(n_edge, cap), c->stream));
                        lt.done(TOMO_KERNEL_GENERAL, cells[0], bytes[0]);
                        if (!first) first = lt.a;
                        continue;
                    }
                    if (wave) {
                        const int nt = st.ntiles;
                        LaunchTimer lt(c);
                        HIPCHK(launch_twave(T, S, st.d_tiles, nt, d_it, st.d_sims, o, std::min(nt, 2 * cap),
                                            c->stream));
                        lt.done(TOMO_KERNEL_TILE, cells[0] + cells[1] + cells[2] + cells[3],
                                bytes[0] + bytes[1] + bytes[2] + bytes[3]);
                        if (!first) first = lt.a;
                        continue;
                    }
                    if (fused) {
                        const int nt = st.ntiles;
                        LaunchTimer lt(c);
                        HIPCHK(launch_tfused(T, st.d_tiles, nt, d_it, st.d_sims, o, std::max(W, c->Wi),
                                             std::min(nt, 2 * cap), c->stream));
                        lt.done(TOMO_KERNEL_FUSED, cells[0] + cells[1] + cells[2] + cells[3],
                                bytes[0] + bytes[1] + bytes[2] + bytes[3]);
                        if (!first) first = lt.a;
                        continue;
                    }
                    // every class of block o starts after all of block o-1 (ev_bar)
                    int off = 0;
                    for (int k = 0; k < 3; ++k) {
                        const int n = cnt[k];
                        const TileItem* tl = st.d_tiles + off;
                        off += n;
                        if (n == 0) continue;
                        hipStream_t es = multi ? c->side[k] : c->stream;
                        if (multi) HIPCHK(hipStreamWaitEvent(es, c->ev_bar, 0));
                        LaunchTimer lt(c);
                        if (k == 0) HIPCHK(launch_tedge(T, tl, n, d_it, st.d_sims, o, W, std::min(n, 2 * cap), es));
                        if (k == 1) HIPCHK(launch_trow(T, tl, n, d_it, st.d_sims, o, c->Wi, std::min(n, 2 * cap), es));
                        if (k == 2) HIPCHK(launch_tcol(T, tl, n, d_it, st.d_sims, o, W, std::min(n, 2 * cap), es));
                        lt.done(TOMO_KERNEL_CORNER + k, cells[k], bytes[k]);
                        if (!first) first = lt.a;
                        if (multi) HIPCHK(hipEventRecord(c->ev_side[k], es));
                    }
                    if (n_inner > 0) {
                        LaunchTimer lt(c);
                        HIPCHK(launch_tinner(T, st.d_tiles + n_edge, n_inner, d_it, st.d_sims, o, c->Wi,
                                             std::min(n_inner, 2 * cap), c->stream));
                        lt.done(TOMO_KERNEL_INNER, cells[3], bytes[3]);
                        if (!first) first = lt.a;
                    }
                    if (multi) {
                        for (int k = 0; k < 3; ++k)
                            if (cnt[k] > 0) HIPCHK(hipStreamWaitEvent(c->stream, c->ev_side[k], 0));
                        HIPCHK(hipEventRecord(c->ev_bar, c->stream));
                    }
                }
                // the fold runs after every update launch of the epoch: its start ends the update phase
                hipEvent_t fold_start = nullptr;
                for (int prec = 0; prec < 2; ++prec) {
                    if (!(prec ? fp64_any : fp32_any)) continue;
                    LaunchTimer lt(c);
                    HIPCHK(launch_fold(prec == 1, d_it, (int)ms.size(), st.d_sims, maxP, maxF, c->stream));
                    lt.done(TOMO_KERNEL_FOLD);
                    if (!fold_start) fold_start = lt.a;
                }
                if (c->prof && first && fold_start) c->pending.push_back({first, fold_start, -1, ucells, ubytes, false});
                return TOMO_OK;
            };
            // A full epoch of an unchanged set of simulations is captured once into a
            // hipGraph and replayed (timed runs launch directly: their per-launch
            // events cannot live inside a graph).
            for (int e = 0; e < E; ++e) {
                h_it = st.h_items + e * nsim;
                d_it = st.d_items + e * nsim;
                int maxn = 0, minn = K;
                for (int i = 0; i < nsim; ++i) {
                    maxn = std::max(maxn, h_it[i].nsteps);
                    minn = std::min(minn, h_it[i].nsteps);
                }
                nl = (maxn + T - 1) / T;
                if (c->use_graph && !st.fixed && !resident && !c->prof && minn == K) {
                    const std::vector<int64_t> gk = {mode, T, S, c->Si, W, c->Wi, c->ER, c->form, K, c->grid_mult, ntiles, maxP,
                                                     fp32_any, fp64_any, (int64_t)ms.size(), (int64_t)(intptr_t)d_it,
                                                     (int64_t)(intptr_t)st.d_tiles, (int64_t)(intptr_t)st.d_sims};
                    if (!c->gexec || gk != c->gkey || key != c->gkey_serials) {
                        if (c->gexec) {
                            HIPCHK(hipGraphExecDestroy(c->gexec));
                            c->gexec = nullptr;
                        }
                        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
                        const int rc2 = launch_epoch();
                        hipGraph_t g = nullptr;
                        const hipError_t ec = hipStreamEndCapture(c->stream, &g);
                        if (rc2) {
                            if (g) (void)hipGraphDestroy(g);
                            return rc2;
                        }
                        HIPCHK(ec);
                        const hipError_t ei = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
                        (void)hipGraphDestroy(g);
                        HIPCHK(ei);
                        c->gkey = gk;
                        c->gkey_serials = key;
                    }
                    HIPCHK(hipGraphLaunch(c->gexec, c->stream));
                } else {
                    const int rc2 = launch_epoch();
                    if (rc2) return rc2;
                }
                if (c->pending.size() > 4096) {
                    int rc2 = drain_timings(c);
                    if (rc2) return rc2;
                }
                for (int i = 0; i < nsim; ++i) {
                    Slot* s = ms[i];
                    const int ns = h_it[i].nsteps;
                    s->done += ns;
                    s->parity ^= ((ns + T - 1) / T) & 1;
                }
            }  // staged epochs
            std::vector<Slot*> next;
            for (Slot* s : ms)
                if (s->done < s->target) next.push_back(s);
            ms.swap(next);
        }
    }
    if (res_used) {
        int err = 0;
        HIPCHK(hipMemcpyAsync(&err, res_err(st), sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (err) return fail(TOMO_E_HIP, "resident update: a workgroup gave up waiting for its neighbours");
    }
    return TOMO_OK;
}

// fold kernels of both precisions skip the items of the other precision
}  // namespace

// The fold kernel template is chosen per slot precision; items of the other
// precision are harmless because both layouts are 16 bytes per entry, but the
// arithmetic would be wrong, so keep one precision per run.
static int check_uniform_precision(const std::vector<Slot*>& sl) {
    int f64 = -1;
    for (Slot* s : sl) {
        int v = (s->d.flags & TOMO_FLAG_DFT_FP64) ? 1 : 0;
        if (f64 >= 0 && v != f64) return fail(TOMO_E_INVALID, "mixed DFT precision in one run");
        f64 = v;
    }
    return TOMO_OK;
}

// Project the far field of `sl` into D/P (device, [slot][freq][angle] at
// offset out_off[i]).  Consecutive simulations that share grid, contour,
// frequencies and angles and whose records are contiguous form one group: one
// matrix-core launch per group.
// The kernels index `d_sims` (the context's table by default); `staging`, when
// given, holds the launch's small tables instead of a device allocation.
static int project(tomo_ctx* c, const std::vector<Slot*>& sl, const std::vector<int>& out_off, float* dD,
                   float* dP, const DevSim* d_sims = nullptr, char* staging = nullptr, size_t staging_cap = 0) {
    if (sl.empty()) return TOMO_OK;
    if (!d_sims) d_sims = c->d_sims;
    auto same = [](const tomo_sim_config& x, const tomo_sim_config& y) {
        return x.mode == y.mode && x.nx == y.nx && x.ny == y.ny && x.ntff_offset == y.ntff_offset &&
               x.dx == y.dx && x.dy == y.dy && x.n_freq == y.n_freq && x.n_angles == y.n_angles &&
               x.f_start == y.f_start && x.f_stop == y.f_stop && x.pulse_tau == y.pulse_tau &&
               x.pulse_t0 == y.pulse_t0 && x.ff_norm == y.ff_norm &&
               ((x.flags ^ y.flags) & TOMO_FLAG_DFT_FP64) == 0;
    };
    struct Group { size_t i0, n; };
    std::vector<Group> groups;
    for (size_t i = 0; i < sl.size(); ++i) {
        const tomo_sim_config& ci = sl[i]->su.cfg;
        const int rec = ci.n_freq * ci.n_angles;
        if (!groups.empty()) {
            Group& g = groups.back();
            const Slot* f = sl[g.i0];
            if (same(f->su.cfg, ci) && out_off[i] == out_off[g.i0] + (int)(i - g.i0) * rec) {
                ++g.n;
                continue;
            }
        }
        groups.push_back({i, 1});
    }
    // staging: per group the simulation indices and per frequency {k, normaliser}
    std::vector<int> idx(sl.size());
    std::vector<float2> kinv;
    std::vector<size_t> kofs;
    for (size_t i = 0; i < sl.size(); ++i) idx[i] = sl[i]->idx;
    for (const Group& g : groups) {
        const SimSetup& su = sl[g.i0]->su;
        kofs.push_back(kinv.size());
        for (int f = 0; f < su.cfg.n_freq; ++f)
            kinv.push_back(make_float2((float)(2 * kPI * su.freq[f] / c0()), (float)su.inv_norm[f]));
    }
    const size_t bytes = align256(idx.size() * sizeof(int)) + kinv.size() * sizeof(float2);
    char* d = staging;
    if (d && bytes > staging_cap) return fail(TOMO_E_NOMEM, "projection tables exceed the fixed staging");
    if (!d) HIPCHK(hipMalloc(&d, bytes));
    int* d_idx = (int*)d;
    float2* d_k = (float2*)(d + align256(idx.size() * sizeof(int)));
    hipError_t e = hipMemcpyAsync(d_idx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_k, kinv.data(), kinv.size() * sizeof(float2), hipMemcpyHostToDevice, c->stream);
    for (size_t gi = 0; gi < groups.size() && e == hipSuccess; ++gi) {
        const Group& g = groups[gi];
        const tomo_sim_config& cf = sl[g.i0]->su.cfg;
        LaunchTimer lt(c);
        e = launch_project_mfma(d_idx + g.i0, (int)g.n, cf.n_freq, d_sims, d_k + kofs[gi], cf.n_angles,
                                (float)eta0(), dD + out_off[g.i0], dP ? dP + out_off[g.i0] : nullptr, c->stream);
        if (e == hipSuccess) lt.done(TOMO_KERNEL_PROJECT);
    }
    // the uploads read host vectors of this frame
    const hipError_t es = hipStreamSynchronize(c->stream);
    if (!staging) (void)hipFree(d);
    HIPCHK(e);
    HIPCHK(es);
    return TOMO_OK;
}

// Run the given simulations (all of one configuration) to completion and
// project them; D_dev receives [slot][freq][angle].
static int run_and_project(tomo_ctx* c, const std::vector<Slot*>& sl, float* dD, float* dP) {
    int rc = check_uniform_precision(sl);
    if (rc) return rc;
    rc = run_epochs(c, sl, stage(c));
    if (rc) return rc;
    std::vector<int> off;
    for (size_t i = 0; i < sl.size(); ++i) off.push_back((int)(i * sl[i]->su.cfg.n_freq * sl[i]->su.cfg.n_angles));
    return project(c, sl, off, dD, dP);
}

static int get_target(tomo_ctx* c, const tomo_sim_config& cfg, std::vector<float>& out);

// Evaluate n_vec vectors; `plain` runs one simulation at a time.
static int eval_many(tomo_ctx* c, const tomo_sim_config& cfg, const double* params, size_t n_vec, size_t dim,
                     const float* target, size_t n_target, double* fit_out, float* D_out, float* P_out, bool plain) {
    int rc = validate(cfg);
    if (rc) return rc;
    if (!params && n_vec) return fail(TOMO_E_INVALID, "null parameters");
    const int F = cfg.n_freq, A = cfg.n_angles, E = cfg.n_exc;
    const size_t L = (size_t)E * F * A;
    std::vector<float> tgt;
    const float* tph = nullptr;
    if (target) {
        if (n_target < L) return fail(TOMO_E_INVALID, "target shorter than the far-field record");
        tgt.assign(target, target + L);
        if ((cfg.flags & TOMO_FLAG_PHASE_FITNESS) && n_target >= 2 * L) tph = target + L;
    } else {
        if (cfg.mode != TOMO_MODE_PHYSICAL) return fail(TOMO_E_INVALID, "ref-compat mode needs a target (measurement)");
        rc = get_target(c, cfg, tgt);
        if (rc) return rc;
    }
    const size_t nrec = n_vec * L;
    DevMem mD, mP, mT, mF;
    HIPCHK(mD.alloc(nrec * 4 + 256));
    HIPCHK(mP.alloc(nrec * 4 + 256));
    HIPCHK(mT.alloc(2 * L * 4 + 256));
    HIPCHK(mF.alloc(n_vec * 8 + 256));
    float *dD = mD.as<float>(), *dP = mP.as<float>(), *dT = mT.as<float>();
    double* dF = mF.as<double>();
    HIPCHK(hipMemcpyAsync(dT, tgt.data(), L * 4, hipMemcpyHostToDevice, c->stream));
    if (tph) HIPCHK(hipMemcpyAsync(dT + L, tph, L * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    rc = TOMO_OK;
    if (plain) {
        for (size_t v = 0; v < n_vec && !rc; ++v) {
            for (int e = 0; e < E && !rc; ++e) {
                Slot* s = nullptr;
                rc = slot_create(c, cfg, params + v * dim, dim, e, &s);
                if (rc) break;
                std::vector<Slot*> one{s};
                rc = run_and_project(c, one, dD + v * L + (size_t)e * F * A, dP + v * L + (size_t)e * F * A);
                slot_destroy(c, s);
            }
        }
    } else {
        // as many whole vectors per pass as the HBM budget admits
        size_t v0 = 0;
        while (v0 < n_vec && !rc) {
            std::vector<Slot*> sl;
            size_t v = v0;
            for (; v < n_vec && !rc; ++v) {
                std::vector<Slot*> mine;
                for (int e = 0; e < E; ++e) {
                    Slot* s = nullptr;
                    int r2 = slot_create(c, cfg, params + v * dim, dim, e, &s);
                    if (r2) {
                        for (Slot* m : mine) slot_destroy(c, m);
                        if (r2 == TOMO_E_NOMEM && !sl.empty()) goto run_pass;
                        rc = r2;
                        break;
                    }
                    mine.push_back(s);
                }
                if (rc) break;
                sl.insert(sl.end(), mine.begin(), mine.end());
            }
        run_pass:
            if (!rc) rc = run_and_project(c, sl, dD + v0 * L, dP + v0 * L);
            for (Slot* s : sl) slot_destroy(c, s);
            v0 = v;
        }
    }
    if (!rc) {
        const int terms = (tph ? 1 : 0) | ((cfg.flags & TOMO_FLAG_REL_FITNESS) ? 2 : 0);
        const double norm = (cfg.flags & TOMO_FLAG_FITNESS_PER_EXC) ? (double)E : (double)E * F;
        LaunchTimer lt(c);
        hipError_t e = launch_fitness(dD, dP, dT, tph ? dT + L : nullptr, (int)L, (int)n_vec, norm, terms, dF,
                                      c->stream);
        if (e == hipSuccess) lt.done(TOMO_KERNEL_FITNESS);
        if (e != hipSuccess) rc = fail(TOMO_E_HIP, hipGetErrorString(e));
    }
    if (!rc) {
        std::vector<double> hf(n_vec);
        hipError_t e = hipMemcpyAsync(hf.data(), dF, n_vec * 8, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && D_out) e = hipMemcpyAsync(D_out, dD, nrec * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && P_out) e = hipMemcpyAsync(P_out, dP, nrec * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(TOMO_E_HIP, hipGetErrorString(e));
        for (size_t v = 0; v < n_vec && !rc; ++v) {
            if (cfg.mode == TOMO_MODE_PHYSICAL && !std::isfinite(hf[v])) rc = fail(TOMO_E_NONFINITE, "non-finite fitness");
            if (fit_out) fit_out[v] = hf[v];
        }
    }
    // the copies above completed (or failed) before the buffers are released
    (void)hipStreamSynchronize(c->stream);
    return rc;
}

static int get_target(tomo_ctx* c, const tomo_sim_config& cfg, std::vector<float>& out) {
    const std::string key = cfg_key(cfg);
    auto it = c->targets.find(key);
    if (it != c->targets.end()) {
        out = it->second;
        return TOMO_OK;
    }
    const std::vector<double> tv = target_vector(cfg);
    const size_t L = (size_t)cfg.n_exc * cfg.n_freq * cfg.n_angles;
    std::vector<Slot*> sl;
    int rc = TOMO_OK;
    for (int e = 0; e < cfg.n_exc && !rc; ++e) {
        Slot* s = nullptr;
        rc = slot_create(c, cfg, tv.data(), tv.size(), e, &s);
        if (!rc) sl.push_back(s);
    }
    float *dD = nullptr, *dP = nullptr;
    if (!rc && hipMalloc(&dD, L * 4) != hipSuccess) rc = fail(TOMO_E_HIP, "hipMalloc");
    if (!rc && hipMalloc(&dP, L * 4) != hipSuccess) rc = fail(TOMO_E_HIP, "hipMalloc");
    if (!rc) rc = run_and_project(c, sl, dD, dP);
    if (!rc) {
        out.resize(L);
        if (hipMemcpy(out.data(), dD, L * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(TOMO_E_HIP, "copy");
    }
    for (Slot* s : sl) slot_destroy(c, s);
    if (dD) (void)hipFree(dD);
    if (dP) (void)hipFree(dP);
    if (!rc) c->targets[key] = out;
    return rc;
}

// ---------------------------------------------------------------- serving
static int64_t request_bytes(const tomo_sim_config& cfg, int chunk) {
    if (validate(cfg)) return -1;
    SimSetup su;
    su.cfg = cfg;
    su.pitch = ((cfg.nx + 1) + 31) & ~31;
    su.P = contour_len(cfg);
    // upper bound of the compact Jp storage: the dispersive disc's bounding box
    const double R = cfg.breast_radius;
    const int rc = std::min(cfg.nx + 1, (int)std::ceil(2 * R / cfg.dx) + 3);
    su.jp_len = rc * std::min(cfg.ny + 1, (int)std::ceil(2 * R / cfg.dy) + 3);
    return (int64_t)(slot_bytes(su, chunk) * (size_t)cfg.n_exc);
}

static int finish_request(tomo_ctx* c, tomo_request* r) {
    const tomo_sim_config& cfg = r->cfg;
    const size_t L = (size_t)cfg.n_exc * cfg.n_freq * cfg.n_angles;
    std::vector<float> tgt;
    int rc = TOMO_OK;
    if (cfg.mode == TOMO_MODE_PHYSICAL) rc = get_target(c, cfg, tgt);
    else rc = fail(TOMO_E_INVALID, "serving needs physical mode");
    float *dD = nullptr, *dP = nullptr, *dT = nullptr;
    double* dF = nullptr;
    if (!rc && (hipMalloc(&dD, L * 4) || hipMalloc(&dP, L * 4) || hipMalloc(&dT, L * 4) || hipMalloc(&dF, 256)))
        rc = fail(TOMO_E_HIP, "hipMalloc");
    if (!rc) {
        std::vector<int> off;
        for (size_t i = 0; i < r->slots.size(); ++i) off.push_back((int)(i * cfg.n_freq * cfg.n_angles));
        rc = project(c, r->slots, off, dD, dP);
    }
    double fit = 0;
    if (!rc) {
        if (hipMemcpyAsync(dT, tgt.data(), L * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            launch_fitness(dD, dP, dT, nullptr, (int)L, 1, (double)cfg.n_exc * cfg.n_freq, 0, dF, c->stream) !=
                hipSuccess ||
            hipMemcpyAsync(&fit, dF, 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = fail(TOMO_E_HIP, "fitness");
    }
    if (dD) (void)hipFree(dD);
    if (dP) (void)hipFree(dP);
    if (dT) (void)hipFree(dT);
    if (dF) (void)hipFree(dF);
    for (Slot* s : r->slots) slot_destroy(c, s);
    r->slots.clear();
    r->fitness = fit;
    r->status = rc ? rc : (std::isfinite(fit) ? TOMO_OK : TOMO_E_NONFINITE);
    r->state = r->status == TOMO_OK ? 2 : 3;
    r->t_done = now_s();
    return rc;
}

// One epoch of the request loop: admit, advance every resident simulation by
// one chunk, retire finished requests.
static int progress_once(tomo_ctx* c, int* n_active) {
    // admission by HBM footprint, FIFO
    while (!c->queue.empty()) {
        auto qi = c->reqs.find(c->queue.front());
        tomo_request* r = qi == c->reqs.end() ? nullptr : qi->second.get();
        if (!r || r->state == 4) { c->queue.pop_front(); continue; }
        if (c->used + (uint64_t)r->bytes > c->budget) {
            if (c->active.empty()) {  // cannot ever fit
                r->state = 3;
                r->status = fail(TOMO_E_NOMEM, "request larger than the HBM budget");
                r->t_done = now_s();
                c->queue.pop_front();
                continue;
            }
            break;
        }
        c->queue.pop_front();
        int rc = TOMO_OK;
        for (int e = 0; e < r->cfg.n_exc && !rc; ++e) {
            Slot* s = nullptr;
            rc = slot_create(c, r->cfg, r->params.data(), r->params.size(), e, &s);
            if (!rc) { s->req = (long)r->ticket; r->slots.push_back(s); }
        }
        if (rc) {
            for (Slot* s : r->slots) slot_destroy(c, s);
            r->slots.clear();
            r->state = 3;
            r->status = rc;
            r->t_done = now_s();
            continue;
        }
        r->state = 1;
        for (Slot* s : r->slots) c->active.push_back(s);
    }
    if (!c->active.empty()) {
        // one epoch: cap every target at done + chunk
        std::vector<Slot*> run;
        for (Slot* s : c->active) {
            s->target = std::min(s->su.cfg.steps, s->done + c->K);
            run.push_back(s);
        }
        int rc = run_epochs(c, run, stage(c));
        if (rc) return rc;
        // retire
        std::vector<Slot*> still;
        std::vector<tomo_request*> fin;
        // every active slot belongs to a running request, whose record tomo_poll
        // keeps (it reports running requests as pending)
        auto owner = [&](const Slot* s) -> tomo_request* {
            auto it = c->reqs.find((uint64_t)s->req);
            return it == c->reqs.end() ? nullptr : it->second.get();
        };
        for (Slot* s : c->active) {
            if (s->done >= s->su.cfg.steps) {
                tomo_request* r = owner(s);
                if (!r) continue;
                bool all = true;
                for (Slot* q : r->slots) all = all && q->done >= q->su.cfg.steps;
                if (all && std::find(fin.begin(), fin.end(), r) == fin.end()) fin.push_back(r);
            }
        }
        for (Slot* s : c->active) {
            tomo_request* r = owner(s);
            if (r && std::find(fin.begin(), fin.end(), r) == fin.end()) still.push_back(s);
        }
        c->active.swap(still);
        for (tomo_request* r : fin) finish_request(c, r);
    }
    if (n_active) *n_active = (int)c->active.size();
    return TOMO_OK;
}

// ---------------------------------------------------------------- allocation-free evaluation
// Device memory of tomo_eval_noalloc: every buffer of its evaluation lives
// here, so that the entry allocates nothing.  Each device the library runs on
// holds its own copy of the array until the process ends.  Host code never
// uses the array's name as an address: the device address comes from
// hipGetSymbolAddress, once per device, with that device current.
static __device__ __align__(256) char scratch[TOMO_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;                // guards g_scratch_ptr
char* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` per device
std::mutex g_noalloc_mu[kMaxDevices];   // one allocation-free evaluation at a time per device

// Device address of `scratch` on `device`, which must be current.
int scratch_address(int device, char** out) {
    if (device < 0 || device >= kMaxDevices) return fail(TOMO_E_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(TOMO_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = (char*)ptr;
    }
    *out = g_scratch_ptr[device];
    return TOMO_OK;
}

// Small tables of the projection launch: simulation indices and per
// frequency {k, normaliser} (one simulation at a time here).
constexpr size_t kNoallocProjBytes = 256 + TOMO_MAX_FREQ * sizeof(float2);

// tomo_eval's plain path with every device buffer carved from `base`
// (TOMO_NOALLOC_SCRATCH_BYTES): fixed tables first, the simulation slot at the
// end, the tile list in between.  The caller holds the context's and the
// device's locks.
int eval_noalloc(tomo_ctx* c, char* base, const tomo_sim_config& cfg, const double* params, size_t n_params,
                 const float* target, size_t n_target, double* fit_out, float* D_out, float* P_out) {
    int rc = validate(cfg);
    if (rc) return rc;
    if (!params) return fail(TOMO_E_INVALID, "null parameters");
    const int F = cfg.n_freq, A = cfg.n_angles, E = cfg.n_exc;
    const size_t L = (size_t)E * F * A;
    if (!target || n_target < L)
        return fail(TOMO_E_INVALID, "tomo_eval_noalloc needs the target far field (n_exc*n_freq*n_angles values)");
    const float* tph = ((cfg.flags & TOMO_FLAG_PHASE_FITNESS) && n_target >= 2 * L) ? target + L : nullptr;
    if (c->K % c->T) return fail(TOMO_E_INVALID, "chunk must be a multiple of the fused step count");
    // host setup of every excitation first: it sizes the slot
    std::vector<std::unique_ptr<Slot>> slots;
    size_t sbytes = 0;
    for (int e = 0; e < E; ++e) {
        std::unique_ptr<Slot> s(new Slot());
        rc = build_setup(cfg, params, n_params, e, s->su);
        if (rc) return rc;
        sbytes = std::max(sbytes, slot_bytes(s->su, c->K));
        slots.push_back(std::move(s));
    }
    size_t off = 0;
    auto carve = [&](size_t b) { char* r = base + off; off += align256(b); return r; };
    DevSim* tbl = (DevSim*)carve(sizeof(DevSim));
    float* dD = (float*)carve(L * 4);
    float* dP = (float*)carve(L * 4);
    float* dT = (float*)carve(2 * L * 4);
    double* dF = (double*)carve(sizeof(double));
    char* pstage = carve(kNoallocProjBytes);
    LaunchItem* d_items = (LaunchItem*)carve(sizeof(LaunchItem) * kMaxStagedEpochs);
    const size_t total = TOMO_NOALLOC_SCRATCH_BYTES;
    const size_t slot_off = (total - sbytes) & ~(size_t)255;
    if (sbytes > total || slot_off < off + align256(64 * sizeof(TileItem)))
        return fail(TOMO_E_NOMEM, "configuration needs " + std::to_string(off + sbytes) +
                                      " bytes of scratch; tomo_eval_noalloc holds " + std::to_string(total));
    TileItem* d_tiles = (TileItem*)(base + off);
    const int tiles_cap = (int)std::min<size_t>((slot_off - off) / sizeof(TileItem), 1u << 24);
    char* slot_mem = base + slot_off;
    // pageable host staging: a fixed staging waits for the stream before rewriting it
    std::vector<LaunchItem> h_items(kMaxStagedEpochs);
    std::vector<TileItem> h_tiles((size_t)tiles_cap);
    Staging st;
    st.fixed = true;
    st.h_items = h_items.data();
    st.d_items = d_items;
    st.items_cap = kMaxStagedEpochs;
    st.h_tiles = h_tiles.data();
    st.d_tiles = d_tiles;
    st.tiles_cap = tiles_cap;
    st.d_sims = tbl;
    hipStream_t sm = c->stream;
    HIPCHK(hipMemcpyAsync(dT, target, L * 4, hipMemcpyHostToDevice, sm));
    if (tph) HIPCHK(hipMemcpyAsync(dT + L, tph, L * 4, hipMemcpyHostToDevice, sm));
    for (int e = 0; e < E; ++e) {
        Slot* s = slots[e].get();
        s->mem = slot_mem;  // each excitation reuses the slot memory after the previous one's work (stream order)
        s->bytes = sbytes;
        s->idx = 0;
        s->serial = c->next_serial++;
        rc = slot_fill(c, s, tbl);
        if (rc) return rc;
        s->exc = e;
        s->target = cfg.steps;
        const std::vector<Slot*> one{s};
        rc = run_epochs(c, one, st);
        if (rc) return rc;
        rc = project(c, one, {0}, dD + (size_t)e * F * A, dP + (size_t)e * F * A, tbl, pstage, kNoallocProjBytes);
        if (rc) return rc;
    }
    const int terms = (tph ? 1 : 0) | ((cfg.flags & TOMO_FLAG_REL_FITNESS) ? 2 : 0);
    const double norm = (cfg.flags & TOMO_FLAG_FITNESS_PER_EXC) ? (double)E : (double)E * F;
    HIPCHK(launch_fitness(dD, dP, dT, tph ? dT + L : nullptr, (int)L, 1, norm, terms, dF, sm));
    double fit = 0;
    HIPCHK(hipMemcpyAsync(&fit, dF, sizeof(double), hipMemcpyDeviceToHost, sm));
    if (D_out) HIPCHK(hipMemcpyAsync(D_out, dD, L * 4, hipMemcpyDeviceToHost, sm));
    if (P_out) HIPCHK(hipMemcpyAsync(P_out, dP, L * 4, hipMemcpyDeviceToHost, sm));
    HIPCHK(hipStreamSynchronize(sm));
    if (cfg.mode == TOMO_MODE_PHYSICAL && !std::isfinite(fit)) return fail(TOMO_E_NONFINITE, "non-finite fitness");
    if (fit_out) *fit_out = fit;
    return TOMO_OK;
}
}  // namespace

// ---------------------------------------------------------------- single simulation access
struct tomo_sim {
    std::shared_ptr<SimLink> link;
};

// Context and slot of a live handle, or an error when its context was destroyed.
static int sim_live(tomo_sim* sim, tomo_ctx** c, Slot** s) {
    if (!sim || !sim->link) return fail(TOMO_E_INVALID, "null sim");
    if (!sim->link->ctx) return fail(TOMO_E_STATE, "the simulation's context was destroyed");
    *c = sim->link->ctx;
    *s = sim->link->s;
    return TOMO_OK;
}

// ================================================================ C ABI
extern "C" {

const char* tomo_last_error(void) { return g_last_error.c_str(); }

int tomo_preset(int preset_id, int grid, tomo_sim_config* out) { return preset(preset_id, grid, out); }

int64_t tomo_farfield_len(const tomo_sim_config* cfg) {
    return cfg ? (int64_t)cfg->n_exc * cfg->n_freq * cfg->n_angles : -1;
}

int32_t tomo_contour_len(const tomo_sim_config* cfg) { return cfg ? contour_len(*cfg) : -1; }

int tomo_ctx_create(int device, uint64_t hbm_budget, tomo_ctx** out) {
    if (!out) return fail(TOMO_E_INVALID, "null output");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(TOMO_E_INVALID, "no such HIP device");
    HIPCHK(hipSetDevice(device));
    std::unique_ptr<tomo_ctx> c(new tomo_ctx());
    c->device = device;
    // set-up of the allocation-free entry: this device's scratch address
    int rs = scratch_address(device, &c->scratch);
    if (rs) return rs;
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (int k = 0; k < 3; ++k) {
        HIPCHK(hipStreamCreateWithFlags(&c->side[k], hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&c->ev_side[k], hipEventDisableTiming | hipEventDisableSystemFence));
    }
    // ordering between this device's streams only: no system-scope fence needed
    HIPCHK(hipEventCreateWithFlags(&c->ev_bar, hipEventDisableTiming | hipEventDisableSystemFence));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    c->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    size_t fr = 0, tot = 0;
    HIPCHK(hipMemGetInfo(&fr, &tot));
    c->budget = hbm_budget ? hbm_budget : (uint64_t)(0.9 * (double)fr);
    HIPCHK(hipEventCreateWithFlags(&c->stg.staged, hipEventDisableTiming));
    HIPCHK(hipEventRecord(c->stg.staged, c->stream));
    int rc = ensure_table(c.get(), 256);
    if (rc) return rc;
    *out = c.release();
    return TOMO_OK;
}

int tomo_ctx_destroy(tomo_ctx* c) {
    if (!c) return TOMO_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    // Every resident slot belongs to exactly one owner: a request (active or
    // finished-but-unprojected) or a tomo_sim handle.  Free each once.
    c->active.clear();
    for (auto& kv : c->reqs) {
        for (Slot* s : kv.second->slots) slot_destroy(c, s);
        kv.second->slots.clear();
    }
    c->reqs.clear();
    c->queue.clear();
    for (auto& link : c->sims) {
        slot_destroy(c, link->s);
        link->s = nullptr;
        link->ctx = nullptr;  // the handle now reports a destroyed context
    }
    c->sims.clear();
    if (c->d_sims) (void)hipFree(c->d_sims);
    if (c->stg.h_items) (void)hipHostFree(c->stg.h_items);
    if (c->stg.d_items) (void)hipFree(c->stg.d_items);
    if (c->stg.h_tiles) (void)hipHostFree(c->stg.h_tiles);
    if (c->stg.d_tiles) (void)hipFree(c->stg.d_tiles);
    if (c->stg.staged) (void)hipEventDestroy(c->stg.staged);
    if (c->stg.h_res) (void)hipHostFree(c->stg.h_res);
    if (c->stg.d_res) (void)hipFree(c->stg.d_res);
    if (c->stg.d_flags) (void)hipFree(c->stg.d_flags);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    (void)drain_timings(c);
    release_cache(c);
    for (int k = 0; k < 3; ++k) {
        if (c->ev_side[k]) (void)hipEventDestroy(c->ev_side[k]);
        if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
    }
    if (c->ev_bar) (void)hipEventDestroy(c->ev_bar);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return TOMO_OK;
}

static int set_schedule_locked(tomo_ctx* c, const tomo_schedule& s) {
    const int T = s.tsteps ? s.tsteps : c->T, S = s.seg_rows ? s.seg_rows : c->S;
    const int W = s.edge_waves ? s.edge_waves : c->W, Wi = s.inner_waves ? s.inner_waves : c->Wi;
    const int K = s.chunk ? s.chunk : c->K, G = s.grid_mult ? s.grid_mult : c->grid_mult;
    const int F = s.form ? s.form : c->form;
    const int ER = s.edge_rows ? s.edge_rows : c->ER;
    if (!tblock_supported(T)) return fail(TOMO_E_INVALID, "unsupported fused step count");
    if (W < 1 || W > 4 || Wi < 1 || Wi > 4 || S < 4 || S > 4096 || K < T || K > 64 || K % T || G < 1 || G > 64 ||
        (F != TOMO_FORM_STREAMS && F != TOMO_FORM_FUSED && F != TOMO_FORM_RESIDENT && F != TOMO_FORM_TILE) ||
        ER < 1 || ER > 4096 ||
        (s.inner_rows != 0 && (s.inner_rows < 4 || s.inner_rows > 4096)))
        return fail(TOMO_E_INVALID, "bad schedule");
    c->T = T; c->S = S; c->W = W; c->Wi = Wi; c->K = K; c->grid_mult = G; c->form = F; c->ER = ER;
    c->Si = s.inner_rows;
    return TOMO_OK;
}

int tomo_set_schedule(tomo_ctx* c, const tomo_schedule* s) {
    if (!c || !s) return fail(TOMO_E_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    if (!c->active.empty()) return fail(TOMO_E_STATE, "schedule cannot change while requests are resident");
    if (s->chunk && s->chunk != c->K && c->cap - (int)c->free_idx.size() > 0)
        return fail(TOMO_E_STATE, "the DFT chunk cannot change while simulations are resident");
    return set_schedule_locked(c, *s);
}

int tomo_get_schedule(tomo_ctx* c, tomo_schedule* s) {
    if (!c || !s) return fail(TOMO_E_INVALID, "null argument");
    s->tsteps = c->T; s->seg_rows = c->S; s->edge_waves = c->W; s->inner_waves = c->Wi;
    s->chunk = c->K; s->grid_mult = c->grid_mult; s->form = c->form; s->edge_rows = c->ER;
    s->inner_rows = c->Si;
    return TOMO_OK;
}

// Seconds for `epochs` epochs of the resident slots under the current schedule.
static int time_epochs(tomo_ctx* c, std::vector<Slot*>& sl, int epochs, double* secs) {
    hipEvent_t a = pool_event(c), b = pool_event(c);
    if (!a || !b) return fail(TOMO_E_HIP, "event");
    for (Slot* s : sl) s->target = std::min(s->su.cfg.steps, s->done + c->K);
    int rc = run_epochs(c, sl, stage(c));  // untimed: rebuilds and uploads the tile list
    if (rc) return rc;
    HIPCHK(hipEventRecord(a, c->stream));
    for (int e = 0; e < epochs; ++e) {
        for (Slot* s : sl) s->target = std::min(s->su.cfg.steps, s->done + c->K);
        rc = run_epochs(c, sl, stage(c));
        if (rc) return rc;
    }
    HIPCHK(hipEventRecord(b, c->stream));
    HIPCHK(hipEventSynchronize(b));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    c->ev_pool.push_back(a);
    c->ev_pool.push_back(b);
    *secs = ms / 1e3;
    return TOMO_OK;
}

int tomo_autotune(tomo_ctx* c, const tomo_sim_config* cfg, int n_sims, tomo_schedule* best, double* rate) {
    if (!c || !cfg || n_sims < 1) return fail(TOMO_E_INVALID, "bad argument");
    std::lock_guard<std::mutex> g(c->mu);
    if (!c->active.empty()) return fail(TOMO_E_STATE, "autotune needs an idle context");
    tomo_sim_config tc = *cfg;
    tc.steps = 1 << 17;  // never runs out during the trials
    std::vector<double> p;
    if (tc.mode == TOMO_MODE_PHYSICAL) {
        p = target_vector(tc);
    } else {
        const int n2 = tc.numpatches * tc.numpatches;
        p.assign((size_t)3 * n2, 0.0);
        for (int i = 0; i < n2; ++i) { p[i] = 3.0; p[n2 + i] = 10.0; }
    }
    std::vector<Slot*> sl;
    int rc = TOMO_OK;
    for (int i = 0; i < n_sims && !rc; ++i) {
        Slot* s = nullptr;
        rc = slot_create(c, tc, p.data(), p.size(), i % tc.n_exc, &s);
        if (!rc) sl.push_back(s);
    }
    const bool prof = c->prof;
    c->prof = false;
    tomo_schedule cur{c->T, c->S, c->W, c->Wi, c->K, c->grid_mult, c->form, c->ER, c->Si};
    // each trial times at least 512 steps: a few launches alone are dominated
    // by first-launch and dispatch effects
    const int epochs = std::max(2, (512 + c->K - 1) / c->K);
    double cells = 0;
    for (Slot* s : sl) cells += (double)s->d.nx * s->d.ny * c->K * epochs;
    double best_t = 1e30;
    auto trial = [&](const tomo_schedule& cand) -> int {
        int r = set_schedule_locked(c, cand);
        if (r) return r;
        double t = 0;
        r = time_epochs(c, sl, epochs, &t);
        if (r) return r;
        if (t < best_t) { best_t = t; cur = cand; }
        return TOMO_OK;
    };
    if (!rc) rc = trial(cur);
    // Seeds measured on MI355X, so that the descent does not start from a
    // poor corner of the space: one simulation (fused, short tiles) and many
    // (four streams, tall tiles; T = 6 where the chunk allows it).
    if (tc.mode == TOMO_MODE_PHYSICAL) {
        const tomo_schedule seeds[3] = {{4, 4, 1, 1, c->K, 16, TOMO_FORM_FUSED, 1, 10},
                                        {4, 32, 1, 4, c->K, 16, TOMO_FORM_STREAMS, 24, 0},
                                        {6, 52, 1, 4, c->K, 32, TOMO_FORM_STREAMS, 24, 0}};
        for (const tomo_schedule& sd : seeds)
            if (!rc && c->K % sd.tsteps == 0) rc = trial(sd);
    }
    const int Ts[] = {1, 2, 4, 6, 8}, Ss[] = {4, 6, 8, 12, 16, 32, 48, 64, 128}, Ws[] = {1, 2, 4}, Gs[] = {2, 4, 8, 16};
    const int ERs[] = {1, 2, 3, 4, 8, 16, 24}, SIs[] = {0, 8, 10, 12, 14, 16, 24};
    const int Fs[] = {TOMO_FORM_STREAMS, TOMO_FORM_FUSED};
    for (int pass = 0; pass < 2 && !rc; ++pass) {
        if (tc.mode == TOMO_MODE_PHYSICAL)
            for (int v : Fs) { tomo_schedule x = cur; x.form = v; if (!rc) rc = trial(x); }
        for (int v : Ts) { tomo_schedule x = cur; x.tsteps = v; if (!rc && c->K % v == 0) rc = trial(x); }
        for (int v : Ss) { tomo_schedule x = cur; x.seg_rows = v; if (!rc) rc = trial(x); }
        for (int v : SIs) { tomo_schedule x = cur; x.inner_rows = v; if (!rc) rc = trial(x); }
        for (int v : ERs) { tomo_schedule x = cur; x.edge_rows = v; if (!rc) rc = trial(x); }
        for (int v : Ws) { tomo_schedule x = cur; x.edge_waves = v; if (!rc) rc = trial(x); }
        for (int v : Ws) { tomo_schedule x = cur; x.inner_waves = v; if (!rc) rc = trial(x); }
        for (int v : Gs) { tomo_schedule x = cur; x.grid_mult = v; if (!rc) rc = trial(x); }
    }
    if (!rc) rc = set_schedule_locked(c, cur);
    c->prof = prof;
    for (Slot* s : sl) slot_destroy(c, s);
    c->stg.tile_key.clear();
    if (rc) return rc;
    if (best) *best = cur;
    if (rate) *rate = cells / best_t;
    return TOMO_OK;
}

int tomo_profile(tomo_ctx* c, int enable) {
    if (!c) return fail(TOMO_E_INVALID, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    int rc = drain_timings(c);
    c->prof = enable != 0;
    c->stats = tomo_stats{};
    return rc;
}

int tomo_use_graphs(tomo_ctx* c, int enable) {
    if (!c) return fail(TOMO_E_INVALID, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    c->use_graph = enable != 0;
    return TOMO_OK;
}

int tomo_get_stats(tomo_ctx* c, tomo_stats* out) {
    if (!c || !out) return fail(TOMO_E_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    int rc = drain_timings(c);
    *out = c->stats;
    out->requests = (int64_t)c->reqs.size();
    out->resident_refused = c->res_refused ? 1 : 0;
    return rc;
}

int tomo_eval(tomo_ctx* c, const tomo_sim_config* cfg, const double* params, size_t n_params, const float* target,
              size_t n_target, double* fitness_out, float* D_out, float* P_out) {
    if (!c || !cfg) return fail(TOMO_E_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    (void)hipSetDevice(c->device);
    return eval_many(c, *cfg, params, 1, n_params, target, n_target, fitness_out, D_out, P_out, true);
}

int tomo_eval_noalloc(tomo_ctx* c, const tomo_sim_config* cfg, const double* params, size_t n_params,
                      const float* target, size_t n_target, double* fitness_out, float* D_out, float* P_out) {
    if (!c || !cfg) return fail(TOMO_E_INVALID, "null argument");
    if (!c->scratch) return fail(TOMO_E_STATE, "no scratch address for this device");
    std::lock_guard<std::mutex> g(c->mu);
    std::lock_guard<std::mutex> gd(g_noalloc_mu[c->device]);
    (void)hipSetDevice(c->device);
    // per-launch timing would create events: it stays off inside this call
    const bool prof = c->prof;
    c->prof = false;
    const int rc = eval_noalloc(c, c->scratch, *cfg, params, n_params, target, n_target, fitness_out, D_out, P_out);
    c->prof = prof;
    return rc;
}

int tomo_eval_batch(tomo_ctx* c, const tomo_sim_config* cfg, const double* params, size_t n_vec, size_t dim,
                    const float* target, size_t n_target, double* fitness_out, float* D_out) {
    if (!c || !cfg) return fail(TOMO_E_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    (void)hipSetDevice(c->device);
    return eval_many(c, *cfg, params, n_vec, dim, target, n_target, fitness_out, D_out, nullptr, false);
}

int tomo_target_farfield(tomo_ctx* c, const tomo_sim_config* cfg, float* D_out) {
    if (!c || !cfg || !D_out) return fail(TOMO_E_INVALID, "null argument");
    if (cfg->mode != TOMO_MODE_PHYSICAL) return fail(TOMO_E_INVALID, "physical mode only");
    std::lock_guard<std::mutex> g(c->mu);
    std::vector<float> t;
    int rc = get_target(c, *cfg, t);
    if (!rc) std::memcpy(D_out, t.data(), t.size() * 4);
    return rc;
}

int64_t tomo_request_bytes(const tomo_sim_config* cfg) { return cfg ? request_bytes(*cfg, 32) : -1; }

int tomo_submit(tomo_ctx* c, const tomo_sim_config* cfg, const double* params, size_t n, uint64_t* ticket) {
    if (!c || !cfg || !ticket || (!params && n)) return fail(TOMO_E_INVALID, "null argument");
    if (cfg->mode != TOMO_MODE_PHYSICAL) return fail(TOMO_E_INVALID, "the request loop serves physical mode");
    int rc = validate(*cfg);
    if (rc) return rc;
    if (n < 8) return fail(TOMO_E_SHORT, "physical mode needs at least 8 parameters");
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(params[i])) return fail(TOMO_E_INVALID, "non-finite parameter");
    std::lock_guard<std::mutex> g(c->mu);
    std::unique_ptr<tomo_request> r(new tomo_request());
    r->ticket = c->next_ticket++;
    r->cfg = *cfg;
    r->params.assign(params, params + n);
    r->t_submit = now_s();
    r->bytes = request_bytes(*cfg, c->K);
    *ticket = r->ticket;
    c->queue.push_back(r->ticket);
    c->reqs[r->ticket] = std::move(r);
    return TOMO_OK;
}

int tomo_poll(tomo_ctx* c, uint64_t ticket, double* fitness_out, double* latency_s) {
    if (!c) return fail(TOMO_E_INVALID, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    auto it = c->reqs.find(ticket);
    if (it == c->reqs.end()) return fail(TOMO_E_INVALID, "unknown ticket");
    tomo_request* r = it->second.get();
    if (r->state == 0 || r->state == 1) return TOMO_E_PENDING;
    if (r->state == 4) {  // reported once, then the record is gone like a finished one's
        c->reqs.erase(it);
        return fail(TOMO_E_CANCELLED, "request cancelled");
    }
    if (fitness_out) *fitness_out = r->fitness;
    if (latency_s) *latency_s = r->t_done - r->t_submit;
    const int st = r->status;
    if (st != TOMO_OK) g_last_error = "request failed";
    c->reqs.erase(it);
    return st;
}

int tomo_progress(tomo_ctx* c, int* n_active) {
    if (!c) return fail(TOMO_E_INVALID, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    (void)hipSetDevice(c->device);
    return progress_once(c, n_active);
}

int tomo_wait(tomo_ctx* c, uint64_t ticket, double* fitness_out, double* latency_s) {
    for (;;) {
        int rc = tomo_poll(c, ticket, fitness_out, latency_s);
        if (rc != TOMO_E_PENDING) return rc;
        int na = 0;
        rc = tomo_progress(c, &na);
        if (rc) return rc;
    }
}

int tomo_cancel(tomo_ctx* c, uint64_t ticket) {
    if (!c) return fail(TOMO_E_INVALID, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    auto it = c->reqs.find(ticket);
    if (it == c->reqs.end()) return fail(TOMO_E_INVALID, "unknown ticket");
    tomo_request* r = it->second.get();
    if (r->state == 0) {
        // leaves the admission queue now: tomo_poll may erase the record before the next epoch
        c->queue.erase(std::remove(c->queue.begin(), c->queue.end(), ticket), c->queue.end());
        r->state = 4;
        r->t_done = now_s();
        return TOMO_OK;
    }
    if (r->state == 1) {
        std::vector<Slot*> still;
        for (Slot* s : c->active)
            if (s->req != (long)r->ticket) still.push_back(s);
        c->active.swap(still);
        for (Slot* s : r->slots) slot_destroy(c, s);
        r->slots.clear();
        r->state = 4;
        r->t_done = now_s();
        return TOMO_OK;
    }
    return fail(TOMO_E_STATE, "request already finished");
}

int tomo_sim_create(tomo_ctx* c, const tomo_sim_config* cfg, const double* params, size_t n, int exc,
                    tomo_sim** out) {
    if (!c || !cfg || !out) return fail(TOMO_E_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    Slot* s = nullptr;
    int rc = slot_create(c, *cfg, params, n, exc, &s);
    if (rc) return rc;
    std::shared_ptr<SimLink> link = std::make_shared<SimLink>();
    link->ctx = c;
    link->s = s;
    c->sims.push_back(link);
    *out = new tomo_sim{link};
    return TOMO_OK;
}

int tomo_sim_destroy(tomo_sim* sim) {
    if (!sim) return TOMO_OK;
    tomo_ctx* c = sim->link ? sim->link->ctx : nullptr;
    if (c) {  // else the context was destroyed first and already freed the slot
        std::lock_guard<std::mutex> g(c->mu);
        slot_destroy(c, sim->link->s);
        sim->link->s = nullptr;
        sim->link->ctx = nullptr;
        c->sims.erase(std::remove(c->sims.begin(), c->sims.end(), sim->link), c->sims.end());
    }
    delete sim;
    return TOMO_OK;
}

int tomo_sim_step(tomo_sim* sim, int nsteps) {
    if (nsteps < 0) return fail(TOMO_E_INVALID, "bad argument");
    tomo_ctx* c = nullptr;
    Slot* s = nullptr;
    int rc = sim_live(sim, &c, &s);
    if (rc) return rc;
    std::lock_guard<std::mutex> g(c->mu);
    s->target = std::min(s->su.cfg.steps, s->done + nsteps);
    std::vector<Slot*> one{s};
    rc = run_epochs(c, one, stage(c));
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return TOMO_OK;
}

int tomo_sim_fields(tomo_sim* sim, float* ez, float* hx, float* hy, float* jp) {
    tomo_ctx* c = nullptr;
    Slot* s = nullptr;
    {
        const int rc0 = sim_live(sim, &c, &s);
        if (rc0) return rc0;
    }
    std::lock_guard<std::mutex> g(c->mu);
    const DevSim& d = s->d;
    const int nx = d.nx, ny = d.ny, par = s->parity;
    hipStream_t st = c->stream;
    if (ez) HIPCHK(hipMemcpy2DAsync(ez, (nx + 1) * 4, d.ez[par], d.pitch * 4, (nx + 1) * 4, ny + 1, hipMemcpyDeviceToHost, st));
    if (hx) HIPCHK(hipMemcpy2DAsync(hx, nx * 4, d.hx[par], d.pitch * 4, nx * 4, ny, hipMemcpyDeviceToHost, st));
    if (hy) HIPCHK(hipMemcpy2DAsync(hy, nx * 4, d.hy[par], d.pitch * 4, nx * 4, ny, hipMemcpyDeviceToHost, st));
    if (jp) {
        std::vector<float> cj((size_t)std::max(s->su.jp_len, 1));
        HIPCHK(hipMemcpyAsync(cj.data(), d.jp[par], cj.size() * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        std::fill(jp, jp + (size_t)(nx + 1) * (ny + 1), 0.f);
        for (int y = 0; y <= ny; ++y) {
            const int lo = s->su.span[3 * y], hi = s->su.span[3 * y + 1], off = s->su.span[3 * y + 2];
            for (int x = lo; x < hi; ++x) jp[(size_t)y * (nx + 1) + x] = cj[off + x - lo];
        }
    }
    HIPCHK(hipStreamSynchronize(st));
    return TOMO_OK;
}

int tomo_sim_materials(tomo_sim* sim, uint8_t* mat, float* coef, int* n_mat) {
    tomo_ctx* c = nullptr;
    Slot* s = nullptr;
    {
        const int rc0 = sim_live(sim, &c, &s);
        if (rc0) return rc0;
    }
    std::lock_guard<std::mutex> g(c->mu);
    const DevSim& d = s->d;
    if (mat)
        HIPCHK(hipMemcpy2D(mat, d.nx + 1, d.mat, d.pitch, d.nx + 1, d.ny + 1, hipMemcpyDeviceToHost));
    if (coef) std::memcpy(coef, d.coef, sizeof(float) * d.nmat * TOMO_NCOEF);
    if (n_mat) *n_mat = d.nmat;
    return TOMO_OK;
}

int tomo_sim_dft(tomo_sim* sim, float* J, float* M) {
    tomo_ctx* c = nullptr;
    Slot* s = nullptr;
    {
        const int rc0 = sim_live(sim, &c, &s);
        if (rc0) return rc0;
    }
    std::lock_guard<std::mutex> g(c->mu);
    const DevSim& d = s->d;
    const size_t n = (size_t)d.nfreq * d.P;
    const bool f64 = (d.flags & TOMO_FLAG_DFT_FP64) != 0;
    std::vector<float> a(4 * n), b(4 * n);
    HIPCHK(hipMemcpy(a.data(), d.accJ, 16 * n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(b.data(), d.accM, 16 * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) {
        if (f64) {
            const double* da = (const double*)a.data();
            const double* db = (const double*)b.data();
            if (J) { J[2 * i] = (float)da[2 * i]; J[2 * i + 1] = (float)da[2 * i + 1]; }
            if (M) { M[2 * i] = (float)db[2 * i]; M[2 * i + 1] = (float)db[2 * i + 1]; }
        } else {
            if (J) { J[2 * i] = a[4 * i]; J[2 * i + 1] = a[4 * i + 1]; }
            if (M) { M[2 * i] = b[4 * i]; M[2 * i + 1] = b[4 * i + 1]; }
        }
    }
    return TOMO_OK;
}

int tomo_sim_farfield(tomo_sim* sim, float* D, float* P) {
    if (!D) return fail(TOMO_E_INVALID, "null argument");
    tomo_ctx* c = nullptr;
    Slot* s = nullptr;
    {
        const int rc0 = sim_live(sim, &c, &s);
        if (rc0) return rc0;
    }
    std::lock_guard<std::mutex> g(c->mu);
    const size_t L = (size_t)s->su.cfg.n_freq * s->su.cfg.n_angles;
    float *dD = nullptr, *dP = nullptr;
    HIPCHK(hipMalloc(&dD, L * 4));
    HIPCHK(hipMalloc(&dP, L * 4));
    std::vector<Slot*> one{s};
    int rc = project(c, one, {0}, dD, dP);
    if (!rc) {
        if (hipMemcpy(D, dD, L * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(TOMO_E_HIP, "copy");
        if (!rc && P && hipMemcpy(P, dP, L * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = fail(TOMO_E_HIP, "copy");
    }
    (void)hipFree(dD);
    (void)hipFree(dP);
    return rc;
}

}  // extern "C"
