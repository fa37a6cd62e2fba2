// kernels.h — host-callable launchers of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "setup.h"
#include "tomo_internal.h"

namespace tomo {

// Steps fused per launch that have a compiled kernel.
bool tblock_supported(int T);

// Time the next update, fold, projection or fitness launch made by the calling
// thread: its dispatch packet carries `start` and `stop` (hipExtLaunchKernelGGL),
// so timing puts no marker packets between kernels.  (nullptr, nullptr) disarms.
void arm_launch_events(hipEvent_t start, hipEvent_t stop);

// Time-blocked E/H update of every tile in `tiles` (T steps, one launch):
// the general kernel (CPML, contour samples, edges, ref-compat physics) and
// the interior kernel (physical mode, tiles away from every boundary).
hipError_t launch_tblock(int T, int mode, const TileItem* tiles, int ntiles, const LaunchItem* items,
                         const DevSim* sims, int ordinal, int waves, int grid, hipStream_t st);
hipError_t launch_tedge(int T, const TileItem* tiles, int ntiles, const LaunchItem* items, const DevSim* sims,
                        int ordinal, int waves, int grid, hipStream_t st);
hipError_t launch_trow(int T, const TileItem* tiles, int ntiles, const LaunchItem* items, const DevSim* sims,
                       int ordinal, int waves, int grid, hipStream_t st);
hipError_t launch_tcol(int T, const TileItem* tiles, int ntiles, const LaunchItem* items, const DevSim* sims,
                       int ordinal, int waves, int grid, hipStream_t st);
hipError_t launch_tinner(int T, const TileItem* tiles, int ntiles, const LaunchItem* items, const DevSim* sims,
                         int ordinal, int waves, int grid, hipStream_t st);
// Rows a row-sweep tile streams (its output rows y1 - y0 plus T ghost rows on
// each side) are at most kMaxTileRows: the sweeps of the interior, row-edge,
// column-edge and corner tiles hold the rows' dispersive spans one per lane.
constexpr int kMaxTileRows = 64;
// All tile classes of physical-mode simulations in one launch (TileItem::cls).
hipError_t launch_tfused(int T, const TileItem* tiles, int ntiles, const LaunchItem* items, const DevSim* sims,
                         int ordinal, int waves, int grid, hipStream_t st);

// Tile update form: one workgroup of kTileWaveCount waves per TileItem holds a
// 64-lane band x (kTileWaveCount * rows_per_wave) rows in registers and
// advances it one time block (T steps); every tile row span y1 - y0 must be
// at most kTileWaveCount * rows_per_wave - 2T.
constexpr int kTileWaveCount = 8;
bool twave_supported(int T, int rows_per_wave);
hipError_t launch_twave(int T, int rows_per_wave, const TileItem* tiles, int ntiles, const LaunchItem* items,
                        const DevSim* sims, int ordinal, int grid, hipStream_t st);

// Resident update form: one cooperative launch advances every simulation of
// `items` by its epoch (all time blocks), one workgroup per ResTile, all
// co-resident.  `done` (ntiles counters) must start at zero; `err` is set to 1
// if a workgroup gave up waiting (after spin_limit polls) and stays set.
constexpr int kResidentWaves = 8;  // waves per workgroup of the resident kernel
bool tres_supported(int T, int rows_per_wave);
// Tiles one cooperative launch can hold on `cus` compute units (0: none).
int tres_max_tiles(int T, int rows_per_wave, int mode, int cus);
hipError_t launch_tres(int T, int rows_per_wave, int mode, const ResTile* tiles, int ntiles,
                       const LaunchItem* items, const DevSim* sims, int* done, int* err, int spin_limit,
                       hipStream_t st);

// Fold one chunk of surface samples into the running DFT (Kahan FP32 or FP64).
hipError_t launch_fold(bool fp64, const LaunchItem* items, int nitems, const DevSim* sims, int max_p, int max_f,
                       hipStream_t st);

// Far-field projection of a group of simulations that share grid, contour and
// frequencies, on the matrix cores: D/P at [(sim * n_freq + f) * n_angles + a];
// kinv[f] = {wavenumber, far-field normaliser}.
hipError_t launch_project_mfma(const int* sim_idx, int n_sims, int n_freq, const DevSim* sims, const float2* kinv,
                               int n_angles, float eta0, float* D, float* Pph, hipStream_t st);

// FP64 fitness reduction, one block per evaluation:
// fit[e] = -sum_i (D[e*L+i] - tgt[i])^2 / norm  (+ optional wrapped phase term).
// terms bit 0: phase term; bit 1: relative misfit (divide each term by tgt^2).
hipError_t launch_fitness(const float* D, const float* Pph, const float* tgt, const float* tgt_phase, int L,
                          int n_eval, double norm, int terms, double* fit, hipStream_t st);

// Material rasterisation (bit-exact with the oracle).
hipError_t launch_raster_ref(uint8_t* mat, int pitch, int nx, int ny, const int* lo, const int* hi,
                             const int* xc, const int* yc, int npatch, hipStream_t st);
hipError_t launch_raster_phys(uint8_t* mat, int pitch, int nx, int ny, float dxf, float dyf,
                              const PhantomGeom& g, hipStream_t st);

// Zero-fill (a kernel, so that it is safe inside captured graphs).
hipError_t launch_zero(void* p, size_t bytes, hipStream_t st);

}  // namespace tomo
