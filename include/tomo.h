/*
 * tomo.h — C ABI of the MI355X-native 2D TMz FDTD fitness evaluation.
 *
 * The reference exposes exactly one interface on this path:
 *     double FDTD_GPU(const std::vector<double>&)   (FDTD_GPU.cuh:8, FDTD_GPU.cu:650)
 *     double FDTD_CPU(const std::vector<double>&)   (FDTD_CPU.h:8,  FDTD_CPU.cxx:669)
 * a free C++ function that maps a PSO/DE parameter vector to a fitness (higher
 * is better), allocating and freeing everything inside the call, with all
 * physics fixed by compile-time macros (FDTD_common.hxx:3-65).
 *
 * This header replaces that with plain C entry points (no C++ types, no torch
 * types, status codes instead of exit(-1)):
 *   - tomo_sim_config : the run-time replacement of the FDTD_common.hxx macros;
 *   - tomo_preset     : the reference configuration and the benchmark workloads;
 *   - tomo_eval       : one fitness evaluation, the C-ABI twin of FDTD_GPU();
 *   - tomo_eval_batch : many evaluations in one batched launch sequence;
 *   - tomo_submit/poll/wait/progress : continuous-batching request loop;
 *   - tomo_sim_*      : single-simulation access used by parity tests
 *                       (fields, running-DFT accumulators, material map).
 * The C++ drop-in shim `double FDTD_GPU(const std::vector<double>&)` is
 * exported by the same library (see tomography_amd/csrc/fdtd_gpu_shim.cpp).
 *
 * Every function returns TOMO_OK (0) or a negative TOMO_E* code; the message
 * of the last failure on the calling thread is available from
 * tomo_last_error().  Nothing throws across this boundary and nothing exits.
 */
#ifndef TOMO_H
#define TOMO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define TOMO_OK 0
#define TOMO_E_INVALID -1   /* bad argument / configuration / parameter     */
#define TOMO_E_SHORT -2     /* parameter vector too short (reference: vector::at throws) */
#define TOMO_E_HIP -3       /* HIP runtime error                            */
#define TOMO_E_NOMEM -4     /* HBM budget or allocation failure             */
#define TOMO_E_NONFINITE -5 /* NaN/Inf detected in a physical-mode result   */
#define TOMO_E_CANCELLED -6 /* request cancelled                           */
#define TOMO_E_PENDING -7   /* request not finished yet (tomo_poll)         */
#define TOMO_E_STATE -8     /* call not valid in the current state          */

/* ---- modes and flags --------------------------------------------------- */
/* Reference-compatible physics: scattered-field plane-wave excitation, the
 * reference CPML quirks, Debye Jp as in E_field_update (FDTD_GPU.cu:358-434),
 * NaN clamp, NTFF/far field as in N2FPostProcess (FDTD_GPU.cu:553-634). */
#define TOMO_MODE_REF_COMPAT 0
/* Physical mode (benchmark workloads): total field, soft differentiated-
 * Gaussian point source per antenna, standard CPML, Debye ADE everywhere. */
#define TOMO_MODE_PHYSICAL 1

/* Reference CPU-path variants (FDTD_CPU.cxx), selectable in ref-compat mode. */
#define TOMO_FLAG_FWF_ABS_ORIGIN 0x1  /* incident wave uses absolute cell indices (FDTD_CPU.cxx:316) */
#define TOMO_FLAG_CPU_COEFS 0x2       /* no Debye; Cezeic/Cezeip CPU form (FDTD_CPU.cxx:893-905)      */
#define TOMO_FLAG_NO_NAN_CLAMP 0x4    /* do not clamp NaN to 0 (FDTD_CPU.cxx:272-277)                 */
#define TOMO_FLAG_PARAMS_EPS_SIGMA 0x8 /* parameter layout {eps_r[n^2], sigma[n^2]} (FDTD_CPU.cxx:681-698) */
#define TOMO_FLAG_PHASE_FITNESS 0x10  /* add the phase term sketched at FDTD_GPU.cu:1009-1018          */
#define TOMO_FLAG_DFT_FP64 0x20       /* running-DFT accumulators in FP64 instead of Kahan FP32      */
#define TOMO_FLAG_REL_FITNESS 0x40    /* relative misfit (D-m)^2/m^2 per term (FDTD_CPU.cxx:844)      */
#define TOMO_FLAG_FITNESS_PER_EXC 0x80 /* normalise by n_exc only, as fitness() at FDTD_common.cxx:53-61 */
#define TOMO_FLAG_CPU_REF_VARIANT                                                              \
    (TOMO_FLAG_FWF_ABS_ORIGIN | TOMO_FLAG_CPU_COEFS | TOMO_FLAG_NO_NAN_CLAMP |             \
     TOMO_FLAG_PARAMS_EPS_SIGMA | TOMO_FLAG_REL_FITNESS)

/* ---- presets ----------------------------------------------------------- */
#define TOMO_PRESET_REF_GPU 0   /* FDTD_common.hxx macros + FDTD_GPU semantics, 27 parameters  */
#define TOMO_PRESET_REF_CPU 1   /* FDTD_common.hxx macros + FDTD_CPU semantics, 162 parameters */
#define TOMO_PRESET_CFG1 2      /* 256^2, dx 0.5 mm, 2048 steps, 1 antenna, 8 f x 180 angles   */
#define TOMO_PRESET_CFG2 3      /* 1024^2, dx 0.25 mm, 8192 steps, 1 antenna                   */
#define TOMO_PRESET_CFG3 4      /* 512^2, dx 0.5 mm, 4096 steps, 16 antennas, 16 f x 360 angles */
#define TOMO_PRESET_CFG4 5      /* uneven-batch template: 256^2..2048^2 chosen per request      */
#define TOMO_PRESET_CFG5 6      /* serving template (same physics as CFG4)                      */

/*
 * Run-time replacement of the FDTD_common.hxx macros.  Geometry is square in
 * every preset; nx != ny is accepted in physical mode only.
 */
typedef struct tomo_sim_config {
    int32_t mode;        /* TOMO_MODE_*                                          */
    int32_t flags;       /* TOMO_FLAG_*                                          */
    int32_t nx, ny;      /* Yee cells; Ez nodes are (nx+1) x (ny+1)              */
    int32_t steps;       /* time steps per excitation (number_of_time_steps)     */
    int32_t n_exc;       /* excitation angles (ref) or antennas (physical)       */
    int32_t ncpml;       /* CPML thickness in cells (ncells)                     */
    int32_t ntff_offset; /* NTFF contour distance from the edge (NF2FFdistfromboundary) */
    int32_t n_freq;      /* numberoffrequencies                                  */
    int32_t n_angles;    /* numberofobservationangles                            */
    int32_t numpatches;  /* ref: patch grid edge (3 GPU path, 9 CPU path)        */
    int32_t n_params;    /* minimum parameter-vector length                      */
    double dx, dy, dt;
    double f_start, f_stop;       /* startfrequency, stopfrequency               */
    double cpml_order;            /* npml (ref) / grading order m (physical)     */
    double kappa_max;             /* kmax                                        */
    double alpha_max;             /* alpha_max                                   */
    double sigma_max;             /* sigma_max (ref: (npml+1)/(150 pi dx))        */
    double pulse_tau, pulse_t0;   /* tau, t0 of the excitation pulse             */
    double debye_tau;             /* ref: relaxation_time                        */
    double breast_radius;         /* ref: disc radius; physical: outer radius    */
    double skin_thickness;        /* physical                                    */
    double src_radius;            /* physical: antenna ring radius               */
    double ff_norm;               /* ref: the 33.329 far-field normaliser        */
    uint64_t seed;                /* physical: phantom seed                      */
} tomo_sim_config;

/* Fill *out with a preset.  `grid` overrides nx = ny (and the step count as
 * clamp(8*grid, 2048, 16384)) for the CFG4/CFG5 templates; pass 0 otherwise. */
int tomo_preset(int preset, int grid, tomo_sim_config *out);

/* Length of the far-field record of one evaluation: n_exc * n_freq * n_angles,
 * laid out [exc][freq][angle] exactly like D in FDTD_GPU.cu:975. */
int64_t tomo_farfield_len(const tomo_sim_config *cfg);

/* Number of NTFF contour points: 2*nx + 2*ny - 8*D - 4 (size_NF2FF_total). */
int32_t tomo_contour_len(const tomo_sim_config *cfg);

/* ---- context ----------------------------------------------------------- */
typedef struct tomo_ctx tomo_ctx;

/* Create a context on HIP device `device` with a device-memory budget for
 * resident simulations (bytes; 0 = 90% of free memory). */
int tomo_ctx_create(int device, uint64_t hbm_budget, tomo_ctx **out);
/* Destroy a context: frees every resident simulation it holds (requests in
 * any state, and the slots of tomo_sim handles).  A tomo_sim handle outlives
 * its context safely: afterwards tomo_sim_destroy only frees the handle and
 * every other tomo_sim_* call returns TOMO_E_STATE.  Must not run
 * concurrently with another call on the same context or its handles. */
int tomo_ctx_destroy(tomo_ctx *ctx);
const char *tomo_last_error(void);

/* Kernel schedule (the autotuner writes it; a zero field keeps the current
 * value).  tsteps: time steps fused per launch (1, 2, 4, 6, 8); seg_rows:
 * output rows per tile; edge_waves / inner_waves: waves per workgroup of the
 * edge and interior update kernels (1-4); chunk: running-DFT fold interval in
 * steps (a multiple of tsteps, <= 64); grid_mult: persistent workgroups per CU;
 * form: how a physical-mode time block is launched, TOMO_FORM_STREAMS (one
 * kernel per tile class on four streams, joined per block) or
 * TOMO_FORM_FUSED (one launch of a kernel that runs every class);
 * edge_rows: output rows per corner and row-edge tile (the latency-critical
 * boundary tiles; at most seg_rows). */
#define TOMO_FORM_STREAMS 1
#define TOMO_FORM_FUSED 2
/* TOMO_FORM_RESIDENT: one cooperative launch per DFT chunk in which each
 * workgroup keeps one tile of a simulation (a 64-lane band, seg_rows rows per
 * wave, 8 waves) in registers through every time block, exchanging only a
 * T-deep frame with its neighbours; seg_rows 8 or 12, tsteps 2, 4 or 8.  Runs
 * that hold more tiles than the device can keep co-resident use the fused
 * form instead, and so does a context the device refuses a cooperative launch
 * (tomo_stats.resident_refused). */
#define TOMO_FORM_RESIDENT 3
/* TOMO_FORM_TILE: one launch per time block in which each workgroup of 8 waves
 * loads a whole tile (a 64-lane band x 8 * seg_rows rows, seg_rows 4, 5 or 6
 * rows per wave) into registers, advances it tsteps steps (2, 4, 6 or 8) with
 * the waves coupled through LDS, and stores its owned cells.  Physical mode;
 * reference-compatible simulations and unsupported pairs use the fused form. */
#define TOMO_FORM_TILE 4
typedef struct tomo_schedule {
    int32_t tsteps, seg_rows, edge_waves, inner_waves, chunk, grid_mult, form, edge_rows;
    /* Output rows of an interior tile of a physical-mode simulation; 0 = seg_rows
     * (which then sets the column-edge and the interior tiles alike). */
    int32_t inner_rows;
} tomo_schedule;
int tomo_set_schedule(tomo_ctx *ctx, const tomo_schedule *s);
int tomo_get_schedule(tomo_ctx *ctx, tomo_schedule *s);
/* The schedule the FDTD_GPU / FDTD_CPU shims install on their context: the
 * fastest measured at the drop-in size (401^2, 4 excitations advancing
 * together): resident tiles of 12 rows, 8 steps per frame exchange, 64-step
 * DFT chunks.  The DFT chunk sets the FP32 partial sums of the running DFT,
 * so contexts compare bit for bit only under equal chunks. */
#define TOMO_DROPIN_SCHEDULE {8, 12, 1, 1, 64, 16, TOMO_FORM_RESIDENT, 1, 0}

/* Autotuner: time candidate schedules on n_sims resident simulations of
 * `cfg` (coordinate descent over form, tsteps, seg_rows, inner_rows,
 * edge_rows, edge/inner waves and grid_mult; the chunk is kept), install the fastest and return it with its
 * measured update rate in cell-steps per second. */
int tomo_autotune(tomo_ctx *ctx, const tomo_sim_config *cfg, int n_sims, tomo_schedule *best,
                  double *cell_steps_per_s);

/* Kernel timing: when enabled, every kernel launch is bracketed by HIP events
 * on the stream it is launched on, and its duration is credited to its
 * kernel.  The update kernels also count the cell-steps they advance and
 * their algorithmic HBM bytes: per cell-step 24 (Ez, Hx, Hy read and written
 * once) + 1 (material id) + 8 on Debye cells (Jp) + 16 on CPML-strip cells
 * (the psi pair).  `update` is the wall time of the whole update phase (all
 * streams, from the first launch of an epoch to the join of its last block). */
enum {
    TOMO_KERNEL_GENERAL = 0, /* k_tblock: every tile of a reference-compatible simulation */
    TOMO_KERNEL_CORNER = 1,  /* k_tedge: corner tiles (and all tiles of small physical grids) */
    TOMO_KERNEL_ROW = 2,     /* k_trow: top / bottom frame between the corners */
    TOMO_KERNEL_COL = 3,     /* k_tcol: left / right frame between the corners */
    TOMO_KERNEL_INNER = 4,   /* k_tinner: clean interior bands */
    TOMO_KERNEL_FOLD = 5,    /* k_fold: running-DFT fold of a chunk of samples */
    TOMO_KERNEL_PROJECT = 6, /* k_project_mfma: far-field projection on the matrix cores */
    TOMO_KERNEL_FITNESS = 7, /* k_fitness: FP64 fitness reduction */
    TOMO_KERNEL_FUSED = 8,   /* k_tfused: every tile class in one launch (TOMO_FORM_FUSED) */
    TOMO_KERNEL_RESIDENT = 9, /* k_tres: a whole epoch in one cooperative launch (TOMO_FORM_RESIDENT) */
    TOMO_KERNEL_TILE = 10,    /* k_twave: register tiles, one launch per time block (TOMO_FORM_TILE) */
    TOMO_KERNEL_COUNT = 11
};
typedef struct tomo_kernel_stats {
    int64_t launches;
    double ms;         /* summed launch durations */
    double cell_steps; /* update kernels only */
    double alg_bytes;  /* update kernels only */
} tomo_kernel_stats;
typedef struct tomo_stats {
    tomo_kernel_stats kernel[TOMO_KERNEL_COUNT];
    tomo_kernel_stats update; /* launches = epochs */
    /* request records the context holds: queued, resident, or finished or
     * cancelled and not yet collected by tomo_poll / tomo_wait */
    int64_t requests;
    /* 1 once the device has refused this context a cooperative launch (one
     * process at a time may hold a device's cooperative queue): its
     * TOMO_FORM_RESIDENT epochs then run the fused form, with the same results */
    int64_t resident_refused;
} tomo_stats;
int tomo_profile(tomo_ctx *ctx, int enable); /* (re)starts the counters */
/*
 * Replay an epoch whose simulations all advance a full DFT chunk from one
 * captured hipGraph (default off: measured no faster than direct launches).
 * Profiling launches directly.  The graph is
 * re-captured whenever the set of resident simulations or the schedule changes.
 */
int tomo_use_graphs(tomo_ctx *ctx, int enable);
int tomo_get_stats(tomo_ctx *ctx, tomo_stats *out);

/* ---- evaluation -------------------------------------------------------- */
/*
 * One fitness evaluation (the FDTD_GPU() equivalent): rasterise the parameter
 * vector into a phantom, run n_exc simulations one after another (one
 * simulation per launch sequence: the "plain" path), project the far field
 * and reduce the fitness in FP64 against `target` (n_target values laid out
 * like D; NULL in physical mode selects the seeded target phantom).
 * D_out / P_out (optional, tomo_farfield_len floats) receive the far field
 * and its phase.
 */
int tomo_eval(tomo_ctx *ctx, const tomo_sim_config *cfg, const double *params,
              size_t n_params, const float *target, size_t n_target,
              double *fitness_out, float *D_out, float *P_out);

/*
 * No counterpart in the reference.  A variant of tomo_eval for callers that
 * cannot allocate device memory: code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  Same arguments and the same
 * results as tomo_eval, bit for bit, with two limits: `target` is required
 * (physical mode: fetch it once beforehand with tomo_target_farfield), and one
 * excitation's simulation plus its tables must fit TOMO_NOALLOC_SCRATCH_BYTES
 * (grids up to 1024^2 do; larger ones return TOMO_E_NOMEM).  The call makes no
 * device or pinned-host allocation: every device buffer of the evaluation
 * lives in one array compiled into the library, one copy per device, held
 * until the process ends.  Calls on one device share that array (which keeps
 * the last call's contents; what must start at zero is cleared each call), so
 * overlapping calls would race: the library serialises them per device within
 * a process.  Per-launch timing (tomo_profile) is not recorded inside it.
 */
#define TOMO_NOALLOC_SCRATCH_BYTES (64u << 20)
int tomo_eval_noalloc(tomo_ctx *ctx, const tomo_sim_config *cfg, const double *params,
                      size_t n_params, const float *target, size_t n_target,
                      double *fitness_out, float *D_out, float *P_out);

/*
 * Batched evaluation of n_vec parameter vectors (row-major, `dim` each) with
 * a shared configuration: all n_vec * n_exc simulations (as many as the HBM
 * budget admits at once) advance together in the same time-blocked launches.
 * Results equal tomo_eval's for each vector bit for bit.
 */
int tomo_eval_batch(tomo_ctx *ctx, const tomo_sim_config *cfg, const double *params,
                    size_t n_vec, size_t dim, const float *target, size_t n_target,
                    double *fitness_out, float *D_out);

/*
 * The objective entry (the FDTD_common.hxx-level role of FDTD_GPU/FDTD_CPU):
 * parameter vector + preset -> fitness, shared by every caller.  `measurement`
 * may be NULL: REF_GPU / REF_CPU then use the packaged measurement files (the
 * arrays FDTD_GPU.cu:996 and FDTD_CPU.cxx:789 hard-code; see the C++ shims)
 * and physical presets their seeded target phantom.  `flags` are OR-ed into
 * the preset's (e.g. TOMO_FLAG_FITNESS_PER_EXC | TOMO_FLAG_REL_FITNESS for the
 * generic fitness() of FDTD_common.cxx:53-61).
 */
int tomo_fitness(tomo_ctx *ctx, int preset, int32_t flags, const double *params, size_t n_params,
                 const float *measurement, size_t n_measurement, double *fitness_out);

/* Far field of the seeded target phantom of a physical configuration. */
int tomo_target_farfield(tomo_ctx *ctx, const tomo_sim_config *cfg, float *D_out);

/* ---- continuous-batching request loop ---------------------------------- */
/*
 * tomo_submit queues one evaluation request (its own configuration, so
 * requests of different grid sizes mix) and returns a ticket.  Requests are
 * admitted into resident device slots when the HBM budget allows, advance in
 * shared time-blocked epochs with everything else resident, and retire as
 * soon as their last step is done.  tomo_progress runs one epoch (callers
 * without a thread of their own); tomo_wait blocks until a ticket finishes
 * (driving epochs itself).  Latency = submit -> completion, in seconds.
 */
int tomo_submit(tomo_ctx *ctx, const tomo_sim_config *cfg, const double *params,
                size_t n_params, uint64_t *ticket_out);
int tomo_poll(tomo_ctx *ctx, uint64_t ticket, double *fitness_out, double *latency_s);
int tomo_wait(tomo_ctx *ctx, uint64_t ticket, double *fitness_out, double *latency_s);
int tomo_cancel(tomo_ctx *ctx, uint64_t ticket);
int tomo_progress(tomo_ctx *ctx, int *n_active_out);
/* Resident device bytes one request of this configuration needs. */
int64_t tomo_request_bytes(const tomo_sim_config *cfg);

/* ---- asynchronous search drivers ---------------------------------------- */
/*
 * Population search over the request loop, replacing the reference's TAO
 * master/worker drivers (test.cxx:79-121: ParticleSwarmMPI / DifferentialEvolutionMPI
 * calling FDTD_GPU through MPI workers).  Every evaluation is a tomo_submit
 * request; up to `inflight` are resident at once and a new one is issued as
 * soon as a result returns (asynchronous, steady state).
 *   TOMO_SEARCH_DE  : DE/rand/1/bin, differential weight F, crossover rate CR;
 *                     a trial replaces its member when its fitness is >=.
 *   TOMO_SEARCH_PSO : particle swarm with constriction (chi 0.7298, c1 = c2 = 1.49618).
 * Physical-mode configurations only (the request loop serves physical mode).
 * best_x receives the best vector (dim values).
 */
#define TOMO_SEARCH_DE 1
#define TOMO_SEARCH_PSO 2
typedef struct tomo_search_config {
    int32_t method;    /* TOMO_SEARCH_* */
    int32_t pop;       /* population size (>= 4) */
    int32_t dim;       /* parameter-vector length (>= 8) */
    int32_t max_evals; /* evaluations to run (>= pop) */
    int32_t inflight;  /* evaluations resident at once (0 = pop) */
    int32_t pad;
    double F, CR;      /* DE weight and crossover rate (0 = 0.5, 0.9) */
    uint64_t seed;
} tomo_search_config;
typedef struct tomo_search_result {
    double best_fitness;
    int64_t evals, failed;
    double wall_s, evals_per_s, p50_latency_s, p99_latency_s;
} tomo_search_result;
int tomo_search_run(tomo_ctx *ctx, const tomo_sim_config *cfg, const double *lo, const double *hi,
                    const tomo_search_config *sc, double *best_x, tomo_search_result *res);

/* ---- single-simulation access (parity tests) --------------------------- */
typedef struct tomo_sim tomo_sim;
/* Build one simulation of excitation/antenna `exc` of the phantom `params`. */
int tomo_sim_create(tomo_ctx *ctx, const tomo_sim_config *cfg, const double *params,
                    size_t n_params, int exc, tomo_sim **out);
int tomo_sim_destroy(tomo_sim *sim);
/* Advance by nsteps (must keep the running DFT chunk boundaries consistent:
 * the accumulated DFT is folded at every chunk boundary and at the end). */
int tomo_sim_step(tomo_sim *sim, int nsteps);
/* Copy fields to host: Ez (nx+1)*(ny+1), Hx and Hy nx*ny, row-major with the
 * reference indexing x + y*(nx+1) and x + y*nx (FDTD_common.hxx:65). */
int tomo_sim_fields(tomo_sim *sim, float *ez, float *hx, float *hy, float *jp);
/* Material id map, (nx+1)*(ny+1) bytes, and the per-material coefficient
 * table (n_mat x 8 floats: Ceze, Cezh, Cezeic, Cezeip, Cezjp, Kp, Beta_p/dt,
 * dispersive flag); `coef` must hold TOMO_MAX_MATERIALS x 8 floats. */
#define TOMO_MAX_MATERIALS 128
int tomo_sim_materials(tomo_sim *sim, uint8_t *mat, float *coef, int *n_mat);
/* Running-DFT accumulators after folding: J and M as interleaved complex
 * [freq][point] (2*n_freq*P floats each), signs and dt applied as in
 * calculate_JandM (FDTD_GPU.cu:231-280). */
int tomo_sim_dft(tomo_sim *sim, float *J, float *M);
/* Far field D and phase P ([freq][angle]) of this simulation. */
int tomo_sim_farfield(tomo_sim *sim, float *D, float *P);

#ifdef __cplusplus
}
#endif
#endif /* TOMO_H */
