/*
 * cosenu_hip.h -- C ABI of the MI355X (gfx950) implementation of COSEnu's
 * time-step hot path.
 *
 * One context (cosenu_ctx) = one simulation state on one GPU, the role of the
 * reference's `class NuOsc` (src/nuosc.hpp:7-191).  Every entry point returns
 * 0 on success or a negative cosenu_status; the message of the last failure
 * on the calling thread is cosenu_last_error().  Nothing here calls exit().
 *
 * Arithmetic is IEEE FP64 throughout.  Two modes:
 *   COSENU_STRICT  the reference's operation order, true divisions, the
 *                  O(nvz) collective loop, no fused multiply-add: bit-identical
 *                  to the reference CPU build (rhs_fv.hpp / rhs_fd.hpp).
 *   COSENU_FAST    fused streaming stage kernels: one-sided WENO7 with a single
 *                  division, the collective integral from two velocity moments
 *                  per z, fused multiply-add.  Agrees with STRICT normwise.
 *
 * Host arrays in the reference's layout are accepted directly:
 *   planes  8 arrays (ee, xx, ex_re, ex_im, bee, bxx, bex_re, bex_im), each
 *           nvz rows; row i starts at i*row_stride and holds nz values from
 *           offset row_offset.  The reference's FieldVar is row_stride =
 *           nz+2*gz, row_offset = gz (structures.hpp:5-37, nuosc.hpp:145-148).
 *   interleaved  for i in nvz, for j in nz: 8 doubles in the field order above,
 *           the body of state.bin / state0.bin (snaps.hpp:20-35, 54-68).
 * The device layout is internal and never exposed.
 *
 * Threading: a context is used by one host thread at a time; distinct
 * contexts are independent.  cosenu_step() is asynchronous (it only enqueues
 * work on the context's stream); the calls that return data synchronise.
 */
#ifndef COSENU_HIP_H
#define COSENU_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define COSENU_ABI_VERSION 1

typedef enum {
    COSENU_OK = 0,
    COSENU_EINVAL = -1,   /* bad argument / unsupported combination */
    COSENU_ENOMEM = -2,   /* device or host allocation failed */
    COSENU_EHIP = -3,     /* HIP runtime error */
    COSENU_ESTATE = -4,   /* call out of order (e.g. no profile set) */
    COSENU_ENOTSUP = -5   /* feature not available in this mode */
} cosenu_status;

/* Scheme (reference: FV / FD preset, main.cpp:46-50). */
enum { COSENU_FV = 0, COSENU_FD = 1 };
/* Boundary (reference: PERIODIC_BC / OPEN_BC preset, nuosc.hpp:212-278). */
enum { COSENU_PERIODIC = 0, COSENU_OPEN = 1 };
/* Physics switches (reference presets VAC_OSC_ON, MAT_OSC_ON, COLL_OSC_ON,
 * ADVEC_OFF; COSENU_KO5 = FD without KO_ORD_3, i.e. 5th-order dissipation). */
enum { COSENU_VAC = 1, COSENU_MAT = 2, COSENU_COLL = 4, COSENU_ADVEC_OFF = 8, COSENU_KO5 = 16 };
enum { COSENU_STRICT = 0, COSENU_FAST = 1 };
/* Which state a transfer addresses. */
enum { COSENU_STATE = 0,   /* v_stat: the evolving state */
       COSENU_STATE0 = 1   /* v_stat0: the t=0 state used by the diagnostics */ };

/* Run parameters: the keys of the reference's job.config (parser.hpp:46-97)
 * plus the compile-time presets as runtime fields.  gz is accepted for
 * compatibility; the device path keeps no ghost cells.  ko is the
 * Kreiss-Oliger coefficient (nuosc.hpp:60, 0.1 in the reference). */
typedef struct cosenu_params {
    int scheme, bc, flags, nz, nvz, gz;
    double z0, z1, v0, v1, CFL;
    double pmo, omega, theta, mu, ko;
    int mode;   /* COSENU_STRICT or COSENU_FAST */
    int device; /* HIP device ordinal */
} cosenu_params;

/* Derived grid constants, as NuOsc's constructor computes them
 * (nuosc.hpp:78-96, 132-143). */
typedef struct cosenu_info {
    double dz, dv, dt, ct, st;
    double phy_time;
    long long steps_done;
    size_t device_bytes; /* device memory held by the context */
    int z_chunk;         /* FAST mode: z cells per streaming work item */
    int fast_variant;    /* FAST mode kernel: 5 = all four RK4 stages per
                            launch (default for FD with 3rd-order dissipation,
                            nvz <= 256, either boundary), 4 = two RK4 stages per
                            launch (default otherwise, either boundary),
                            3 = barrier-free (default for small periodic
                            problems), 2 = one stage per launch, two lanes
                            per bin (nu / nubar halves; nz < 16); 0 in STRICT.
                            nvz > 256: 5 = all four stages per launch with a
                            team of 2 blocks per z range (default for FD with
                            3rd-order dissipation on rows of 512 bins, i.e.
                            nvz 449-512, either boundary), 6 = two RK4 stages
                            per launch with a team of 2 or 4 blocks per z
                            range (default for other rows of 512 or 1024
                            bins, nvz 961-1024), else 3, in a general
                            form where a group of 32 bins straddles v = 0,
                            bins are padded (nvz not a multiple of 64) or the
                            boundary is open */
    int range_teams;     /* FAST mode: > 0 when each of this many teams streams
                            one contiguous z range (z_chunk = its length) */
} cosenu_info;

/* Diagnostics (analysis.hpp:3-185) plus conservation checks.  All sums are
 * over the interior cells times dz*dv like the reference's. */
typedef struct cosenu_diag {
    double dP_max;    /* max over cells of | |P|-|P0| | for nu and nubar (dcon) */
    double dP_avg;    /* G-weighted mean of | |P|-|P0| | (dcon) */
    double dbP_avg;   /* bG-weighted mean for nubar (dcon) */
    double M0_norm;   /* |M0| (cal_Mn, n = 0) */
    double M0[3];     /* M0 components e1, e2, e3 */
    double Pee;       /* <P_ee>  = sum ee / sum ee0 (surv_prob) */
    double Pbee;      /* <P_bee> = sum bee / sum bee0 */
    double trace;     /* sum (ee+xx) dz dv */
    double btrace;    /* sum (bee+bxx) dz dv */
    double lepton;    /* sum [(ee-xx)-(bee-bxx)] dz dv (= M0 e3) */
    double nonfinite; /* number of non-finite state values (0 when healthy) */
} cosenu_diag;

typedef struct cosenu_ctx cosenu_ctx;

/* Message of the last failed call on this thread ("" if none). */
const char *cosenu_last_error(void);
int cosenu_abi_version(void);

/* NuOsc::NuOsc + set_vac_pars + set_collective_pars (nuosc.hpp:47-143).
 * Allocates the device buffers; the state starts at zero. */
int cosenu_create(const cosenu_params *p, cosenu_ctx **out);
/* NuOsc::~NuOsc (nuosc.hpp:119-130). */
int cosenu_destroy(cosenu_ctx *ctx);
int cosenu_get_info(cosenu_ctx *ctx, cosenu_info *info);

/* The angular profiles G0->G, G0->bG (structures.hpp:124-137), set by
 * NuOsc::initialize (initialize.hpp:27-28) or NuOsc::read_G0
 * (snaps.hpp:161-183).  Planes of nvz rows as described above. */
int cosenu_set_profile(cosenu_ctx *ctx, const double *G, const double *bG, ptrdiff_t row_stride,
                       int row_offset);

/* State transfers.  `which` = COSENU_STATE or COSENU_STATE0.  Uploading
 * COSENU_STATE replaces the v_stat that initialize/read_state fill
 * (initialize.hpp:30-38, snaps.hpp:75-124); COSENU_STATE0 replaces the
 * v_stat0 copy (main.cpp:131-150). */
int cosenu_upload_planes(cosenu_ctx *ctx, int which, const double *const planes[8],
                         ptrdiff_t row_stride, int row_offset);
int cosenu_download_planes(cosenu_ctx *ctx, int which, double *const planes[8],
                           ptrdiff_t row_stride, int row_offset);
int cosenu_upload_interleaved(cosenu_ctx *ctx, int which, const double *buf);
int cosenu_download_interleaved(cosenu_ctx *ctx, int which, double *buf);

/* Initial states built on the device (the reference builds them on the host,
 * initialize.hpp:23-43).  Every initial state of the reference and of the
 * benchmark workloads is separable: field f at (z, v) is a[f*nvz + v] *
 * b[f*nz + z], one rounded product, exactly what the host computes (e.g.
 * ee = (0.5 G(v)) (1 + sqrt(1 - eps(z)^2))).  a is 8 x nvz, b is 8 x nz.
 * cosenu_set_profile_v takes z-independent profiles G(v), bG(v) (nvz each). */
int cosenu_init_separable(cosenu_ctx *ctx, int which, const double *a, const double *b);
int cosenu_set_profile_v(cosenu_ctx *ctx, const double *G, const double *bG);

/* NuOsc::copy_state (nuosc.hpp:193-210) between the two device states, e.g.
 * copy_state(v_stat, v_stat0) after initialize (main.cpp:139). */
int cosenu_copy_state(cosenu_ctx *ctx, int from, int to);

/* NuOsc::step_rk4 (nuosc.hpp:320-348) nsteps times.  Enqueues only.  Runs of
 * at least COSENU_GRAPH_STEPS steps (environment, default 16) replay a
 * captured HIP graph of that many steps; the rest are launched one by one. */
int cosenu_step(cosenu_ctx *ctx, int nsteps);
/* Capture now the graph that cosenu_step replays (no counterpart in the
 * reference): the first long cosenu_step then pays no capture and
 * instantiation.  Runs nothing on the device; a no-op while per-launch timing
 * is enabled or graphs are switched off (COSENU_GRAPH_STEPS=0). */
int cosenu_prepare(cosenu_ctx *ctx);
/* Block until all work of the context has finished. */
int cosenu_sync(cosenu_ctx *ctx);

/* out = calRHS(v_stat) (rhs_fv.hpp:115-236 / rhs_fd.hpp:6-148), written into
 * host planes.  Synchronous; for tests and for callers of calRHS. */
int cosenu_rhs(cosenu_ctx *ctx, double *const out[8], ptrdiff_t row_stride, int row_offset);

/* analyse + surv_prob (analysis.hpp:3-185) on the current state against
 * v_stat0 and the profiles, reduced on the device.  Synchronous. */
int cosenu_diagnostics(cosenu_ctx *ctx, cosenu_diag *d);
/* Angle-resolved survival probability: pee[i] = sum_z ee(z,v_i) /
 * sum_z ee0(z,v_i), pbee likewise for nubar.  Arrays of nvz.  Synchronous. */
int cosenu_survival_profile(cosenu_ctx *ctx, double *pee, double *pbee);

/* Snapshots without stopping the stepping: computes P_ee(z,v) = ee/ee0 into
 * a device buffer on the compute stream, then copies it on a second stream
 * into `host` (nvz rows of nz doubles, v-major like dump_rho's data).  `host`
 * must come from cosenu_host_alloc.  cosenu_snapshot_wait blocks until the
 * copy has landed. */
int cosenu_snapshot_pee_async(cosenu_ctx *ctx, double *host);
int cosenu_snapshot_wait(cosenu_ctx *ctx);
void *cosenu_host_alloc(size_t bytes);
void cosenu_host_free(void *p);

/* Per-kernel timing with HIP events recorded around every launch of a step
 * on the context's stream.  ms[k] = total milliseconds of launch k of the
 * step since timing was enabled, counts[k] = launches timed: launch k = RK
 * stage k+1 for the one-stage kernels; with two stages per launch (FAST
 * variant 4) launch 0 = stages 1+2, launch 1 = stages 3+4 and slots 2, 3
 * stay empty; with all four in one launch (variant 5) slot 0 only.
 * Enabling clears.  While enabled, steps are launched one by
 * one (no graph replay).  cosenu_stage_kernel_name(ctx, k) names the kernel
 * of launch k. */
int cosenu_set_timing(cosenu_ctx *ctx, int enable);
int cosenu_get_timing(cosenu_ctx *ctx, double ms[4], long long counts[4]);

/* Name of the HIP kernel that runs RK stage s (0..3) in this context, as
 * profilers report it. */
const char *cosenu_stage_kernel_name(cosenu_ctx *ctx, int stage);

/* ------------------------------------------------------------------------
 * Autotuner (no counterpart in the reference).  Times whole FAST RK4 steps
 * for each kernel variant and z-chunk length that divides nz on a
 * scratch context of p's shape, and records the fastest; contexts created
 * afterwards with the same (padded bins, scheme, nz) use it.  When the
 * environment variable COSENU_TUNE_FILE names a file, the table is read from
 * and written to it (text: VP scheme nz variant z_chunk ms). */
int cosenu_autotune(const cosenu_params *p, int *variant, int *z_chunk, double *ms);

/* ------------------------------------------------------------------------
 * Batched runs (no counterpart in the reference, whose sweeps are one process
 * per job, manage.py:361-428).  Many independent FAST-mode runs, of uneven nz,
 * nvz and physics, advance together: runs are grouped by (padded bins,
 * scheme), each group's states are packed into shared buffers, and one launch
 * per RK stage and group covers all (run, z-chunk) work items; runs of
 * cosenu_batch_step are replayed from captured HIP graphs.  Every run is
 * bit-identical to the same run in its own context.  Per-run calls mirror the
 * single-run ones. */
typedef struct cosenu_batch cosenu_batch;
typedef struct cosenu_batch_info {
    int nruns, ngroups;
    long long cells;      /* sum over runs of nz*nvz */
    long long steps_done;
    size_t device_bytes;
    long long work_items; /* (run, z-chunk) items per launch */
    int z_chunk;          /* chunk length of the first group */
    int fast_variant;     /* 4: two RK4 stages per launch (default), 2: one;
                             5: groups of FD runs with 3rd-order dissipation
                             take all four stages per launch (the others two) */
} cosenu_batch_info;

int cosenu_batch_create(const cosenu_params *runs, int nruns, cosenu_batch **out);
int cosenu_batch_destroy(cosenu_batch *b);
int cosenu_batch_get_info(cosenu_batch *b, cosenu_batch_info *info);

/* Tile scheduler tuning for batches (no counterpart in the reference, whose
 * sweeps are one process per job, manage.py:361-428).  Builds a scratch
 * batch of the given runs, times whole FAST steps for work-item chunks of
 * 32..512 z cells (at most a group's largest nz; a run's last item is
 * shorter) and records the fastest for every (padded bins, scheme)
 * group of the runs in the tune table (entry nz = -1; COSENU_TUNE_FILE as for
 * cosenu_autotune).  Batches created afterwards use it.  Needs the device
 * memory of the batch; *ms is per batch step. */
int cosenu_batch_autotune(const cosenu_params *runs, int nruns, int *z_chunk, double *ms);
int cosenu_batch_set_profile(cosenu_batch *b, int run, const double *G, const double *bG,
                             ptrdiff_t row_stride, int row_offset);
int cosenu_batch_upload_planes(cosenu_batch *b, int run, int which, const double *const planes[8],
                               ptrdiff_t row_stride, int row_offset);
int cosenu_batch_download_planes(cosenu_batch *b, int run, int which, double *const planes[8],
                                 ptrdiff_t row_stride, int row_offset);
int cosenu_batch_step(cosenu_batch *b, int nsteps);
int cosenu_batch_sync(cosenu_batch *b);
int cosenu_batch_diagnostics(cosenu_batch *b, int run, cosenu_diag *d);
int cosenu_batch_survival_profile(cosenu_batch *b, int run, double *pee, double *pbee);

/* ------------------------------------------------------------------------
 * Learned surrogate (no counterpart in the reference): a fully connected
 * network dims[0] -> ... -> dims[nlayers] with GELU between layers, BF16
 * weights and activations, FP32 accumulation on the matrix cores, FP32
 * master weights and AdamW state on the device.  Parameters are exchanged as
 * one flat FP32 array, layer by layer: W_l (dims[l] x dims[l+1], row-major,
 * y = x W + b) then b_l.  Inputs X are M x dims[0], outputs/targets
 * M x dims[nlayers], FP32 row-major host arrays.  Errors: negative status and
 * cosenu_mlp_last_error(). */
typedef struct cosenu_mlp cosenu_mlp;
const char *cosenu_mlp_last_error(void);
int cosenu_mlp_create(const int *dims, int nlayers, int max_batch, int device, cosenu_mlp **out);
int cosenu_mlp_destroy(cosenu_mlp *m);
long long cosenu_mlp_num_params(cosenu_mlp *m);
int cosenu_mlp_set_params(cosenu_mlp *m, const float *flat);
int cosenu_mlp_get_params(cosenu_mlp *m, float *flat);
int cosenu_mlp_forward(cosenu_mlp *m, const float *X, int M, float *Y);
/* Inference that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops).  X (M x dims[0]) and Y (M x dims[nlayers]) are
 * FP32 DEVICE arrays; the BF16 copy of X and the hidden activations live in
 * one 16 MiB array of the library's code object per GPU, so M x (dims[0] +
 * 2 x widest hidden layer) x 2 bytes (each part rounded up to 256 bytes)
 * must fit (about 14,000 rows of the 67-256-256-256-64 network), else
 * COSENU_EINVAL and a message.  Enqueues on `stream` (a hipStream_t; NULL =
 * the network's own stream, ordered after its training steps: then
 * cosenu_mlp_sync waits for it) and returns; Y equals cosenu_mlp_forward's
 * output bit for bit.  Every call on one GPU uses that GPU's one array:
 * calls that overlap in time on one GPU (other streams, other networks)
 * race. */
int cosenu_mlp_forward_noalloc(cosenu_mlp *m, const float *X, int M, float *Y, void *stream);
/* mean squared error over all M x dims[nlayers] outputs and its gradient */
int cosenu_mlp_gradients(cosenu_mlp *m, const float *X, const float *T, int M, float *grad,
                         float *loss);
/* one AdamW step (decoupled weight decay, none on biases) on that loss */
int cosenu_mlp_train_step(cosenu_mlp *m, const float *X, const float *T, int M, float lr,
                          float beta1, float beta2, float eps, float weight_decay, float *loss);
int cosenu_mlp_sync(cosenu_mlp *m);

#ifdef __cplusplus
}
#endif

#endif /* COSENU_HIP_H */
