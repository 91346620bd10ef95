// Shared definitions of the COSEnu MI355X kernels.
//
// Device state layout ("row layout"): for every z cell one row of 8 fields,
// each field a run of VP doubles over the velocity bins, VP = padded_bins(nvz)
// (64, 128 or 256 up to 256 bins, else nvz rounded up to a multiple of 64).
// Element (z, f, v) lives at
//     ((size_t)z * 8 + f) * VP + v.
// A 64-lane load of one field at one z is one contiguous 512-byte segment.
// Bins v >= nvz are padding, kept at zero and never written by the stage
// kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace cosenu {

enum Field { EE = 0, XX, EXR, EXI, BEE, BXX, BEXR, BEXI };

constexpr int kFlagVac = 1, kFlagMat = 2, kFlagColl = 4, kFlagAdvecOff = 8, kFlagKO5 = 16;

__host__ __device__ inline size_t row_index(int z, int f, int VP) { return ((size_t)z * 8 + f) * VP; }

// Bins per field row on the device.  Up to 256 bins the row is 64, 128 or
// 256 wide, the widths every FAST kernel is built for (a team holds VP/32
// wavefronts), so nvz = 129..256 takes 256 with zero padding bins; beyond
// 256 the row is nvz rounded up to a multiple of 64 (the barrier-free
// kernel, one wavefront per 64 bins).
inline int padded_bins(int nvz)
{
    if (nvz <= 64)
        return 64;
    if (nvz <= 128)
        return 128;
    if (nvz <= 256)
        return 256;
    return ((nvz + 63) / 64) * 64;
}

// Everything a stage kernel needs.  Stage s in 1..4 evaluates K = RHS(yin).
// Two forms of the RK4 combination (nuosc.hpp:320-348):
//  classic (STRICT), the reference's algebra:
//   s=1: acc = K;            yout = y + a*K        (a = 0.5 dt)
//   s=2: acc = acc + 2.0*K;  yout = y + a*K        (a = 0.5 dt)
//   s=3: acc = acc + 2.0*K;  yout = y + a*K        (a = dt)
//   s=4:                     yout = y + a*(K+acc)  (a = dt/6, yout == y)
//  stage-value form (FAST kernels): no accumulator; the slopes are
//  recovered from the stage states, K1 = 2(Y2-y)/dt, K2 = 2(Y3-y)/dt,
//  K3 = (Y4-y)/dt, so
//   s=1..3: yout = y + a*K   (Y2, Y3, Y4)
//   s=4:    yout = y + ((Y2-y) + 2(Y3-y) + (Y4-y))/3 + a*K   (a = dt/6)
//  which reads y, Y2, Y3 pointwise at stage 4 and moves 832 instead of
//  1088 bytes per cell and step.
struct StageArgs {
    const double *yin; // stage input Y_s (read with its z-neighbours)
    const double *y;   // state at the start of the step (pointwise)
    double *acc;       // classic form: RK accumulator
    double *yout;      // stage output
    const double *y2, *y3; // stage-value form, stage 4: Y2 and Y3
    const double *vz;  // [VP] velocity of each bin (padding bins: 0)
    const double *Hm;  // [nz] matter potential (MAT only)
    // barrier-free FAST kernel: per-z partial moments of the stage input
    // (read) and of the stage output (written), kMomSlot doubles per (z, bin
    // group of 32)
    const double *mom_in;
    double *mom_out;
    int nz, nvz, VP, flags, bc, scheme;
    double a;          // stage coefficient (see above)
    double dz, dv, ct, st, pmo, mu, ko;
};

// Partial moments of one bin group of 32 at one z: (S, vS, R, vR, D, vD) in
// the first 6 of kMomSlot doubles (S = ex_im + bex_im, R = ex_re - bex_re,
// D = ee - xx - bee + bxx summed over the group's bins, v = bin velocity).
constexpr int kMomSlot = 8;

// One run of a batched launch: its slices of the packed buffers and its
// physics constants (the runs of one batch may differ in every field).
struct BatchRun {
    double *y, *acc, *A, *B;
    const double *vz, *Hm;
    int nz, nvz, VP, flags, bc, scheme;
    double a[4]; // stage coefficients 0.5 dt, 0.5 dt, dt, dt/6
    double dz, dv, ct, st, pmo, mu, ko;
};

// The buffer roles of RK stage `stage` (1..4).  Classic form: y -> A -> B ->
// A -> y with the accumulator in `acc`.  Stage-value form: y -> A (Y2) -> B
// (Y3) -> acc (Y4) -> y, stage 4 reading A and B pointwise.
__host__ __device__ inline StageArgs batch_stage_args(const BatchRun &r, int stage, bool yform)
{
    StageArgs a;
    a.y = r.y;
    a.acc = r.acc;
    a.y2 = r.A;
    a.y3 = r.B;
    if (yform) {
        a.yin = stage == 1 ? r.y : (stage == 2 ? r.A : (stage == 3 ? r.B : r.acc));
        a.yout = stage == 1 ? r.A : (stage == 2 ? r.B : (stage == 3 ? r.acc : r.y));
    } else {
        a.yin = stage == 1 ? r.y : (stage == 3 ? r.B : r.A);
        a.yout = stage == 1 ? r.A : (stage == 2 ? r.B : (stage == 3 ? r.A : r.y));
    }
    a.a = r.a[stage - 1];
    a.vz = r.vz;
    a.Hm = r.Hm;
    a.mom_in = nullptr;
    a.mom_out = nullptr;
    a.nz = r.nz; a.nvz = r.nvz; a.VP = r.VP; a.flags = r.flags; a.bc = r.bc; a.scheme = r.scheme;
    a.dz = r.dz; a.dv = r.dv; a.ct = r.ct; a.st = r.st; a.pmo = r.pmo; a.mu = r.mu; a.ko = r.ko;
    return a;
}

struct RhsArgs { // plain RHS evaluation: out = RHS(in), for calRHS
    const double *in;
    double *out;
    const double *vz, *Hm;
    int nz, nvz, VP, flags, bc, scheme;
    double dz, dv, ct, st, pmo, mu, ko;
};

// Launchers (defined in the .hip files).
hipError_t launch_stage_strict(const StageArgs &a, int stage, hipStream_t s);
hipError_t launch_rhs_strict(const RhsArgs &a, hipStream_t s);
const char *stage_strict_kernel_name(int stage);

// FAST mode, one RK4 stage per launch (variant 2, stage_fast.hip): two lanes
// per velocity bin (nu and nubar halves).  Returns hipErrorInvalidValue for
// unsupported shapes.  nteams > 0: nteams teams, each streaming one
// contiguous z range; otherwise chunks of z_chunk cells.
hipError_t launch_stage_fast(const StageArgs &a, int stage, int z_chunk, int nteams, hipStream_t s);
hipError_t launch_rhs_fast(const RhsArgs &a, int z_chunk, int nteams, hipStream_t s);
// teams of the split-lane kernel resident on the current device at once
int fast_resident_teams(int VP, int scheme, int bc);

// Barrier-free FAST kernel (variant 3, stage_stream.hip): every wavefront
// marches its own 32 bins and takes the per-z moments of its input from
// a.mom_in, written by the launch that produced the input.  The uniform form
// (general = false) needs nvz a multiple of 64, periodic boundaries and bins
// of one sign of v in every group of 32 (stream_supported) and marches each
// group in its upwind direction; the general form takes any nvz (rows of VP,
// a multiple of 64), groups that straddle v = 0 and either boundary.
int stream_supported(int nvz, int VP, int bc, const double *vz_host);
hipError_t launch_stage_stream(const StageArgs &a, int stage, int z_chunk, int nteams,
                               bool general, hipStream_t s);
// partial moments of state y into mom (kMomSlot doubles per z and bin group;
// bins v >= nvz count zero)
hipError_t launch_stream_moments(const double *y, const double *vz, double *mom, int nz, int nvz,
                                 int VP, hipStream_t s);
// waves of the barrier-free kernel resident on the current device at once
int stream_resident_waves(int scheme, bool general, bool stage4 = false);
const char *stage_stream_kernel_name(int stage, bool general);

// Two RK4 stages per launch (variant 4, stage_pair.hip): pair 0 runs stages
// 1 and 2 (a1, a2 = their stage arguments), pair 1 stages 3 and 4; the
// second stage's input never goes through HBM.  a2.yout of pair 1 must not
// alias a1.y (the new state goes to a second buffer).  Periodic boundaries,
// VP in {64, 128, 256}.
int pair_supported(int VP, int bc);
// uniform: nvz == VP and every group of 32 bins has one sign of v
// (stream_supported's test) -- selects kernels without padding masks.
hipError_t launch_stage_pair(const StageArgs &a1, const StageArgs &a2, int pair, int z_chunk,
                             int nteams, bool uniform, hipStream_t s);
int pair_resident_teams(int VP, int scheme);
// Batched two-stage launch over work items (run, first z cell) of runs with
// the same VP and scheme; each run's .y is its current state and .acc the
// buffer that receives the new one (pair 1).
// std_phys: every run of the launch has the benchmark physics (vacuum +
// collective, no matter term, 3rd-order dissipation): the instances with
// those switches fixed at compile time
hipError_t launch_stage_pair_batch(const BatchRun *runs, const int2 *items, int nitems, int VP,
                                   int scheme, int L, int pair, bool uniform, bool std_phys,
                                   hipStream_t s);
const char *stage_pair_kernel_name(int pair);
// Two RK4 stages per launch for rows of 512 or 1024 bins (variant 6,
// stage_pair.hip): a team of VP/256 blocks per z range, the per-row moments
// summed over the blocks through `slots` (pair_wide_slot_words(VP, nteams)
// words, zero before the first launch of a team layout) with `ctl` (2
// words, zero before the first launch of a layout: the ticket counter and
// the error word -- error != 0 after a block gave up waiting for a partner,
// which voids the launch's results).  Range mode only: nteams contiguous z
// ranges.
int pair_wide_supported(int VP, int bc);
size_t pair_wide_slot_words(int VP, int nteams);
int pair_wide_resident_blocks(int scheme);
// give_up_test: every wait of the launch looks for a tag no block writes and
// gives up after 10 us (a test of the error path; the launch's results are
// void)
hipError_t launch_stage_pair_wide(const StageArgs &a1, const StageArgs &a2, int pair, int nteams,
                                  bool uniform, unsigned long long *slots,
                                  unsigned long long *ctl, bool give_up_test, hipStream_t s);
const char *stage_pair_wide_kernel_name(int pair);

// FAST mode, FD with 3rd-order dissipation: all four RK4 stages in one
// launch (variant 5, stage_quad.hip).  `a` = stage 1's arguments (yin = y);
// ca = dt/2, dt/2, dt, dt/6; the new state goes to yout.
int quad_supported(int VP, int bc, int scheme, int flags);
hipError_t launch_stage_quad(const StageArgs &a, const double ca[4], double *yout, int z_chunk,
                             int nteams, bool uniform, hipStream_t s);
int quad_resident_teams(int VP);
const char *stage_quad_kernel_name();
// Batched runs (every run FD with 3rd-order dissipation, periodic, nz >= 16):
// one launch per step; the new states go to the runs' acc buffers.  mat: some
// run of the launch has the matter term.
hipError_t launch_stage_quad_batch(const BatchRun *runs, const int2 *items, int nitems, int VP,
                                   int L, bool uniform, bool mat, bool std_phys, hipStream_t s);
// Rows of 512 bins (variant 5 on wide grids): a team of 2 blocks per z
// range, the moments exchanged as in launch_stage_pair_wide
// through `slots` (quad_wide_slot_words(VP, nteams) words) and `ctl`.  Range
// mode only.
int quad_wide_supported(int VP, int bc, int scheme, int flags);
size_t quad_wide_slot_words(int VP, int nteams);
int quad_wide_resident_blocks();
hipError_t launch_stage_quad_wide(const StageArgs &a, const double ca[4], double *yout,
                                  int nteams, bool uniform, unsigned long long *slots,
                                  unsigned long long *ctl, bool give_up_test, hipStream_t s);
const char *stage_quad_wide_kernel_name();
const char *stage_fast_kernel_name(int VP, int stage);
// Batched FAST stage over `nitems` work items (run index, first z cell) of
// runs with the same VP and scheme; `runs` and `items` are device arrays.
hipError_t launch_stage_fast_batch(const BatchRun *runs, const int2 *items, int nitems, int VP,
                                   int scheme, int L, int stage, hipStream_t s);
int fast_supported(int VP, int bc);

// Layout conversion and reductions (aux.hip).
hipError_t launch_import_plane(double *dst, const double *src, int f, int nz, int nvz, int VP,
                               ptrdiff_t row_stride, int row_offset, hipStream_t s);
hipError_t launch_export_plane(double *dst, const double *src, int f, int nz, int nvz, int VP,
                               ptrdiff_t row_stride, int row_offset, hipStream_t s);
hipError_t launch_import_interleaved(double *dst, const double *src, int nz, int nvz, int VP,
                                     hipStream_t s);
hipError_t launch_export_interleaved(double *dst, const double *src, int nz, int nvz, int VP,
                                     hipStream_t s);
// G/bG planes -> profile rows [z][2][VP]
hipError_t launch_import_profile(double *dst, const double *G, const double *bG, int nz, int nvz,
                                 int VP, ptrdiff_t row_stride, int row_offset, hipStream_t s);

constexpr int kDiagSlots = 20;
// Partial sums of the diagnostics over z-blocks, then a fixed-order final
// reduction: out[kDiagSlots] on the device.
hipError_t launch_diagnostics(const double *y, const double *y0, const double *prof, int nz,
                              int nvz, int VP, double *partials, int npart, double *out,
                              hipStream_t s);
int diag_partials_needed(int nz);
hipError_t launch_survival_profile(const double *y, const double *y0, int nz, int nvz, int VP,
                                   double *partials, int npart, double *out, hipStream_t s);
hipError_t launch_pee_plane(double *dst, const double *y, const double *y0, int nz, int nvz,
                            int VP, hipStream_t s);
// state(z, f, v) = a[f*VP + v] * b[f*nz + z]; a, b on the device
hipError_t launch_separable(double *dst, const double *a, const double *b, int nz, int nvz, int VP,
                            hipStream_t s);
hipError_t launch_profile_v(double *dst, const double *G, const double *bG, int nz, int nvz, int VP,
                            hipStream_t s);
// nrows rows of VP doubles: bins v < nvz pseudo-random in [0.5, 1), padding 0
hipError_t launch_fill_pattern(double *dst, long nrows, int nvz, int VP, hipStream_t s);

} // namespace cosenu
