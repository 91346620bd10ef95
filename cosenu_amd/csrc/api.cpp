// C ABI of the COSEnu MI355X library (include/cosenu_hip.h): context and
// batch lifetime, layout conversion on upload/download, the RK4 drivers,
// timing, diagnostics, snapshots and the z-chunk autotuner.  Host code only;
// kernels live in the .hip files.
#include "cosenu_hip.h"
#include "common.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <numeric>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

using namespace cosenu;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(COSENU_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));          \
    } while (0)

#define RC_TRY(expr)                                                                               \
    do {                                                                                           \
        int r_ = (expr);                                                                           \
        if (r_)                                                                                    \
            return r_;                                                                             \
    } while (0)

int env_int(const char *name, int dflt)
{
    const char *s = std::getenv(name);
    return (s && *s) ? std::atoi(s) : dflt;
}

// ---------------------------------------------------------------------------
// One state in device row layout plus what its transfers and reductions need.
struct View {
    double *y = nullptr, *y0 = nullptr, *prof = nullptr; // state, t=0 state, profiles
    double *sa = nullptr, *sb = nullptr;                  // staging (stage buffers)
    int nz = 0, nvz = 0, VP = 0;
    size_t n_state = 0; // doubles in one state slice
};

double *which_ptr(const View &v, int which) { return which == COSENU_STATE0 ? v.y0 : v.y; }

bool valid_layout(const View &v, ptrdiff_t stride, int off)
{
    return off >= 0 && stride >= (ptrdiff_t)off + v.nz && (size_t)stride * v.nvz <= v.n_state;
}

int view_upload_planes(const View &v, int which, const double *const planes[8], ptrdiff_t stride,
                       int off, hipStream_t s)
{
    if (!planes || !valid_layout(v, stride, off) || (which != COSENU_STATE && which != COSENU_STATE0))
        return fail(COSENU_EINVAL, "upload_planes: bad argument");
    const size_t w = (size_t)v.nz * sizeof(double), pitch = (size_t)stride * sizeof(double);
    for (int f = 0; f < 8; f++) {
        if (!planes[f])
            return fail(COSENU_EINVAL, "upload_planes: null plane");
        double *stage = (f & 1) ? v.sb : v.sa;
        HIP_TRY(hipMemcpy2DAsync(stage + off, pitch, planes[f] + off, pitch, w, v.nvz,
                                 hipMemcpyHostToDevice, s));
        HIP_TRY(launch_import_plane(which_ptr(v, which), stage, f, v.nz, v.nvz, v.VP, stride, off, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

int view_download_planes(const View &v, const double *src, double *const planes[8],
                         ptrdiff_t stride, int off, hipStream_t s)
{
    if (!planes || !valid_layout(v, stride, off))
        return fail(COSENU_EINVAL, "download_planes: bad argument");
    const size_t w = (size_t)v.nz * sizeof(double), pitch = (size_t)stride * sizeof(double);
    for (int f = 0; f < 8; f++) {
        if (!planes[f])
            return fail(COSENU_EINVAL, "download_planes: null plane");
        HIP_TRY(launch_export_plane(v.sa, src, f, v.nz, v.nvz, v.VP, stride, off, s));
        HIP_TRY(hipMemcpy2DAsync(planes[f] + off, pitch, v.sa + off, pitch, w, v.nvz,
                                 hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
    }
    return 0;
}

int view_upload_interleaved(const View &v, int which, const double *buf, hipStream_t s)
{
    if (!buf || (which != COSENU_STATE && which != COSENU_STATE0))
        return fail(COSENU_EINVAL, "upload_interleaved: bad argument");
    const size_t n = (size_t)v.nvz * v.nz * 8;
    HIP_TRY(hipMemcpyAsync(v.sa, buf, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(launch_import_interleaved(which_ptr(v, which), v.sa, v.nz, v.nvz, v.VP, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

int view_download_interleaved(const View &v, int which, double *buf, hipStream_t s)
{
    if (!buf || (which != COSENU_STATE && which != COSENU_STATE0))
        return fail(COSENU_EINVAL, "download_interleaved: bad argument");
    const size_t n = (size_t)v.nvz * v.nz * 8;
    HIP_TRY(launch_export_interleaved(v.sa, which_ptr(v, which), v.nz, v.nvz, v.VP, s));
    HIP_TRY(hipMemcpyAsync(buf, v.sa, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

int view_set_profile(const View &v, const double *G, const double *bG, ptrdiff_t stride, int off,
                     hipStream_t s)
{
    if (!G || !bG || !valid_layout(v, stride, off))
        return fail(COSENU_EINVAL, "set_profile: bad argument");
    const size_t w = (size_t)v.nz * sizeof(double), pitch = (size_t)stride * sizeof(double);
    HIP_TRY(hipMemcpy2DAsync(v.sa + off, pitch, G + off, pitch, w, v.nvz, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpy2DAsync(v.sb + off, pitch, bG + off, pitch, w, v.nvz, hipMemcpyHostToDevice, s));
    HIP_TRY(launch_import_profile(v.prof, v.sa, v.sb, v.nz, v.nvz, v.VP, stride, off, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

int view_diagnostics(const View &v, double dz, double dv, double *part, int npart, double *dout,
                     cosenu_diag *d, hipStream_t s)
{
    HIP_TRY(launch_diagnostics(v.y, v.y0, v.prof, v.nz, v.nvz, v.VP, part, npart, dout, s));
    double h[kDiagSlots];
    HIP_TRY(hipMemcpyAsync(h, dout, sizeof(h), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const double dzdv = dz * dv;
    d->dP_max = h[0] >= h[1] ? h[0] : h[1];
    if (h[0] != h[0] || h[1] != h[1])
        d->dP_max = NAN;
    d->dP_avg = h[2] / h[4];
    d->dbP_avg = h[3] / h[5];
    d->M0[0] = h[6] * dzdv;
    d->M0[1] = h[7] * dzdv;
    d->M0[2] = h[8] * dzdv;
    d->M0_norm = sqrt(d->M0[0] * d->M0[0] + d->M0[1] * d->M0[1] + d->M0[2] * d->M0[2]);
    d->Pee = h[9] / h[11];
    d->Pbee = h[10] / h[12];
    d->trace = h[13] * dzdv;
    d->btrace = h[14] * dzdv;
    d->lepton = h[15] * dzdv;
    d->nonfinite = h[16];
    return 0;
}

int view_survival(const View &v, double *part, int npart, double *dout, double *pee, double *pbee,
                  hipStream_t s)
{
    HIP_TRY(launch_survival_profile(v.y, v.y0, v.nz, v.nvz, v.VP, part, npart, dout, s));
    HIP_TRY(hipMemcpyAsync(pee, dout, v.nvz * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(pbee, dout + v.VP, v.nvz * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

// Physics constants derived like NuOsc's constructor (nuosc.hpp:78-96) and
// set_vac_pars (nuosc.hpp:132-138).
struct Derived {
    double dz, dv, dt, ct, st;
};

Derived derive(const cosenu_params &p)
{
    Derived d;
    d.dz = (p.z1 - p.z0) / p.nz;
    d.dv = (p.v1 - p.v0) / (p.nvz);
    d.dt = p.CFL * d.dz;
    double th = p.theta * M_PI / 180.0;
    d.ct = (1.0 / 2.0) * p.omega * cos(2.0 * th);
    d.st = (1.0 / 2.0) * p.omega * sin(2.0 * th);
    return d;
}

int validate(const cosenu_params *p)
{
    if (p->scheme != COSENU_FV && p->scheme != COSENU_FD)
        return fail(COSENU_EINVAL, "scheme must be COSENU_FV or COSENU_FD");
    if (p->bc != COSENU_PERIODIC && p->bc != COSENU_OPEN)
        return fail(COSENU_EINVAL, "bc must be COSENU_PERIODIC or COSENU_OPEN");
    if (p->nvz < 1 || p->nvz > 1024)
        return fail(COSENU_EINVAL, "nvz must be in 1..1024");
    if (p->nz < 8)
        return fail(COSENU_EINVAL, "nz must be at least 8");
    if (p->flags & ~(COSENU_VAC | COSENU_MAT | COSENU_COLL | COSENU_ADVEC_OFF | COSENU_KO5))
        return fail(COSENU_EINVAL, "unknown flag bits");
    if ((p->flags & COSENU_KO5) && p->scheme != COSENU_FD)
        return fail(COSENU_EINVAL, "COSENU_KO5 applies to the FD scheme only");
    if (p->mode != COSENU_STRICT && p->mode != COSENU_FAST)
        return fail(COSENU_EINVAL, "mode must be COSENU_STRICT or COSENU_FAST");
    if (!(p->z1 > p->z0) || !(p->v1 > p->v0) || !(p->CFL > 0))
        return fail(COSENU_EINVAL, "need z1 > z0, v1 > v0, CFL > 0");
    // FAST mode takes every grid: rows of up to 256 bins on the team kernels
    // (either boundary), wider rows on the barrier-free kernel (its general
    // form for groups of 32 bins that straddle v = 0, padding bins and the
    // open boundary)
    return 0;
}

void fill_grids(const cosenu_params &p, const Derived &d, int VP, std::vector<double> &vz,
                std::vector<double> &Hm)
{
    vz.assign(VP, 0.0);
    Hm.assign(p.nz, 0.0);
    for (int i = 0; i < p.nvz; i++)
        vz[i] = p.v0 + (i + 0.5) * d.dv;
    for (int j = 0; j < p.nz; j++)
        Hm[j] = -1.0e-3 * ((p.z0 + (j + 0.5) * d.dz) + 600.0);
}

// ---------------------------------------------------------------------------
// Autotune table: (VP, scheme, nz) -> (variant, z chunk).  Loaded from and
// saved to the file named by COSENU_TUNE_FILE (text, one entry per line).
struct TuneKey {
    int VP, scheme, nz;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(VP, scheme, nz) < std::tie(o.VP, o.scheme, o.nz);
    }
};
struct TuneVal {
    int variant, chunk;
    double ms;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;
bool g_tune_loaded = false;

void tune_load_locked()
{
    if (g_tune_loaded)
        return;
    g_tune_loaded = true;
    const char *f = std::getenv("COSENU_TUNE_FILE");
    if (!f || !*f)
        return;
    std::ifstream in(f);
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#')
            continue;
        std::istringstream ss(line);
        TuneKey k;
        TuneVal v;
        // entries of kernels this build does not have are skipped
        if (ss >> k.VP >> k.scheme >> k.nz >> v.variant >> v.chunk >> v.ms && v.variant >= 2 &&
            v.variant <= 6)
            g_tune[k] = v;
    }
}

bool tune_lookup(int VP, int scheme, int nz, TuneVal *out)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load_locked();
    auto it = g_tune.find({VP, scheme, nz});
    if (it == g_tune.end())
        return false;
    *out = it->second;
    return true;
}

void tune_store(int VP, int scheme, int nz, TuneVal v)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load_locked();
    g_tune[{VP, scheme, nz}] = v;
    const char *f = std::getenv("COSENU_TUNE_FILE");
    if (!f || !*f)
        return;
    std::ofstream out(f, std::ios::trunc);
    out << "# VP scheme nz variant z_chunk ms_per_step\n";
    for (auto &kv : g_tune)
        out << kv.first.VP << " " << kv.first.scheme << " " << kv.first.nz << " "
            << kv.second.variant << " " << kv.second.chunk << " " << kv.second.ms << "\n";
}

// Contiguous-range mode: `mult` times the resident teams (one team = the
// wavefronts covering all bins of a z range), at most one team per 2 z cells
// (8 for the barrier-free kernel).
//
// Two-stage launches of small problems: a team's launch time is about
// (rows + 9) z steps (the first stage's halo), each step slower with two
// wavefronts per SIMD than with one (x1.44 FV, x1.23 FD measured at nz =
// 1024, nvz = 64).  Where the resident teams would get few rows each, half
// as many teams (one wavefront per SIMD) finish sooner: below 8 rows per
// resident team for FV, 3 for FD.  (nz = 1024, nvz = 64: 512 teams of 2
// rows, 63 us per FV step against 90 us with 128 teams of 8 rows.)
// (stage4: the barrier-free kernel's stage-4 launch, which holds fewer
// wavefronts per SIMD than stages 1-3 and gets ranges of its own)
int range_teams(int nz, int VP, int scheme, int bc, int variant, int mult, bool general,
                bool stage4 = false)
{
    long t = variant == 3   ? (long)stream_resident_waves(scheme, general, stage4) / (VP / 32)
             : variant == 4 ? (long)pair_resident_teams(VP, scheme)
             : variant == 5 ? (long)(VP > 256 ? quad_wide_resident_blocks() / (VP / 256)
                                              : quad_resident_teams(VP))
             : variant == 6 ? (long)pair_wide_resident_blocks(scheme) / (VP / 256)
                            : (long)fast_resident_teams(VP, scheme, bc);
    t *= mult;
    if (variant == 4 && mult == 1 && t >= 2 && (nz + t - 1) / t < (scheme == 0 ? 8 : 3))
        t /= 2;
    t = std::min<long>(t, variant == 3 ? nz / 8 : nz / 2);
    return (int)std::max<long>(t, 1);
}

// Default z chunk: ~4 wavefronts per SIMD over 1024 SIMDs, within [8, 256],
// rounded down to a divisor of nz.
int default_chunk(int nz, int VP)
{
    long target = (long)nz * (VP / 64) / 4096;
    int L = (int)std::max(8L, std::min(256L, target));
    if (L > nz)
        L = nz;
    while (L > 1 && nz % L)
        L--;
    return L;
}

} // namespace

// ===========================================================================
// Single-run context
struct cosenu_ctx {
    cosenu_params p;
    Derived d;
    int nz = 0, nvz = 0, VP = 0;
    double phy_time = 0;
    long long steps = 0;
    size_t bytes = 0;
    hipStream_t stream = nullptr, copy_stream = nullptr;
    double *y = nullptr, *acc = nullptr, *A = nullptr, *B = nullptr, *y0 = nullptr;
    double *states = nullptr; // one allocation holding y, acc, A, B, y0
    double *prof = nullptr, *vz = nullptr, *Hm = nullptr;
    double *diag_part = nullptr, *diag_out = nullptr, *snap = nullptr;
    int npart = 0;
    bool have_profile = false, have_state0 = false;
    int z_chunk = 0, variant = 2;
    int nteams = 0; // > 0: contiguous-range mode with this many teams
    int nteams4 = 0; // variant 3 in range mode: the stage-4 launch's team count
    bool stream_ok = false; // the barrier-free kernel (variant 3) applies
    bool stream_general = false; // ... in its general form (nvz > 256 only)
    bool uniform = false;   // nvz == VP, every group of 32 bins of one sign of v
    double *mom[4] = {nullptr, nullptr, nullptr, nullptr}; // its moments of y, A, B, acc
    // variants 5 and 6 on rows of more than 256 bins: the device-memory
    // moment rings of their multi-block teams (xwords words) and the 2
    // control words (ticket counter, error)
    unsigned long long *xslots = nullptr;
    unsigned long long *xctl = nullptr;
    size_t xwords = 0;
    bool wide_give_up_test = false; // COSENU_WIDE_GIVE_UP_TEST: force the error path
    View view;
    bool timing = false;
    std::vector<hipEvent_t> ev;
    std::vector<int> ev_stage;
    double ms[4] = {0, 0, 0, 0};
    long long cnt[4] = {0, 0, 0, 0};
    // captured step graphs, one per state buffer they start from (the two-
    // and four-stage launches alternate two buffers): graph[k] starts from
    // graph_y[k] and runs graph_steps[k] steps
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_steps[2] = {0, 0};
    const double *graph_y[2] = {nullptr, nullptr};
    hipEvent_t snap_ready = nullptr, snap_done = nullptr;
    bool snap_pending = false;
};

namespace {

template <class T>
int dev_alloc(size_t &bytes, T **p, size_t n)
{
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(COSENU_ENOMEM, "hipMalloc of " + std::to_string(n * sizeof(T)) +
                                       " bytes failed: " + hipGetErrorString(e));
    }
    bytes += n * sizeof(T);
    return 0;
}

void release(cosenu_ctx *c)
{
    double *bufs[] = {c->states, c->prof, c->vz, c->Hm, c->diag_part, c->diag_out, c->snap};
    for (double *b : bufs)
        if (b)
            (void)hipFree(b);
    for (double *b : c->mom)
        if (b)
            (void)hipFree(b);
    if (c->xslots)
        (void)hipFree(c->xslots);
    if (c->xctl)
        (void)hipFree(c->xctl);
    for (hipEvent_t e : c->ev)
        (void)hipEventDestroy(e);
    for (hipGraphExec_t g : c->graph)
        if (g)
            (void)hipGraphExecDestroy(g);
    if (c->snap_ready)
        (void)hipEventDestroy(c->snap_ready);
    if (c->snap_done)
        (void)hipEventDestroy(c->snap_done);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    if (c->copy_stream)
        (void)hipStreamDestroy(c->copy_stream);
}

StageArgs stage_args(const cosenu_ctx *c, int stage)
{
    BatchRun r;
    r.y = c->y; r.acc = c->acc; r.A = c->A; r.B = c->B;
    r.vz = c->vz; r.Hm = c->Hm;
    r.nz = c->nz; r.nvz = c->nvz; r.VP = c->VP;
    r.flags = c->p.flags; r.bc = c->p.bc; r.scheme = c->p.scheme;
    r.a[0] = 0.5 * c->d.dt; r.a[1] = 0.5 * c->d.dt; r.a[2] = c->d.dt; r.a[3] = 1.0 / 6.0 * c->d.dt;
    r.dz = c->d.dz; r.dv = c->d.dv; r.ct = c->d.ct; r.st = c->d.st;
    r.pmo = c->p.pmo; r.mu = c->p.mu; r.ko = c->p.ko;
    // the FAST kernels use the stage-value form
    StageArgs a = batch_stage_args(r, stage, c->p.mode == COSENU_FAST);
    if (c->p.mode == COSENU_FAST && c->variant == 3) {
        // moments travel with the stage states: y, A (Y2), B (Y3), acc (Y4)
        a.mom_in = c->mom[stage - 1];
        a.mom_out = c->mom[stage & 3];
    }
    return a;
}

RhsArgs rhs_args(const cosenu_ctx *c, const double *in, double *out)
{
    RhsArgs r;
    r.in = in; r.out = out; r.vz = c->vz; r.Hm = c->Hm;
    r.nz = c->nz; r.nvz = c->nvz; r.VP = c->VP;
    r.flags = c->p.flags; r.bc = c->p.bc; r.scheme = c->p.scheme;
    r.dz = c->d.dz; r.dv = c->d.dv; r.ct = c->d.ct; r.st = c->d.st;
    r.pmo = c->p.pmo; r.mu = c->p.mu; r.ko = c->p.ko;
    return r;
}

hipError_t launch_stage(const cosenu_ctx *c, int stage, hipStream_t s)
{
    StageArgs a = stage_args(c, stage);
    if (c->p.mode == COSENU_FAST && c->variant == 3)
        return launch_stage_stream(a, stage, c->z_chunk,
                                   stage == 4 && c->nteams4 > 0 ? c->nteams4 : c->nteams,
                                   c->stream_general, s);
    if (c->p.mode == COSENU_FAST)
        return launch_stage_fast(a, stage, c->z_chunk, c->nteams, s);
    return launch_stage_strict(a, stage, s);
}

bool two_stage_launches(const cosenu_ctx *c)
{
    return c->p.mode == COSENU_FAST && (c->variant == 4 || c->variant == 6);
}
bool four_stage_launch(const cosenu_ctx *c) { return c->p.mode == COSENU_FAST && c->variant == 5; }
// the launches whose teams span several blocks (rows of more than 256 bins)
bool wide_teams(const cosenu_ctx *c)
{
    return c->p.mode == COSENU_FAST && c->VP > 256 && (c->variant == 5 || c->variant == 6);
}
// the words of the moment rings of such a launch
size_t wide_words(const cosenu_ctx *c)
{
    return c->variant == 5 ? quad_wide_slot_words(c->VP, c->nteams)
                           : pair_wide_slot_words(c->VP, c->nteams);
}
// the kernels that write the new state to the spare buffer
bool swaps_state(const cosenu_ctx *c) { return two_stage_launches(c) || four_stage_launch(c); }

// Stages 1+2 (pair 0) or 3+4 (pair 1) in one launch; the new state of pair 1
// goes to the spare state buffer (acc), which becomes y after the step.
hipError_t launch_pair(const cosenu_ctx *c, int pair, hipStream_t s)
{
    StageArgs a1 = stage_args(c, pair == 0 ? 1 : 3), a2 = stage_args(c, pair == 0 ? 2 : 4);
    if (pair == 1)
        a2.yout = c->acc;
    if (c->variant == 6)
        return wide_words(c) <= c->xwords
                   ? launch_stage_pair_wide(a1, a2, pair, c->nteams, c->uniform, c->xslots,
                                            c->xctl, c->wide_give_up_test, s)
                   : hipErrorInvalidValue;
    return launch_stage_pair(a1, a2, pair, c->z_chunk, c->nteams, c->uniform, s);
}

// Multi-block teams (wide_teams): moment rings for the current layout
// (allocated or grown here, never inside a graph capture), zeroed, and
// zeroed control words.  The rings need no clearing between launches of one
// layout (their words are tagged with launch and row); a layout starts from
// zeroed rings and a zeroed ticket counter (the kernels derive team, block
// and launch number from it), as a reset counter numbers launches from 0
// again.
int wide_buffers(cosenu_ctx *c)
{
    const size_t words = wide_words(c);
    if (words > c->xwords || !c->xslots) {
        if (c->xslots)
            (void)hipFree(c->xslots);
        c->xslots = nullptr;
        c->xwords = 0;
        RC_TRY(dev_alloc(c->bytes, &c->xslots, words));
        c->xwords = words;
    }
    HIP_TRY(hipMemsetAsync(c->xslots, 0, words * sizeof(unsigned long long), c->stream));
    if (!c->xctl)
        RC_TRY(dev_alloc(c->bytes, &c->xctl, 2));
    HIP_TRY(hipMemsetAsync(c->xctl, 0, 2 * sizeof(unsigned long long), c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// Multi-block teams: an error if a block of a launch since the last check
// gave up waiting for its partners' moments (the launch's results are void).
int wide_check(cosenu_ctx *c)
{
    if (!c->xctl)
        return 0;
    unsigned long long err = 0;
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(&err, c->xctl + 1, sizeof(err), hipMemcpyDeviceToHost));
    if (err) {
        const unsigned long long zero = 0;
        HIP_TRY(hipMemcpy(c->xctl + 1, &zero, sizeof(zero), hipMemcpyHostToDevice));
        return fail(COSENU_EHIP, "launch of a wide grid: a block waited more than 1 s for the "
                                 "moments of its team's other blocks; the state is void");
    }
    return 0;
}

void swap_state(cosenu_ctx *c)
{
    std::swap(c->y, c->acc);
    c->view.y = c->y;
}

// All four stages in one launch; the new state goes to acc.
hipError_t launch_quad(const cosenu_ctx *c, hipStream_t s)
{
    const StageArgs a = stage_args(c, 1);
    const double dt = c->d.dt;
    const double ca[4] = {0.5 * dt, 0.5 * dt, dt, 1.0 / 6.0 * dt};
    if (c->VP > 256)
        return wide_words(c) <= c->xwords
                   ? launch_stage_quad_wide(a, ca, c->acc, c->nteams, c->uniform, c->xslots,
                                            c->xctl, c->wide_give_up_test, s)
                   : hipErrorInvalidValue;
    return launch_stage_quad(a, ca, c->acc, c->z_chunk, c->nteams, c->uniform, s);
}

// The launches of one RK4 step; with timing, events around each launch
// (timing slot k = launch k of the step).
int enqueue_step(cosenu_ctx *c, hipStream_t s)
{
    const bool pairs = two_stage_launches(c), quad = four_stage_launch(c);
    for (int k = 0; k < (quad ? 1 : pairs ? 2 : 4); k++) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (c->timing) {
            HIP_TRY(hipEventCreate(&e0));
            HIP_TRY(hipEventCreate(&e1));
            HIP_TRY(hipEventRecord(e0, s));
        }
        HIP_TRY(quad ? launch_quad(c, s) : pairs ? launch_pair(c, k, s) : launch_stage(c, k + 1, s));
        if (c->timing) {
            HIP_TRY(hipEventRecord(e1, s));
            c->ev.push_back(e0);
            c->ev.push_back(e1);
            c->ev_stage.push_back(k);
        }
    }
    if (pairs || quad)
        swap_state(c);
    return 0;
}

int fold_timing(cosenu_ctx *c)
{
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < c->ev_stage.size(); k++) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev[2 * k], c->ev[2 * k + 1]));
        c->ms[c->ev_stage[k]] += ms;
        c->cnt[c->ev_stage[k]] += 1;
    }
    for (hipEvent_t e : c->ev)
        (void)hipEventDestroy(e);
    c->ev.clear();
    c->ev_stage.clear();
    return 0;
}

// Capture `nsteps` RK4 steps (4 launches each) on `s` into an executable graph.
template <class Launch>
int capture_graph(hipStream_t s, int nsteps, Launch launch_step, hipGraphExec_t *out)
{
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int k = 0; k < nsteps && rc == 0; k++)
        rc = launch_step();
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc)
        return rc;
    if (e != hipSuccess)
        return fail(COSENU_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess)
        return fail(COSENU_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    // move the graph's launch data to the device now, not at its first launch
    e = hipGraphUpload(*out, s);
    if (e != hipSuccess) {
        (void)hipGraphExecDestroy(*out);
        *out = nullptr;
        return fail(COSENU_EHIP, std::string("hipGraphUpload: ") + hipGetErrorString(e));
    }
    return 0;
}

// Steps per captured graph.  Two- and four-stage launches alternate the
// state buffers every step: a graph of an even number of steps ends on the
// buffer it started from.
int graph_steps(const cosenu_ctx *c)
{
    int gsteps = env_int("COSENU_GRAPH_STEPS", 16);
    if (swaps_state(c))
        gsteps &= ~1;
    return gsteps;
}

// The graph of gsteps steps starting from the current state buffer
// (captured if there is none); returns it in *g.
int ensure_graph(cosenu_ctx *c, int gsteps, hipGraphExec_t *g)
{
    int k = c->graph_y[0] == c->y ? 0 : c->graph_y[1] == c->y ? 1 : -1;
    if (k >= 0 && c->graph[k] && c->graph_steps[k] == gsteps) {
        *g = c->graph[k];
        return 0;
    }
    if (k < 0) // a slot not holding the other buffer's graph
        k = (c->graph[0] && c->graph_y[0] == c->acc && swaps_state(c)) ? 1 : 0;
    if (c->graph[k]) {
        (void)hipGraphExecDestroy(c->graph[k]);
        c->graph[k] = nullptr;
    }
    c->graph_y[k] = c->y;
    auto one = [c]() -> int { return enqueue_step(c, c->stream); };
    RC_TRY(capture_graph(c->stream, gsteps, one, &c->graph[k]));
    c->graph_steps[k] = gsteps;
    *g = c->graph[k];
    return 0;
}

} // namespace

extern "C" {

const char *cosenu_last_error(void) { return g_err.c_str(); }
int cosenu_abi_version(void) { return COSENU_ABI_VERSION; }

int cosenu_create(const cosenu_params *p, cosenu_ctx **out)
{
    if (!p || !out)
        return fail(COSENU_EINVAL, "cosenu_create: null argument");
    *out = nullptr;
    RC_TRY(validate(p));
    HIP_TRY(hipSetDevice(p->device));

    cosenu_ctx *c = new cosenu_ctx();
    c->p = *p;
    c->d = derive(*p);
    c->nz = p->nz;
    c->nvz = p->nvz;
    c->VP = padded_bins(p->nvz);
    const size_t n_state = (size_t)c->nz * 8 * c->VP;
    std::vector<double> vz, Hm;
    fill_grids(*p, c->d, c->VP, vz, Hm);
    c->uniform = stream_supported(c->nvz, c->VP, COSENU_PERIODIC, vz.data()) != 0 &&
                 env_int("COSENU_UNIFORM", 1) != 0;
    c->stream_ok = p->mode == COSENU_FAST && c->uniform && p->bc == COSENU_PERIODIC;
    if (p->mode == COSENU_FAST && c->VP > 256 && !c->stream_ok) {
        // groups that straddle v = 0, padding bins or the open boundary: the
        // general form of the barrier-free kernel
        c->stream_ok = true;
        c->stream_general = true;
    }
    if (p->mode == COSENU_FAST) {
        // z_chunk < 0 in the tune table = contiguous-range mode with that many
        // times the resident teams
        int chunk;
        TuneVal tv;
        if (tune_lookup(c->VP, p->scheme, p->nz, &tv)) {
            c->variant = tv.variant;
            chunk = tv.chunk;
        } else if (c->stream_ok && !c->stream_general && c->nz >= 16 &&
                   (long)c->nz * (c->VP / 32) <= 2048) {
            // small problems (at most 2 wavefronts per SIMD in all): the
            // barrier-free kernel in chunks of 2 rows (nz = 1024, nvz = 64:
            // 27 us per FV step against 45 us for two-stage ranges; FD 20 us
            // against 28 us for four-stage ranges)
            c->variant = 3;
            chunk = c->nz % 2 ? 1 : 2;
        } else if (c->VP > 256) {
            // the team kernels hold a team of VP/32 wavefronts in one block
            // of at most 8: rows of 512 bins with FD and 3rd-order
            // dissipation take four stages per launch with a team over 2
            // blocks (nz = 16384, nvz = 512: 0.405 ms per step against 0.84
            // ms for two-stage launches; open, nvz = 480: 0.62 ms), other
            // rows of 512 and 1024 bins two stages per launch
            // with a team over 2 or 4 blocks, other wide grids the
            // barrier-free kernel
            c->variant =
                quad_wide_supported(c->VP, p->bc, p->scheme, p->flags) && c->nz >= 16 ? 5
                : pair_wide_supported(c->VP, p->bc) && c->nz >= 16                  ? 6
                                                                                     : 3;
            chunk = -1;
        } else {
            // FD with 3rd-order dissipation: four stages per launch (nz =
            // 65536, nvz = 128: 0.54 ms per step against 0.84 ms for two-stage
            // launches); else two stages per launch where it applies, else the
            // split-lane kernel; contiguous z ranges, one team per resident slot
            if (quad_supported(c->VP, p->bc, p->scheme, p->flags) && c->nz >= 16)
                c->variant = 5;
            else
                c->variant = pair_supported(c->VP, p->bc) && c->nz >= 16 ? 4 : 2;
            chunk = -1;
        }
        int ev = env_int("COSENU_FAST_VARIANT", 0), ez = env_int("COSENU_ZCHUNK", 0);
        if (ev >= 2 && ev <= 6 && ev != c->variant) {
            c->variant = ev;
            chunk = -1; // that variant's default launch: contiguous ranges
        }
        if (ez != 0)
            chunk = ez;
        if (c->VP > 256) {
            // as for narrower rows, four stages per launch fall back to two
            // (FV, 5th-order dissipation, rows of 1024 bins); the
            // barrier-free kernel takes every wide grid (stream_ok holds)
            const bool qw =
                quad_wide_supported(c->VP, p->bc, p->scheme, p->flags) && c->nz >= 16;
            const bool pw = pair_wide_supported(c->VP, p->bc) && c->nz >= 16;
            if (c->variant == 5 && !qw)
                c->variant = 6;
            if ((c->variant == 6 && !pw) || (c->variant != 5 && c->variant != 6))
                c->variant = 3;
        }
        if (c->variant == 6 && c->VP <= 256)
            c->variant = 4;
        if (c->VP > 256 && (c->variant == 5 || c->variant == 6) && chunk > 0)
            chunk = -1; // ranges only
        if (c->variant == 3 && !c->stream_ok)
            c->variant = 2; // shapes the barrier-free kernel does not take
        if (c->variant == 5 && c->VP <= 256 &&
            (!quad_supported(c->VP, p->bc, p->scheme, p->flags) || c->nz < 16))
            c->variant = 4; // FV, 5th-order dissipation
        if (c->variant == 4 && (!pair_supported(c->VP, p->bc) || c->nz < 16))
            c->variant = 2;
        if (chunk < 0) {
            c->nteams = range_teams(c->nz, c->VP, p->scheme, p->bc, c->variant, -chunk,
                                   c->stream_general);
            c->z_chunk = (c->nz + c->nteams - 1) / c->nteams;
            if (c->variant == 3)
                c->nteams4 = range_teams(c->nz, c->VP, p->scheme, p->bc, 3, -chunk,
                                         c->stream_general, true);
        } else {
            if (chunk <= 0 || chunk > c->nz || c->nz % chunk)
                chunk = default_chunk(c->nz, c->VP);
            c->z_chunk = chunk;
        }
    }

    int rc = 0;
    auto chk = [&](int r) {
        if (!rc)
            rc = r;
    };
    // the five states in one allocation
    chk(dev_alloc(c->bytes, &c->states, 5 * n_state));
    if (!rc) {
        c->y = c->states;
        c->acc = c->states + n_state;
        c->A = c->states + 2 * n_state;
        c->B = c->states + 3 * n_state;
        c->y0 = c->states + 4 * n_state;
    }
    if (!rc) chk(dev_alloc(c->bytes, &c->prof, (size_t)c->nz * 2 * c->VP));
    if (!rc) chk(dev_alloc(c->bytes, &c->vz, (size_t)c->VP));
    if (!rc) chk(dev_alloc(c->bytes, &c->Hm, (size_t)c->nz));
    for (int k = 0; k < 4 && c->stream_ok; k++)
        if (!rc) chk(dev_alloc(c->bytes, &c->mom[k], (size_t)c->nz * (c->VP / 32) * kMomSlot));
    c->npart = diag_partials_needed(c->nz);
    if (!rc) chk(dev_alloc(c->bytes, &c->diag_part, std::max((size_t)c->npart * kDiagSlots,
                                                             (size_t)c->npart * 4 * c->VP)));
    if (!rc) chk(dev_alloc(c->bytes, &c->diag_out, (size_t)std::max(kDiagSlots, 2 * c->VP)));
    if (rc) {
        release(c);
        delete c;
        return rc;
    }
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipEventCreateWithFlags(&c->snap_ready, hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipEventCreateWithFlags(&c->snap_done, hipEventDisableTiming);
    // Zero every buffer: padding bins must stay zero for the stage kernels.
    double *bufs[] = {c->y, c->acc, c->A, c->B, c->y0};
    for (double *b : bufs)
        if (e == hipSuccess)
            e = hipMemset(b, 0, n_state * sizeof(double));
    if (e == hipSuccess)
        e = hipMemset(c->prof, 0, (size_t)c->nz * 2 * c->VP * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpy(c->vz, vz.data(), c->VP * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(c->Hm, Hm.data(), c->nz * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        release(c);
        delete c;
        return fail(COSENU_EHIP, std::string("cosenu_create: ") + hipGetErrorString(e));
    }
    if (wide_teams(c)) {
        c->wide_give_up_test = env_int("COSENU_WIDE_GIVE_UP_TEST", 0) != 0;
        const int rcw = wide_buffers(c);
        if (rcw) {
            release(c);
            delete c;
            return rcw;
        }
    }
    c->view.y = c->y;
    c->view.y0 = c->y0;
    c->view.prof = c->prof;
    c->view.sa = c->A;
    c->view.sb = c->B;
    c->view.nz = c->nz;
    c->view.nvz = c->nvz;
    c->view.VP = c->VP;
    c->view.n_state = n_state;
    *out = c;
    return 0;
}

int cosenu_destroy(cosenu_ctx *c)
{
    if (!c)
        return 0;
    (void)hipSetDevice(c->p.device);
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamSynchronize(c->copy_stream);
    release(c);
    delete c;
    return 0;
}

int cosenu_get_info(cosenu_ctx *c, cosenu_info *info)
{
    if (!c || !info)
        return fail(COSENU_EINVAL, "cosenu_get_info: null argument");
    info->dz = c->d.dz;
    info->dv = c->d.dv;
    info->dt = c->d.dt;
    info->ct = c->d.ct;
    info->st = c->d.st;
    info->phy_time = c->phy_time;
    info->steps_done = c->steps;
    info->device_bytes = c->bytes;
    info->z_chunk = c->z_chunk;
    info->fast_variant = c->p.mode == COSENU_FAST ? c->variant : 0;
    info->range_teams = c->nteams;
    return 0;
}

int cosenu_set_profile(cosenu_ctx *c, const double *G, const double *bG, ptrdiff_t stride, int off)
{
    if (!c)
        return fail(COSENU_EINVAL, "cosenu_set_profile: null context");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    RC_TRY(view_set_profile(c->view, G, bG, stride, off, c->stream));
    c->have_profile = true;
    return 0;
}

int cosenu_upload_planes(cosenu_ctx *c, int which, const double *const planes[8], ptrdiff_t stride,
                         int off)
{
    if (!c)
        return fail(COSENU_EINVAL, "cosenu_upload_planes: null context");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    RC_TRY(view_upload_planes(c->view, which, planes, stride, off, c->stream));
    if (which == COSENU_STATE0)
        c->have_state0 = true;
    return 0;
}

int cosenu_download_planes(cosenu_ctx *c, int which, double *const planes[8], ptrdiff_t stride,
                           int off)
{
    if (!c || (which != COSENU_STATE && which != COSENU_STATE0))
        return fail(COSENU_EINVAL, "cosenu_download_planes: bad argument");
    HIP_TRY(hipSetDevice(c->p.device));
    RC_TRY(wide_check(c));
    return view_download_planes(c->view, which_ptr(c->view, which), planes, stride, off, c->stream);
}

int cosenu_upload_interleaved(cosenu_ctx *c, int which, const double *buf)
{
    if (!c)
        return fail(COSENU_EINVAL, "cosenu_upload_interleaved: null context");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    RC_TRY(view_upload_interleaved(c->view, which, buf, c->stream));
    if (which == COSENU_STATE0)
        c->have_state0 = true;
    return 0;
}

int cosenu_download_interleaved(cosenu_ctx *c, int which, double *buf)
{
    if (!c)
        return fail(COSENU_EINVAL, "cosenu_download_interleaved: null context");
    HIP_TRY(hipSetDevice(c->p.device));
    RC_TRY(wide_check(c));
    return view_download_interleaved(c->view, which, buf, c->stream);
}

int cosenu_set_profile_v(cosenu_ctx *c, const double *G, const double *bG)
{
    if (!c || !G || !bG)
        return fail(COSENU_EINVAL, "cosenu_set_profile_v: null argument");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    // per-bin values staged at the start of the (idle) stage buffer A
    HIP_TRY(hipMemcpyAsync(c->A, G, c->nvz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->A + c->VP, bG, c->nvz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_profile_v(c->prof, c->A, c->A + c->VP, c->nz, c->nvz, c->VP, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->have_profile = true;
    return 0;
}

int cosenu_init_separable(cosenu_ctx *c, int which, const double *a, const double *b)
{
    if (!c || !a || !b || (which != COSENU_STATE && which != COSENU_STATE0))
        return fail(COSENU_EINVAL, "cosenu_init_separable: bad argument");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    // factors staged in the (idle) stage buffer A: 8*VP per-bin, then 8*nz per-z
    std::vector<double> av((size_t)8 * c->VP, 0.0);
    for (int f = 0; f < 8; f++)
        for (int v = 0; v < c->nvz; v++)
            av[(size_t)f * c->VP + v] = a[(size_t)f * c->nvz + v];
    HIP_TRY(hipMemcpyAsync(c->A, av.data(), av.size() * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipMemcpyAsync(c->A + av.size(), b, (size_t)8 * c->nz * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_separable(which_ptr(c->view, which), c->A, c->A + av.size(), c->nz, c->nvz, c->VP,
                             c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (which == COSENU_STATE0)
        c->have_state0 = true;
    return 0;
}

int cosenu_copy_state(cosenu_ctx *c, int from, int to)
{
    if (!c || (from != COSENU_STATE && from != COSENU_STATE0) ||
        (to != COSENU_STATE && to != COSENU_STATE0))
        return fail(COSENU_EINVAL, "cosenu_copy_state: bad argument");
    HIP_TRY(hipSetDevice(c->p.device));
    if (from != to)
        HIP_TRY(hipMemcpyAsync(which_ptr(c->view, to), which_ptr(c->view, from),
                               c->view.n_state * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (to == COSENU_STATE0)
        c->have_state0 = true;
    return 0;
}

int cosenu_prepare(cosenu_ctx *c)
{
    if (!c)
        return fail(COSENU_EINVAL, "cosenu_prepare: null context");
    HIP_TRY(hipSetDevice(c->p.device));
    const int gsteps = graph_steps(c);
    if (c->timing || gsteps <= 0)
        return 0;
    hipGraphExec_t g = nullptr;
    RC_TRY(ensure_graph(c, gsteps, &g));
    if (swaps_state(c)) {
        // the graph that starts from the other buffer (after an odd number of
        // steps); capture runs nothing, the swap is bookkeeping only
        swap_state(c);
        const int rc = ensure_graph(c, gsteps, &g);
        swap_state(c);
        RC_TRY(rc);
    }
    return 0;
}

int cosenu_step(cosenu_ctx *c, int nsteps)
{
    if (!c || nsteps < 0)
        return fail(COSENU_EINVAL, "cosenu_step: bad argument");
    HIP_TRY(hipSetDevice(c->p.device));
    if (nsteps > 0 && c->p.mode == COSENU_FAST && c->variant == 3)
        // the state may have been replaced since the last step: its moments
        // are rebuilt here, after that every stage 4 leaves them current
        HIP_TRY(launch_stream_moments(c->y, c->vz, c->mom[0], c->nz, c->nvz, c->VP, c->stream));
    const int gsteps = graph_steps(c);
    int done = 0;
    if (!c->timing && gsteps > 0 && nsteps >= gsteps) {
        hipGraphExec_t g = nullptr;
        RC_TRY(ensure_graph(c, gsteps, &g));
        for (; done + gsteps <= nsteps; done += gsteps)
            HIP_TRY(hipGraphLaunch(g, c->stream));
    }
    for (; done < nsteps; done++) {
        RC_TRY(enqueue_step(c, c->stream));
        if (c->timing && c->ev.size() > 8192)
            RC_TRY(fold_timing(c));
    }
    // phy_time += dt per step, as step_rk4 does (nuosc.hpp:347)
    for (int k = 0; k < nsteps; k++)
        c->phy_time += c->d.dt;
    c->steps += nsteps;
    return 0;
}

int cosenu_sync(cosenu_ctx *c)
{
    if (!c)
        return fail(COSENU_EINVAL, "cosenu_sync: null context");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return wide_check(c);
}

int cosenu_rhs(cosenu_ctx *c, double *const out[8], ptrdiff_t stride, int off)
{
    if (!c)
        return fail(COSENU_EINVAL, "cosenu_rhs: null context");
    if (!valid_layout(c->view, stride, off))
        return fail(COSENU_EINVAL, "cosenu_rhs: bad layout");
    HIP_TRY(hipSetDevice(c->p.device));
    RhsArgs r = rhs_args(c, c->y, c->B);
    // a single RHS of a wide grid: the barrier-free kernel (the multi-block
    // kernels have no one-stage form)
    if (c->p.mode == COSENU_FAST && (c->variant == 3 || wide_teams(c))) {
        HIP_TRY(launch_stream_moments(c->y, c->vz, c->mom[0], c->nz, c->nvz, c->VP, c->stream));
        StageArgs a = stage_args(c, 1);
        a.yout = c->B;
        a.mom_in = c->mom[0];
        a.mom_out = nullptr;
        // the multi-block kernels' teams are not the barrier-free kernel's:
        // its own ranges
        const int nt = c->variant == 3 ? c->nteams
                                       : range_teams(c->nz, c->VP, c->p.scheme, c->p.bc, 3, 1,
                                                     c->stream_general);
        const int zc = c->variant == 3 ? c->z_chunk : (c->nz + nt - 1) / nt;
        HIP_TRY(launch_stage_stream(a, 0, zc, nt, c->stream_general, c->stream));
    } else if (c->p.mode == COSENU_FAST) { // variants 4, 5 evaluate single RHSs with variant 2
        HIP_TRY(launch_rhs_fast(r, c->z_chunk, c->nteams, c->stream));
    } else {
        HIP_TRY(launch_rhs_strict(r, c->stream));
    }
    return view_download_planes(c->view, c->B, out, stride, off, c->stream);
}

int cosenu_diagnostics(cosenu_ctx *c, cosenu_diag *d)
{
    if (!c || !d)
        return fail(COSENU_EINVAL, "cosenu_diagnostics: null argument");
    if (!c->have_profile || !c->have_state0)
        return fail(COSENU_ESTATE, "cosenu_diagnostics: set the profile and the t=0 state first");
    HIP_TRY(hipSetDevice(c->p.device));
    RC_TRY(wide_check(c));
    return view_diagnostics(c->view, c->d.dz, c->d.dv, c->diag_part, c->npart, c->diag_out, d,
                            c->stream);
}

int cosenu_survival_profile(cosenu_ctx *c, double *pee, double *pbee)
{
    if (!c || !pee || !pbee)
        return fail(COSENU_EINVAL, "cosenu_survival_profile: null argument");
    if (!c->have_state0)
        return fail(COSENU_ESTATE, "cosenu_survival_profile: set the t=0 state first");
    HIP_TRY(hipSetDevice(c->p.device));
    RC_TRY(wide_check(c));
    return view_survival(c->view, c->diag_part, c->npart, c->diag_out, pee, pbee, c->stream);
}

void *cosenu_host_alloc(size_t bytes)
{
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        g_err = "hipHostMalloc failed";
        return nullptr;
    }
    return p;
}

void cosenu_host_free(void *p)
{
    if (p)
        (void)hipHostFree(p);
}

int cosenu_snapshot_pee_async(cosenu_ctx *c, double *host)
{
    if (!c || !host)
        return fail(COSENU_EINVAL, "cosenu_snapshot_pee_async: null argument");
    if (!c->have_state0)
        return fail(COSENU_ESTATE, "cosenu_snapshot_pee_async: set the t=0 state first");
    HIP_TRY(hipSetDevice(c->p.device));
    const size_t n = (size_t)c->nvz * c->nz;
    if (!c->snap)
        RC_TRY(dev_alloc(c->bytes, &c->snap, n));
    // the previous copy must have drained before the buffer is rewritten
    HIP_TRY(hipStreamWaitEvent(c->stream, c->snap_done, 0));
    HIP_TRY(launch_pee_plane(c->snap, c->y, c->y0, c->nz, c->nvz, c->VP, c->stream));
    HIP_TRY(hipEventRecord(c->snap_ready, c->stream));
    HIP_TRY(hipStreamWaitEvent(c->copy_stream, c->snap_ready, 0));
    HIP_TRY(hipMemcpyAsync(host, c->snap, n * sizeof(double), hipMemcpyDeviceToHost, c->copy_stream));
    HIP_TRY(hipEventRecord(c->snap_done, c->copy_stream));
    c->snap_pending = true;
    return 0;
}

int cosenu_snapshot_wait(cosenu_ctx *c)
{
    if (!c)
        return fail(COSENU_EINVAL, "cosenu_snapshot_wait: null context");
    HIP_TRY(hipSetDevice(c->p.device));
    if (c->snap_pending)
        HIP_TRY(hipEventSynchronize(c->snap_done));
    c->snap_pending = false;
    return 0;
}

int cosenu_set_timing(cosenu_ctx *c, int enable)
{
    if (!c)
        return fail(COSENU_EINVAL, "cosenu_set_timing: null context");
    HIP_TRY(hipSetDevice(c->p.device));
    RC_TRY(fold_timing(c));
    for (int s = 0; s < 4; s++) {
        c->ms[s] = 0;
        c->cnt[s] = 0;
    }
    c->timing = enable != 0;
    return 0;
}

int cosenu_get_timing(cosenu_ctx *c, double ms[4], long long counts[4])
{
    if (!c || !ms || !counts)
        return fail(COSENU_EINVAL, "cosenu_get_timing: null argument");
    HIP_TRY(hipSetDevice(c->p.device));
    RC_TRY(fold_timing(c));
    for (int s = 0; s < 4; s++) {
        ms[s] = c->ms[s];
        counts[s] = c->cnt[s];
    }
    return 0;
}

const char *cosenu_stage_kernel_name(cosenu_ctx *c, int stage)
{
    if (!c)
        return "";
    if (c->p.mode == COSENU_FAST && c->variant == 3)
        return stage_stream_kernel_name(stage, c->stream_general);
    if (four_stage_launch(c))
        return stage != 0 ? "" : c->VP > 256 ? stage_quad_wide_kernel_name() : stage_quad_kernel_name();
    if (two_stage_launches(c) && c->variant == 6)
        return stage < 2 ? stage_pair_wide_kernel_name(stage) : "";
    if (two_stage_launches(c))
        return stage < 2 ? stage_pair_kernel_name(stage) : "";
    return c->p.mode == COSENU_FAST ? stage_fast_kernel_name(c->VP, stage)
                                    : stage_strict_kernel_name(stage);
}

// Times whole RK4 steps for every kernel variant and launch shape on a
// scratch context of this shape and records the fastest in the tune table.
int cosenu_autotune(const cosenu_params *p, int *variant, int *z_chunk, double *ms_best)
{
    if (!p)
        return fail(COSENU_EINVAL, "cosenu_autotune: null argument");
    cosenu_params q = *p;
    q.mode = COSENU_FAST;
    RC_TRY(validate(&q));
    cosenu_ctx *c = nullptr;
    RC_TRY(cosenu_create(&q, &c));
    int rc = 0;
    double best = 1e300;
    int bv = 2, bl = c->z_chunk;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = fail(COSENU_EHIP, "cosenu_autotune: hipEventCreate failed");
    // timed on a pseudo-random state and profile, not on zeros
    if (!rc && (launch_fill_pattern(c->y, (long)c->nz * 8, c->nvz, c->VP, c->stream) != hipSuccess ||
                launch_fill_pattern(c->prof, (long)c->nz * 2, c->nvz, c->VP, c->stream) != hipSuccess))
        rc = fail(COSENU_EHIP, "cosenu_autotune: fill failed");
    // chunk lengths, then (variants 2-6) contiguous ranges with 1 or 2 times
    // the resident teams, encoded as chunk = -1, -2
    const int cands[] = {1, 2, 4, 8, 16, 32, 64, 128, 256, -1, -2};
    for (int var = 2; var <= 6 && !rc; var++)
        for (int L : cands) {
            if (rc || (L > 0 && (L > c->nz || c->nz % L)) ||
                (var == 3 && !c->stream_ok) ||
                (var != 3 && var != 5 && var != 6 && c->VP > 256) ||
                (var == 4 && (!pair_supported(c->VP, q.bc) || c->nz < 16)) ||
                (var == 5 && c->VP <= 256 &&
                 (!quad_supported(c->VP, q.bc, q.scheme, q.flags) || c->nz < 16)) ||
                (var == 5 && c->VP > 256 &&
                 (L > 0 || !quad_wide_supported(c->VP, q.bc, q.scheme, q.flags) || c->nz < 16)) ||
                (var == 6 && (L > 0 || !pair_wide_supported(c->VP, q.bc) || c->nz < 16)))
                continue;
            c->variant = var;
            if (var == 3 && launch_stream_moments(c->y, c->vz, c->mom[0], c->nz, c->nvz,
                                                  c->VP, c->stream) != hipSuccess) {
                rc = fail(COSENU_EHIP, "cosenu_autotune: launch failed");
                break;
            }
            c->nteams4 = 0;
            if (L > 0) {
                c->z_chunk = L;
                c->nteams = 0;
            } else {
                c->nteams = range_teams(c->nz, c->VP, q.scheme, q.bc, var, -L, c->stream_general);
                c->z_chunk = (c->nz + c->nteams - 1) / c->nteams;
                if (var == 3)
                    c->nteams4 = range_teams(c->nz, c->VP, q.scheme, q.bc, 3, -L,
                                             c->stream_general, true);
                if (wide_teams(c) && (rc = wide_buffers(c)) != 0)
                    break;
            }
            // whole RK4 steps: the variants move different bytes per stage
            const int reps = 3;
            for (int k = 0; k < 1 + reps && !rc; k++) {
                if (k == 1 && hipEventRecord(e0, c->stream) != hipSuccess)
                    rc = fail(COSENU_EHIP, "cosenu_autotune: event record");
                if (!rc)
                    rc = enqueue_step(c, c->stream);
            }
            if (rc)
                break;
            float ms = 0;
            if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = fail(COSENU_EHIP, "cosenu_autotune: timing failed");
                break;
            }
            ms /= reps;
            if (wide_teams(c) && (rc = wide_check(c)) != 0)
                break;
            if (ms < best) {
                best = ms;
                bv = var;
                bl = L;
            }
        }
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    cosenu_destroy(c);
    if (rc)
        return rc;
    tune_store(padded_bins(q.nvz), q.scheme, q.nz, {bv, bl, best}); // ms per step
    if (variant)
        *variant = bv;
    if (z_chunk)
        *z_chunk = bl;
    if (ms_best)
        *ms_best = best;
    return 0;
}

} // extern "C"

// ===========================================================================
// Batched runs
namespace {

struct Group {
    int VP = 0, scheme = 0, flags = 0, L = 0;
    std::vector<int> runs;       // batch run indices
    size_t n_state = 0, n_prof = 0, n_vz = 0, n_hm = 0;
    double *y = nullptr, *acc = nullptr, *A = nullptr, *B = nullptr, *y0 = nullptr;
    double *prof = nullptr, *vz = nullptr, *Hm = nullptr;
    BatchRun *d_runs = nullptr;
    BatchRun *d_runs_sw = nullptr; // the same with y and acc swapped (two-stage launches)
    int2 *d_items = nullptr;
    int nitems = 0;
    bool uniform = false; // every run: nvz == VP, groups of 32 bins of one sign of v
    bool quad = false;    // four stages per launch (FD, 3rd-order dissipation), else two
    bool any_mat = false; // some run has the matter term
    bool std_phys = true; // every run: vacuum + collective, no matter term, KO3
};

} // namespace

struct cosenu_batch {
    int nruns = 0, device = 0;
    std::vector<cosenu_params> P;
    std::vector<Derived> D;
    std::vector<int> run_group;  // group of each run
    std::vector<int> run_slot;   // index within the group
    std::vector<size_t> off_state, off_prof; // offsets inside the group buffers
    std::vector<View> views;
    std::vector<bool> have_profile, have_state0;
    std::vector<Group> groups;
    hipStream_t stream = nullptr;
    double *diag_part = nullptr, *diag_out = nullptr;
    int npart = 1024;
    size_t bytes = 0;
    long long cells = 0, steps = 0;
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    // two stages per launch: the state of run r is in buffer `parity` of
    // (run_y[r], run_sp[r]) and the two alternate every step
    bool pairs = false;
    int parity = 0, graph_parity = 0;
    std::vector<double *> run_y, run_sp;
};

namespace {

void release_batch(cosenu_batch *b)
{
    for (Group &g : b->groups) {
        double *bufs[] = {g.y, g.acc, g.A, g.B, g.y0, g.prof, g.vz, g.Hm};
        for (double *p : bufs)
            if (p)
                (void)hipFree(p);
        if (g.d_runs)
            (void)hipFree(g.d_runs);
        if (g.d_runs_sw)
            (void)hipFree(g.d_runs_sw);
        if (g.d_items)
            (void)hipFree(g.d_items);
    }
    if (b->diag_part)
        (void)hipFree(b->diag_part);
    if (b->diag_out)
        (void)hipFree(b->diag_out);
    if (b->graph)
        (void)hipGraphExecDestroy(b->graph);
    if (b->stream)
        (void)hipStreamDestroy(b->stream);
}

int batch_enqueue_step(cosenu_batch *b, hipStream_t s)
{
    if (b->pairs) {
        for (int pair = 0; pair < 2; pair++)
            for (Group &g : b->groups) {
                const BatchRun *runs = b->parity ? g.d_runs_sw : g.d_runs;
                if (g.quad) {
                    if (pair == 0)
                        HIP_TRY(launch_stage_quad_batch(runs, g.d_items, g.nitems, g.VP, g.L,
                                                        g.uniform, g.any_mat, g.std_phys, s));
                } else {
                    HIP_TRY(launch_stage_pair_batch(runs, g.d_items, g.nitems, g.VP, g.scheme,
                                                    g.L, pair, g.uniform, g.std_phys, s));
                }
            }
        b->parity ^= 1;
        for (int r = 0; r < b->nruns; r++)
            b->views[r].y = b->parity ? b->run_sp[r] : b->run_y[r];
        return 0;
    }
    for (int st = 1; st <= 4; st++)
        for (Group &g : b->groups)
            HIP_TRY(launch_stage_fast_batch(g.d_runs, g.d_items, g.nitems, g.VP, g.scheme, g.L, st, s));
    return 0;
}

// Chunk length of a group's work items: the tuned one (entry nz = -1 of the
// tune table, or COSENU_BATCH_ZCHUNK), else one item per resident team.
int group_chunk(const cosenu_batch *b, const Group &g)
{
    int maxnz = 1;
    for (int r : g.runs)
        maxnz = std::max(maxnz, b->P[r].nz);
    int want = env_int("COSENU_BATCH_ZCHUNK", 0);
    TuneVal tv;
    if (want <= 0 && tune_lookup(g.VP, g.scheme, -1, &tv))
        want = tv.chunk;
    if (want <= 0) {
        long rows = 0;
        for (int r : g.runs)
            rows += b->P[r].nz;
        if (b->pairs) {
            // one work item per resident team, 32 to 512 cells (each item
            // recomputes 8 rows of halo for its second stage, 12 over three
            // stages with four stages per launch)
            const long target = std::max(1, g.quad ? quad_resident_teams(g.VP)
                                                   : pair_resident_teams(g.VP, g.scheme));
            want = (int)std::max(32L, std::min(512L, rows / target));
        } else {
            // about two work items per resident team, chunks of 8 to 32
            // cells: a few runs still spread over the whole device
            const long target = 2L * std::max(1, fast_resident_teams(g.VP, g.scheme, 0));
            want = (int)std::max(8L, std::min(32L, rows / target));
        }
    }
    // chunks need not divide a run's nz: a run's last work item takes the
    // remaining rows (the kernels take ranges shorter than the chunk)
    return std::max(1, std::min(want, maxnz));
}

// (Re)build a group's work items, (run, first z cell), for chunk length L
// (at most the group's largest nz).
hipError_t set_group_chunk(cosenu_batch *b, Group &g, int L)
{
    int maxnz = 1;
    for (int r : g.runs)
        maxnz = std::max(maxnz, b->P[r].nz);
    L = std::max(1, std::min(L, maxnz));
    std::vector<int2> items;
    for (size_t k = 0; k < g.runs.size(); k++)
        for (int j0 = 0; j0 < b->P[g.runs[k]].nz; j0 += L)
            items.push_back(make_int2((int)k, j0));
    if (g.d_items && (int)items.size() > g.nitems) {
        (void)hipFree(g.d_items);
        g.d_items = nullptr;
    }
    hipError_t e = hipSuccess;
    if (!g.d_items)
        e = hipMalloc((void **)&g.d_items, items.size() * sizeof(int2));
    if (e == hipSuccess)
        e = hipMemcpy(g.d_items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        g.L = L;
        g.nitems = (int)items.size();
    }
    return e;
}

} // namespace

extern "C" {

int cosenu_batch_create(const cosenu_params *runs, int nruns, cosenu_batch **out)
{
    if (!runs || nruns < 1 || !out)
        return fail(COSENU_EINVAL, "cosenu_batch_create: bad argument");
    *out = nullptr;
    for (int r = 0; r < nruns; r++) {
        cosenu_params q = runs[r];
        q.mode = COSENU_FAST;
        int rc = validate(&q);
        if (rc)
            return fail(rc, "run " + std::to_string(r) + ": " + g_err);
        if (runs[r].device != runs[0].device)
            return fail(COSENU_EINVAL, "all runs of a batch must use one device");
        if (runs[r].bc != COSENU_PERIODIC)
            return fail(COSENU_ENOTSUP, "run " + std::to_string(r) + ": batches take periodic runs");
        if (runs[r].nvz > 256)
            return fail(COSENU_ENOTSUP, "run " + std::to_string(r) + ": batches take nvz up to 256");
    }
    HIP_TRY(hipSetDevice(runs[0].device));
    cosenu_batch *b = new cosenu_batch();
    b->nruns = nruns;
    b->device = runs[0].device;
    b->P.assign(runs, runs + nruns);
    b->run_group.resize(nruns);
    b->run_slot.resize(nruns);
    b->off_state.resize(nruns);
    b->off_prof.resize(nruns);
    b->views.resize(nruns);
    b->run_y.assign(nruns, nullptr);
    b->run_sp.assign(nruns, nullptr);
    b->have_profile.assign(nruns, false);
    b->have_state0.assign(nruns, false);
    // group by (VP, scheme)
    std::map<std::pair<int, int>, int> gid;
    for (int r = 0; r < nruns; r++) {
        b->P[r].mode = COSENU_FAST;
        b->D.push_back(derive(b->P[r]));
        const int VP = padded_bins(runs[r].nvz);
        auto key = std::make_pair(VP, runs[r].scheme);
        if (!gid.count(key)) {
            gid[key] = (int)b->groups.size();
            Group g;
            g.VP = VP;
            g.scheme = runs[r].scheme;
            b->groups.push_back(g);
        }
        Group &g = b->groups[gid[key]];
        b->run_group[r] = gid[key];
        b->run_slot[r] = (int)g.runs.size();
        b->off_state[r] = g.n_state;
        b->off_prof[r] = g.n_prof;
        g.runs.push_back(r);
        g.n_state += (size_t)runs[r].nz * 8 * VP;
        g.n_prof += (size_t)runs[r].nz * 2 * VP;
        g.n_vz += VP;
        g.n_hm += runs[r].nz;
        b->cells += (long long)runs[r].nz * runs[r].nvz;
    }
    // two stages per launch unless COSENU_BATCH_VARIANT=2 (one stage per
    // launch); groups of FD runs with 3rd-order dissipation four stages per
    // launch unless COSENU_BATCH_VARIANT=4
    const int bvar = env_int("COSENU_BATCH_VARIANT", 5);
    b->pairs = bvar != 2;
    for (int r = 0; r < nruns; r++)
        b->pairs = b->pairs && runs[r].nz >= 16;
    for (const Group &g : b->groups)
        b->pairs = b->pairs && pair_supported(g.VP, COSENU_PERIODIC);
    for (Group &g : b->groups) {
        g.quad = b->pairs && bvar != 4;
        for (int r : g.runs) {
            g.quad = g.quad && quad_supported(g.VP, COSENU_PERIODIC, runs[r].scheme, runs[r].flags);
            g.any_mat = g.any_mat || (runs[r].flags & COSENU_MAT) != 0;
            g.std_phys = g.std_phys && runs[r].flags == (COSENU_VAC | COSENU_COLL);
        }
    }
    int rc = 0;
    for (Group &g : b->groups) {
        if (!rc) rc = dev_alloc(b->bytes, &g.y, g.n_state);
        if (!rc) rc = dev_alloc(b->bytes, &g.acc, g.n_state);
        if (!rc) rc = dev_alloc(b->bytes, &g.A, g.n_state);
        if (!rc) rc = dev_alloc(b->bytes, &g.B, g.n_state);
        if (!rc) rc = dev_alloc(b->bytes, &g.y0, g.n_state);
        if (!rc) rc = dev_alloc(b->bytes, &g.prof, g.n_prof);
        if (!rc) rc = dev_alloc(b->bytes, &g.vz, g.n_vz);
        if (!rc) rc = dev_alloc(b->bytes, &g.Hm, g.n_hm);
    }
    int maxVP = 64;
    for (Group &g : b->groups)
        maxVP = std::max(maxVP, g.VP);
    if (!rc) rc = dev_alloc(b->bytes, &b->diag_part, std::max((size_t)b->npart * kDiagSlots,
                                                              (size_t)b->npart * 4 * maxVP));
    if (!rc) rc = dev_alloc(b->bytes, &b->diag_out, (size_t)std::max(kDiagSlots, 2 * maxVP));
    hipError_t e = hipSuccess;
    if (!rc)
        e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    for (Group &g : b->groups) {
        if (rc || e != hipSuccess)
            break;
        double *bufs[] = {g.y, g.acc, g.A, g.B, g.y0};
        for (double *p : bufs)
            if (e == hipSuccess)
                e = hipMemset(p, 0, g.n_state * sizeof(double));
        if (e == hipSuccess)
            e = hipMemset(g.prof, 0, g.n_prof * sizeof(double));
        // per-run tables
        g.L = group_chunk(b, g);
        g.uniform = true;
        std::vector<BatchRun> hr;
        std::vector<double> vz_all, hm_all;
        size_t vo = 0, ho = 0;
        for (size_t k = 0; k < g.runs.size(); k++) {
            const int r = g.runs[k];
            const cosenu_params &p = b->P[r];
            const Derived &d = b->D[r];
            std::vector<double> vz, Hm;
            fill_grids(p, d, g.VP, vz, Hm);
            g.uniform = g.uniform && stream_supported(p.nvz, g.VP, COSENU_PERIODIC, vz.data());
            vz_all.insert(vz_all.end(), vz.begin(), vz.end());
            hm_all.insert(hm_all.end(), Hm.begin(), Hm.end());
            BatchRun br;
            const size_t so = b->off_state[r];
            br.y = g.y + so; br.acc = g.acc + so; br.A = g.A + so; br.B = g.B + so;
            br.vz = g.vz + vo; br.Hm = g.Hm + ho;
            br.nz = p.nz; br.nvz = p.nvz; br.VP = g.VP;
            br.flags = p.flags; br.bc = p.bc; br.scheme = p.scheme;
            br.a[0] = 0.5 * d.dt; br.a[1] = 0.5 * d.dt; br.a[2] = d.dt; br.a[3] = 1.0 / 6.0 * d.dt;
            br.dz = d.dz; br.dv = d.dv; br.ct = d.ct; br.st = d.st;
            br.pmo = p.pmo; br.mu = p.mu; br.ko = p.ko;
            hr.push_back(br);
            b->run_y[r] = br.y;
            b->run_sp[r] = br.acc;
            View &v = b->views[r];
            v.y = br.y; v.y0 = g.y0 + so; v.prof = g.prof + b->off_prof[r];
            v.sa = br.A; v.sb = br.B;
            v.nz = p.nz; v.nvz = p.nvz; v.VP = g.VP;
            v.n_state = (size_t)p.nz * 8 * g.VP;
            vo += g.VP;
            ho += p.nz;
        }
        std::vector<BatchRun> hs = hr;
        for (BatchRun &br : hs)
            std::swap(br.y, br.acc);
        if (e == hipSuccess) e = hipMalloc((void **)&g.d_runs, hr.size() * sizeof(BatchRun));
        if (e == hipSuccess) e = hipMalloc((void **)&g.d_runs_sw, hs.size() * sizeof(BatchRun));
        if (e == hipSuccess)
            e = hipMemcpy(g.d_runs_sw, hs.data(), hs.size() * sizeof(BatchRun), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(g.d_runs, hr.data(), hr.size() * sizeof(BatchRun), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = set_group_chunk(b, g, g.L);
        if (e == hipSuccess)
            e = hipMemcpy(g.vz, vz_all.data(), vz_all.size() * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(g.Hm, hm_all.data(), hm_all.size() * sizeof(double), hipMemcpyHostToDevice);
    }
    if (!rc && e != hipSuccess)
        rc = fail(COSENU_EHIP, std::string("cosenu_batch_create: ") + hipGetErrorString(e));
    if (rc) {
        release_batch(b);
        delete b;
        return rc;
    }
    *out = b;
    return 0;
}

int cosenu_batch_destroy(cosenu_batch *b)
{
    if (!b)
        return 0;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    release_batch(b);
    delete b;
    return 0;
}

static int batch_run_ok(cosenu_batch *b, int run)
{
    if (!b || run < 0 || run >= b->nruns)
        return fail(COSENU_EINVAL, "batch: bad run index");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

int cosenu_batch_set_profile(cosenu_batch *b, int run, const double *G, const double *bG,
                             ptrdiff_t stride, int off)
{
    RC_TRY(batch_run_ok(b, run));
    RC_TRY(view_set_profile(b->views[run], G, bG, stride, off, b->stream));
    b->have_profile[run] = true;
    return 0;
}

int cosenu_batch_upload_planes(cosenu_batch *b, int run, int which, const double *const planes[8],
                               ptrdiff_t stride, int off)
{
    RC_TRY(batch_run_ok(b, run));
    RC_TRY(view_upload_planes(b->views[run], which, planes, stride, off, b->stream));
    if (which == COSENU_STATE0)
        b->have_state0[run] = true;
    return 0;
}

int cosenu_batch_download_planes(cosenu_batch *b, int run, int which, double *const planes[8],
                                 ptrdiff_t stride, int off)
{
    RC_TRY(batch_run_ok(b, run));
    if (which != COSENU_STATE && which != COSENU_STATE0)
        return fail(COSENU_EINVAL, "cosenu_batch_download_planes: bad `which`");
    return view_download_planes(b->views[run], which_ptr(b->views[run], which), planes, stride, off,
                                b->stream);
}

int cosenu_batch_step(cosenu_batch *b, int nsteps)
{
    if (!b || nsteps < 0)
        return fail(COSENU_EINVAL, "cosenu_batch_step: bad argument");
    HIP_TRY(hipSetDevice(b->device));
    int gsteps = env_int("COSENU_GRAPH_STEPS", 16);
    if (b->pairs)
        gsteps &= ~1; // an even number of steps ends on the buffers it started from
    int done = 0;
    if (gsteps > 0 && nsteps >= gsteps) {
        if (!b->graph || b->graph_steps != gsteps || b->graph_parity != b->parity) {
            if (b->graph) {
                (void)hipGraphExecDestroy(b->graph);
                b->graph = nullptr;
            }
            b->graph_parity = b->parity;
            auto one = [b]() -> int { return batch_enqueue_step(b, b->stream); };
            RC_TRY(capture_graph(b->stream, gsteps, one, &b->graph));
            b->graph_steps = gsteps;
        }
        for (; done + gsteps <= nsteps; done += gsteps)
            HIP_TRY(hipGraphLaunch(b->graph, b->stream));
    }
    for (; done < nsteps; done++)
        RC_TRY(batch_enqueue_step(b, b->stream));
    b->steps += nsteps;
    return 0;
}

int cosenu_batch_sync(cosenu_batch *b)
{
    if (!b)
        return fail(COSENU_EINVAL, "cosenu_batch_sync: null batch");
    HIP_TRY(hipSetDevice(b->device));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

int cosenu_batch_diagnostics(cosenu_batch *b, int run, cosenu_diag *d)
{
    RC_TRY(batch_run_ok(b, run));
    if (!d)
        return fail(COSENU_EINVAL, "cosenu_batch_diagnostics: null argument");
    if (!b->have_profile[run] || !b->have_state0[run])
        return fail(COSENU_ESTATE, "cosenu_batch_diagnostics: set the profile and the t=0 state first");
    return view_diagnostics(b->views[run], b->D[run].dz, b->D[run].dv, b->diag_part,
                            std::min(b->npart, diag_partials_needed(b->P[run].nz)), b->diag_out, d,
                            b->stream);
}

int cosenu_batch_survival_profile(cosenu_batch *b, int run, double *pee, double *pbee)
{
    RC_TRY(batch_run_ok(b, run));
    if (!pee || !pbee)
        return fail(COSENU_EINVAL, "cosenu_batch_survival_profile: null argument");
    if (!b->have_state0[run])
        return fail(COSENU_ESTATE, "cosenu_batch_survival_profile: set the t=0 state first");
    return view_survival(b->views[run], b->diag_part,
                         std::min(b->npart, diag_partials_needed(b->P[run].nz)), b->diag_out, pee,
                         pbee, b->stream);
}

int cosenu_batch_get_info(cosenu_batch *b, cosenu_batch_info *info)
{
    if (!b || !info)
        return fail(COSENU_EINVAL, "cosenu_batch_get_info: null argument");
    info->nruns = b->nruns;
    info->ngroups = (int)b->groups.size();
    info->cells = b->cells;
    info->steps_done = b->steps;
    info->device_bytes = b->bytes;
    info->work_items = 0;
    for (const Group &g : b->groups)
        info->work_items += g.nitems;
    info->z_chunk = b->groups.empty() ? 0 : b->groups[0].L;
    info->fast_variant = b->pairs ? 4 : 2;
    for (const Group &g : b->groups)
        if (g.quad)
            info->fast_variant = 5;
    return 0;
}

int cosenu_batch_autotune(const cosenu_params *runs, int nruns, int *z_chunk, double *ms_best)
{
    cosenu_batch *b = nullptr;
    RC_TRY(cosenu_batch_create(runs, nruns, &b));
    int rc = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = fail(COSENU_EHIP, "cosenu_batch_autotune: hipEventCreate failed");
    // timed on pseudo-random states and profiles, not on zeros
    for (Group &g : b->groups)
        if (!rc && (launch_fill_pattern(g.y, (long)(g.n_state / g.VP), g.VP, g.VP, b->stream) != hipSuccess ||
                    launch_fill_pattern(g.prof, (long)(g.n_prof / g.VP), g.VP, g.VP, b->stream) != hipSuccess))
            rc = fail(COSENU_EHIP, "cosenu_batch_autotune: fill failed");
    double best = 1e300;
    int bl = 0;
    // chunk lengths of the work items; whole batch steps (zero states) timed
    const int cands[] = {32, 64, 128, 256, 512};
    for (int L : cands) {
        for (Group &g : b->groups)
            if (!rc && set_group_chunk(b, g, L) != hipSuccess)
                rc = fail(COSENU_EHIP, "cosenu_batch_autotune: work items");
        const int reps = 3;
        for (int k = 0; k < 1 + reps && !rc; k++) {
            if (k == 1 && hipEventRecord(e0, b->stream) != hipSuccess)
                rc = fail(COSENU_EHIP, "cosenu_batch_autotune: event record");
            if (!rc)
                rc = batch_enqueue_step(b, b->stream);
        }
        if (rc)
            break;
        float ms = 0;
        if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = fail(COSENU_EHIP, "cosenu_batch_autotune: timing failed");
            break;
        }
        if (ms / reps < best) {
            best = ms / reps;
            bl = L;
        }
    }
    if (!rc)
        for (const Group &g : b->groups) // batches created afterwards use it
            tune_store(g.VP, g.scheme, -1, {g.quad ? 5 : b->pairs ? 4 : 2, bl, best});
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    cosenu_batch_destroy(b);
    if (rc)
        return rc;
    if (z_chunk)
        *z_chunk = bl;
    if (ms_best)
        *ms_best = best;
    return 0;
}

} // extern "C"
