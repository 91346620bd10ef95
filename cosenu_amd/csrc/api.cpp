// C ABI of the COSEnu MI355X library (include/cosenu_hip.h): context
// lifetime, layout conversion on upload/download, the RK4 driver, timing,
// diagnostics and snapshots.  Host code only; kernels live in the .hip files.
#include "cosenu_hip.h"
#include "common.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
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

int env_int(const char *name, int dflt)
{
    const char *s = std::getenv(name);
    return (s && *s) ? std::atoi(s) : dflt;
}

} // namespace

struct cosenu_ctx {
    cosenu_params p;
    int nz = 0, nvz = 0, VP = 0;
    double dz = 0, dv = 0, dt = 0, ct = 0, st = 0;
    double phy_time = 0;
    long long steps = 0;
    size_t n_state = 0; // doubles per state in row layout
    size_t bytes = 0;
    hipStream_t stream = nullptr, copy_stream = nullptr;
    double *y = nullptr, *acc = nullptr, *A = nullptr, *B = nullptr, *y0 = nullptr;
    double *prof = nullptr, *vz = nullptr, *Hm = nullptr;
    double *diag_part = nullptr, *diag_out = nullptr, *snap = nullptr;
    int npart = 0;
    bool have_profile = false, have_state0 = false;
    int z_chunk = 0;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev; // start/stop pairs, stage s = (pair index) % 4
    std::vector<int> ev_stage;
    double ms[4] = {0, 0, 0, 0};
    long long cnt[4] = {0, 0, 0, 0};
    // graph of `graph_steps` RK4 steps
    hipGraphExec_t graph = nullptr;
    int graph_steps = 0;
    hipEvent_t snap_ready = nullptr, snap_done = nullptr;
    bool snap_pending = false;
};

namespace {

int dev_alloc(cosenu_ctx *c, double **p, size_t n)
{
    hipError_t e = hipMalloc((void **)p, n * sizeof(double));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(COSENU_ENOMEM, "hipMalloc of " + std::to_string(n * sizeof(double)) +
                                       " bytes failed: " + hipGetErrorString(e));
    }
    c->bytes += n * sizeof(double);
    return 0;
}

void release(cosenu_ctx *c)
{
    double *bufs[] = {c->y, c->acc, c->A, c->B, c->y0, c->prof, c->vz, c->Hm,
                      c->diag_part, c->diag_out, c->snap};
    for (double *b : bufs)
        if (b)
            (void)hipFree(b);
    for (hipEvent_t e : c->ev)
        (void)hipEventDestroy(e);
    if (c->graph)
        (void)hipGraphExecDestroy(c->graph);
    if (c->snap_ready)
        (void)hipEventDestroy(c->snap_ready);
    if (c->snap_done)
        (void)hipEventDestroy(c->snap_done);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    if (c->copy_stream)
        (void)hipStreamDestroy(c->copy_stream);
}

// z cells per streaming chunk in FAST mode: aim at ~4 wavefronts per SIMD
// over 1024 SIMDs, within [8, 256], and a divisor of nz.
int pick_chunk(int nz, int VP)
{
    int forced = env_int("COSENU_ZCHUNK", 0);
    int W = VP / 64;
    long target = (long)nz * W / 4096;
    int L = forced > 0 ? forced : (int)std::max(8L, std::min(256L, target));
    if (L > nz)
        L = nz;
    while (L > 1 && nz % L)
        L--;
    return L;
}

StageArgs stage_args(cosenu_ctx *c, int stage)
{
    StageArgs a;
    a.y = c->y;
    a.acc = c->acc;
    switch (stage) {
    case 1: a.yin = c->y; a.yout = c->A; a.a = 0.5 * c->dt; break;
    case 2: a.yin = c->A; a.yout = c->B; a.a = 0.5 * c->dt; break;
    case 3: a.yin = c->B; a.yout = c->A; a.a = c->dt; break;
    default: a.yin = c->A; a.yout = c->y; a.a = 1.0 / 6.0 * c->dt; break;
    }
    a.vz = c->vz;
    a.Hm = c->Hm;
    a.nz = c->nz; a.nvz = c->nvz; a.VP = c->VP;
    a.flags = c->p.flags; a.bc = c->p.bc; a.scheme = c->p.scheme;
    a.dz = c->dz; a.dv = c->dv; a.ct = c->ct; a.st = c->st; a.pmo = c->p.pmo; a.mu = c->p.mu;
    a.ko = c->p.ko;
    return a;
}

RhsArgs rhs_args(cosenu_ctx *c, const double *in, double *out)
{
    RhsArgs r;
    r.in = in; r.out = out; r.vz = c->vz; r.Hm = c->Hm;
    r.nz = c->nz; r.nvz = c->nvz; r.VP = c->VP;
    r.flags = c->p.flags; r.bc = c->p.bc; r.scheme = c->p.scheme;
    r.dz = c->dz; r.dv = c->dv; r.ct = c->ct; r.st = c->st; r.pmo = c->p.pmo; r.mu = c->p.mu;
    r.ko = c->p.ko;
    return r;
}

hipError_t launch_stage(cosenu_ctx *c, int stage, hipStream_t s)
{
    StageArgs a = stage_args(c, stage);
    if (c->p.mode == COSENU_FAST)
        return launch_stage_fast(a, stage, c->z_chunk, s);
    return launch_stage_strict(a, stage, s);
}

int enqueue_step(cosenu_ctx *c, hipStream_t s)
{
    for (int st = 1; st <= 4; st++) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (c->timing) {
            HIP_TRY(hipEventCreate(&e0));
            HIP_TRY(hipEventCreate(&e1));
            HIP_TRY(hipEventRecord(e0, s));
        }
        HIP_TRY(launch_stage(c, st, s));
        if (c->timing) {
            HIP_TRY(hipEventRecord(e1, s));
            c->ev.push_back(e0);
            c->ev.push_back(e1);
            c->ev_stage.push_back(st - 1);
        }
    }
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

int build_graph(cosenu_ctx *c, int nsteps)
{
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int k = 0; k < nsteps && rc == 0; k++)
        for (int st = 1; st <= 4 && rc == 0; st++) {
            hipError_t e = launch_stage(c, st, c->stream);
            if (e != hipSuccess)
                rc = fail(COSENU_EHIP, std::string("capture launch: ") + hipGetErrorString(e));
        }
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc)
        return rc;
    if (e != hipSuccess)
        return fail(COSENU_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess)
        return fail(COSENU_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    c->graph_steps = nsteps;
    return 0;
}

bool valid_layout(const cosenu_ctx *c, ptrdiff_t stride, int off)
{
    return off >= 0 && stride >= (ptrdiff_t)off + c->nz;
}

double *target(cosenu_ctx *c, int which) { return which == COSENU_STATE0 ? c->y0 : c->y; }

} // namespace

extern "C" {

const char *cosenu_last_error(void) { return g_err.c_str(); }
int cosenu_abi_version(void) { return COSENU_ABI_VERSION; }

int cosenu_create(const cosenu_params *p, cosenu_ctx **out)
{
    if (!p || !out)
        return fail(COSENU_EINVAL, "cosenu_create: null argument");
    *out = nullptr;
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
    const int VP = ((p->nvz + 63) / 64) * 64;
    if (p->mode == COSENU_FAST) {
        if (!fast_supported(VP, p->bc))
            return fail(COSENU_ENOTSUP, "FAST mode supports periodic runs with nvz up to 256");
        if ((p->flags & COSENU_KO5) != 0)
            return fail(COSENU_ENOTSUP, "FAST mode implements the 3rd-order KO dissipation only");
    }
    HIP_TRY(hipSetDevice(p->device));

    cosenu_ctx *c = new cosenu_ctx();
    c->p = *p;
    c->nz = p->nz;
    c->nvz = p->nvz;
    c->VP = VP;
    // NuOsc::NuOsc (nuosc.hpp:78-80) and set_vac_pars (nuosc.hpp:132-138)
    c->dz = (p->z1 - p->z0) / p->nz;
    c->dv = (p->v1 - p->v0) / (p->nvz);
    c->dt = p->CFL * c->dz;
    double th = p->theta * M_PI / 180.0;
    c->ct = (1.0 / 2.0) * p->omega * cos(2.0 * th);
    c->st = (1.0 / 2.0) * p->omega * sin(2.0 * th);
    c->n_state = (size_t)c->nz * 8 * VP;
    c->z_chunk = p->mode == COSENU_FAST ? pick_chunk(c->nz, VP) : 0;

    int rc = 0;
    auto chk = [&](int r) {
        if (!rc)
            rc = r;
    };
    chk(dev_alloc(c, &c->y, c->n_state));
    if (!rc) chk(dev_alloc(c, &c->acc, c->n_state));
    if (!rc) chk(dev_alloc(c, &c->A, c->n_state));
    if (!rc) chk(dev_alloc(c, &c->B, c->n_state));
    if (!rc) chk(dev_alloc(c, &c->y0, c->n_state));
    if (!rc) chk(dev_alloc(c, &c->prof, (size_t)c->nz * 2 * VP));
    if (!rc) chk(dev_alloc(c, &c->vz, VP));
    if (!rc) chk(dev_alloc(c, &c->Hm, c->nz));
    c->npart = diag_partials_needed(c->nz);
    if (!rc) chk(dev_alloc(c, &c->diag_part, std::max((size_t)c->npart * kDiagSlots,
                                                      (size_t)c->npart * 4 * VP)));
    if (!rc) chk(dev_alloc(c, &c->diag_out, std::max(kDiagSlots, 2 * VP)));
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
            e = hipMemset(b, 0, c->n_state * sizeof(double));
    if (e == hipSuccess)
        e = hipMemset(c->prof, 0, (size_t)c->nz * 2 * VP * sizeof(double));
    // Velocity grid and matter profile (nuosc.hpp:82-96).
    std::vector<double> vz(VP, 0.0), Hm(c->nz);
    for (int i = 0; i < p->nvz; i++)
        vz[i] = p->v0 + (i + 0.5) * c->dv;
    for (int j = 0; j < c->nz; j++)
        Hm[j] = -1.0e-3 * ((p->z0 + (j + 0.5) * c->dz) + 600.0);
    if (e == hipSuccess)
        e = hipMemcpy(c->vz, vz.data(), VP * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(c->Hm, Hm.data(), c->nz * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        release(c);
        delete c;
        return fail(COSENU_EHIP, std::string("cosenu_create: ") + hipGetErrorString(e));
    }
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
    info->dz = c->dz;
    info->dv = c->dv;
    info->dt = c->dt;
    info->ct = c->ct;
    info->st = c->st;
    info->phy_time = c->phy_time;
    info->steps_done = c->steps;
    info->device_bytes = c->bytes;
    info->z_chunk = c->z_chunk;
    return 0;
}

int cosenu_set_profile(cosenu_ctx *c, const double *G, const double *bG, ptrdiff_t stride, int off)
{
    if (!c || !G || !bG || !valid_layout(c, stride, off))
        return fail(COSENU_EINVAL, "cosenu_set_profile: bad argument");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    // stage the two planes in A and B
    const size_t w = (size_t)c->nz * sizeof(double), pitch = (size_t)stride * sizeof(double);
    if ((size_t)stride * c->nvz > c->n_state)
        return fail(COSENU_EINVAL, "cosenu_set_profile: row_stride too large");
    HIP_TRY(hipMemcpy2DAsync(c->A + off, pitch, G + off, pitch, w, c->nvz, hipMemcpyHostToDevice,
                             c->stream));
    HIP_TRY(hipMemcpy2DAsync(c->B + off, pitch, bG + off, pitch, w, c->nvz, hipMemcpyHostToDevice,
                             c->stream));
    HIP_TRY(launch_import_profile(c->prof, c->A, c->B, c->nz, c->nvz, c->VP, stride, off, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->have_profile = true;
    return 0;
}

int cosenu_upload_planes(cosenu_ctx *c, int which, const double *const planes[8], ptrdiff_t stride,
                         int off)
{
    if (!c || !planes || !valid_layout(c, stride, off) || (which != COSENU_STATE && which != COSENU_STATE0))
        return fail(COSENU_EINVAL, "cosenu_upload_planes: bad argument");
    if ((size_t)stride * c->nvz > c->n_state)
        return fail(COSENU_EINVAL, "cosenu_upload_planes: row_stride too large");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double *dst = target(c, which);
    const size_t w = (size_t)c->nz * sizeof(double), pitch = (size_t)stride * sizeof(double);
    for (int f = 0; f < 8; f++) {
        if (!planes[f])
            return fail(COSENU_EINVAL, "cosenu_upload_planes: null plane");
        double *stage = (f & 1) ? c->B : c->A;
        HIP_TRY(hipMemcpy2DAsync(stage + off, pitch, planes[f] + off, pitch, w, c->nvz,
                                 hipMemcpyHostToDevice, c->stream));
        HIP_TRY(launch_import_plane(dst, stage, f, c->nz, c->nvz, c->VP, stride, off, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (which == COSENU_STATE0)
        c->have_state0 = true;
    return 0;
}

int cosenu_download_planes(cosenu_ctx *c, int which, double *const planes[8], ptrdiff_t stride,
                           int off)
{
    if (!c || !planes || !valid_layout(c, stride, off) || (which != COSENU_STATE && which != COSENU_STATE0))
        return fail(COSENU_EINVAL, "cosenu_download_planes: bad argument");
    if ((size_t)stride * c->nvz > c->n_state)
        return fail(COSENU_EINVAL, "cosenu_download_planes: row_stride too large");
    HIP_TRY(hipSetDevice(c->p.device));
    const double *src = target(c, which);
    const size_t w = (size_t)c->nz * sizeof(double), pitch = (size_t)stride * sizeof(double);
    for (int f = 0; f < 8; f++) {
        if (!planes[f])
            return fail(COSENU_EINVAL, "cosenu_download_planes: null plane");
        HIP_TRY(launch_export_plane(c->A, src, f, c->nz, c->nvz, c->VP, stride, off, c->stream));
        HIP_TRY(hipMemcpy2DAsync(planes[f] + off, pitch, c->A + off, pitch, w, c->nvz,
                                 hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return 0;
}

int cosenu_upload_interleaved(cosenu_ctx *c, int which, const double *buf)
{
    if (!c || !buf || (which != COSENU_STATE && which != COSENU_STATE0))
        return fail(COSENU_EINVAL, "cosenu_upload_interleaved: bad argument");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const size_t n = (size_t)c->nvz * c->nz * 8;
    HIP_TRY(hipMemcpyAsync(c->A, buf, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_import_interleaved(target(c, which), c->A, c->nz, c->nvz, c->VP, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (which == COSENU_STATE0)
        c->have_state0 = true;
    return 0;
}

int cosenu_download_interleaved(cosenu_ctx *c, int which, double *buf)
{
    if (!c || !buf || (which != COSENU_STATE && which != COSENU_STATE0))
        return fail(COSENU_EINVAL, "cosenu_download_interleaved: bad argument");
    HIP_TRY(hipSetDevice(c->p.device));
    const size_t n = (size_t)c->nvz * c->nz * 8;
    HIP_TRY(launch_export_interleaved(c->A, target(c, which), c->nz, c->nvz, c->VP, c->stream));
    HIP_TRY(hipMemcpyAsync(buf, c->A, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int cosenu_step(cosenu_ctx *c, int nsteps)
{
    if (!c || nsteps < 0)
        return fail(COSENU_EINVAL, "cosenu_step: bad argument");
    HIP_TRY(hipSetDevice(c->p.device));
    const int gsteps = env_int("COSENU_GRAPH_STEPS", 16);
    int done = 0;
    if (!c->timing && gsteps > 0 && nsteps >= gsteps) {
        if (!c->graph || c->graph_steps != gsteps) {
            if (c->graph) {
                (void)hipGraphExecDestroy(c->graph);
                c->graph = nullptr;
            }
            int rc = build_graph(c, gsteps);
            if (rc)
                return rc;
        }
        for (; done + gsteps <= nsteps; done += gsteps)
            HIP_TRY(hipGraphLaunch(c->graph, c->stream));
    }
    for (; done < nsteps; done++) {
        int rc = enqueue_step(c, c->stream);
        if (rc)
            return rc;
        if (c->timing && c->ev.size() > 8192) {
            rc = fold_timing(c);
            if (rc)
                return rc;
        }
    }
    // phy_time += dt per step, as step_rk4 does (nuosc.hpp:347)
    for (int k = 0; k < nsteps; k++)
        c->phy_time += c->dt;
    c->steps += nsteps;
    return 0;
}

int cosenu_sync(cosenu_ctx *c)
{
    if (!c)
        return fail(COSENU_EINVAL, "cosenu_sync: null context");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int cosenu_rhs(cosenu_ctx *c, double *const out[8], ptrdiff_t stride, int off)
{
    if (!c || !out || !valid_layout(c, stride, off))
        return fail(COSENU_EINVAL, "cosenu_rhs: bad argument");
    if ((size_t)stride * c->nvz > c->n_state)
        return fail(COSENU_EINVAL, "cosenu_rhs: row_stride too large");
    HIP_TRY(hipSetDevice(c->p.device));
    RhsArgs r = rhs_args(c, c->y, c->B);
    if (c->p.mode == COSENU_FAST)
        HIP_TRY(launch_rhs_fast(r, c->z_chunk, c->stream));
    else
        HIP_TRY(launch_rhs_strict(r, c->stream));
    const size_t w = (size_t)c->nz * sizeof(double), pitch = (size_t)stride * sizeof(double);
    for (int f = 0; f < 8; f++) {
        HIP_TRY(launch_export_plane(c->A, c->B, f, c->nz, c->nvz, c->VP, stride, off, c->stream));
        HIP_TRY(hipMemcpy2DAsync(out[f] + off, pitch, c->A + off, pitch, w, c->nvz,
                                 hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return 0;
}

int cosenu_diagnostics(cosenu_ctx *c, cosenu_diag *d)
{
    if (!c || !d)
        return fail(COSENU_EINVAL, "cosenu_diagnostics: null argument");
    if (!c->have_profile || !c->have_state0)
        return fail(COSENU_ESTATE, "cosenu_diagnostics: set the profile and the t=0 state first");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(launch_diagnostics(c->y, c->y0, c->prof, c->nz, c->nvz, c->VP, c->diag_part, c->npart,
                               c->diag_out, c->stream));
    double h[kDiagSlots];
    HIP_TRY(hipMemcpyAsync(h, c->diag_out, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const double dzdv = c->dz * c->dv;
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

int cosenu_survival_profile(cosenu_ctx *c, double *pee, double *pbee)
{
    if (!c || !pee || !pbee)
        return fail(COSENU_EINVAL, "cosenu_survival_profile: null argument");
    if (!c->have_state0)
        return fail(COSENU_ESTATE, "cosenu_survival_profile: set the t=0 state first");
    HIP_TRY(hipSetDevice(c->p.device));
    HIP_TRY(launch_survival_profile(c->y, c->y0, c->nz, c->nvz, c->VP, c->diag_part, c->npart,
                                    c->diag_out, c->stream));
    HIP_TRY(hipMemcpyAsync(pee, c->diag_out, c->nvz * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipMemcpyAsync(pbee, c->diag_out + c->VP, c->nvz * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
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
    if (!c->snap) {
        int rc = dev_alloc(c, &c->snap, n);
        if (rc)
            return rc;
    }
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
    int rc = fold_timing(c);
    if (rc)
        return rc;
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
    int rc = fold_timing(c);
    if (rc)
        return rc;
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
    return c->p.mode == COSENU_FAST ? stage_fast_kernel_name(c->VP, stage)
                                    : stage_strict_kernel_name(stage);
}

} // extern "C"
