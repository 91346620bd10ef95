// FAST mode stage kernels: one fused kernel per RK4 stage that streams the
// state through registers.
//
// Work decomposition.  A team of W wavefronts covers the VP = 64*W velocity
// bins (lane = bin).  Each team owns a chunk of L consecutive z cells and
// marches along it: at iteration j every lane holds, per field, the window
// u(j-3) .. u(j+4) of its bin in registers, reconstructs the upwind WENO7
// value at the interface j+1/2, and updates cell j with the flux difference
// against the interface j-1/2 kept from the previous iteration.  Each z row
// of the stage input is read from HBM once per chunk (plus an 8-row halo),
// every output once.
//
// Upwind reconstruction.  The reference evaluates both biased
// reconstructions everywhere and selects one by sign(v) (rhs_fv.hpp:153-161).
// Only one is needed per bin: for v > 0 the left-biased value at j+1/2 from
// u(j-3..j+3) (rflux[j]); for v < 0 the right-biased value at j+1/2 from
// u(j-2..j+4) (lflux[j+1]), which is the same formula applied to the
// mirrored stencil u(j+4), u(j+3), ..., u(j-2) (the smoothness indicators of
// rhs_fv.hpp:49-52 are mirror images of each other).  The nonlinear weights
// gamma_k/(eps+beta_k)^2 are normalised with a single division:
//   sum_k w_k c_k = sum_k g_k P_k c_k / sum_k g_k P_k,  P_k = prod_{m!=k} (eps+beta_m)^2.
//
// Collective term.  sum_k mu (1 - v v_k) f_k dv = mu dv (F0 - v F1) with the
// moments F0 = sum_k f_k and F1 = sum_k v_k f_k of S = exi+bexi,
// R = exr-bexr, D = ee-xx-bee+bxx (rhs_fv.hpp:188-233 summed analytically).
// The 6 moments of z row j+4 are reduced across the wavefront when that row
// enters the window and parked in an LDS ring of 8 rows; cell j reads them 4
// iterations later, so the cross-wave exchange never waits on fresh data.
#include "fast_math.hpp"
#include <cstdio>
#include <cstdlib>

namespace cosenu {
namespace {

using namespace fastk;

constexpr int kWin = 8; // u(j-3) .. u(j+4)

// ---------------------------------------------------------------------------
// Split-lane variant.  The 8 fields of a velocity bin are shared by two lanes
// of one wavefront: lanes 0-31 carry (ee, xx, ex_re, ex_im) of 32 bins and
// lanes 32-63 carry (bee, bxx, bex_re, bex_im) of the same bins.  The nu and
// nubar halves of the right-hand side couple only through the per-z moments,
// which travel through the LDS ring anyway, so every lane needs a window of
// 4 fields (64 VGPRs) instead of 8.  A team is VP/32 wavefronts; with nvz a
// multiple of 64 every wavefront's bins share one sign of v, which selects the
// upwind stencil order with a wave-uniform branch instead of per-lane selects.
//
// With sigma = +1 for the nu lanes and -1 for the nubar lanes, and
// (p, q, r, s) = (ee, xx, ex_re, ex_im) or (bee, bxx, bex_re, bex_im), the
// local terms of rhs_fv.hpp:164-232 read
//   K_p = -2 st s + 2 (r S - sigma s R)           K_q = -K_p (local part)
//   K_r = -2 ct s + sigma h s + (q - p) S + sigma s D
//   K_s =  2 ct r + st (p - q) - sigma h r + sigma (p - q) R - sigma r D
// with S, R, D the collective integrals mu dv (F0 - v F1) and h = Hm(z).

// One team's march over the chunk [j0, j0+L) of the run described by A.
// `tring` is the team's LDS ring.  Teams of a block must call this with the
// same L (they meet at one barrier per z step).  OPENB selects the open
// boundary of nuosc.hpp:247-277: rows below 0 copy row 0, rows above nz-1 are
// extrapolated linearly one after another, and the flux ghosts follow the
// same rule (rflux[-1] = rflux[0]; lflux[nz] = 2 lflux[nz-1] - lflux[nz-2]).
template <int NW, int SCHEME, int STAGE, bool OPENB>
__device__ __forceinline__ void fast2_body(const StageArgs &A, int L, int j0, int jend,
                                           bool active, double (*tring)[NW][8])
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wit = wave % NW;
    const int h = lane >> 5;
    const int v = wit * 32 + (lane & 31);
    const int VP = A.VP, nz = A.nz;
    const size_t zs = (size_t)8 * VP;
    const double sig = h ? -1.0 : 1.0;
    // z positions are the same in every lane of a wavefront: keep them scalar
    j0 = __builtin_amdgcn_readfirstlane(j0);
    jend = __builtin_amdgcn_readfirstlane(jend);

    const double vi = A.vz[v];
    const bool pos = vi > 0.0;
    // wave-uniform upwind direction: +1 / -1 when all live bins agree
    const bool anyp = __any(vi > 0.0), anyn = __any(vi < 0.0);
    const int dir = anyn ? (anyp ? 0 : -1) : 1;
    const bool store = active && v < A.nvz;
    const double *in = A.yin + (size_t)(4 * h) * VP + v; // this lane's first field
    // The team updates cells [j0, jend); the loop below runs L + 1 iterations
    // in every team of a block (jend - j0 <= L) so that all waves meet at
    // every barrier.

    auto zrow = [&](int m) -> int {
        if (OPENB)
            return m;
        int w = m < 0 ? m + nz : m; // m >= -4 > -nz
        while (w >= nz)
            w -= nz;
        return w;
    };
    // the 4 fields of this lane at z row m (m is wave-uniform)
    auto load4 = [&](int m, double o[4]) {
        if (!OPENB || (m >= 0 && m < nz)) {
            const double *r = in + (size_t)zrow(m) * zs;
#pragma unroll
            for (int f = 0; f < 4; f++)
                o[f] = r[(size_t)f * VP];
        } else if (m < 0) {
#pragma unroll
            for (int f = 0; f < 4; f++)
                o[f] = in[(size_t)f * VP];
        } else {
            const double *r1 = in + (size_t)(nz - 2) * zs, *r2 = in + (size_t)(nz - 1) * zs;
#pragma unroll
            for (int f = 0; f < 4; f++) {
                double a = r1[(size_t)f * VP], b = r2[(size_t)f * VP];
                for (int k = nz; k <= m; k++) {
                    const double c = b * 2 - a;
                    a = b;
                    b = c;
                }
                o[f] = b;
            }
        }
    };

    double win[4][kWin];
#pragma unroll
    for (int kk = 0; kk < kWin; kk++) {
        double t4[4];
        load4(j0 - 4 + kk, t4);
#pragma unroll
        for (int f = 0; f < 4; f++)
            win[f][kk] = t4[f];
    }

    const bool coll = (A.flags & kFlagColl) != 0;
    // Padding bins (v >= nvz) hold whatever a staging copy left in the stage
    // buffers and are never written: keep them out of the moments.
    const bool live = v < A.nvz;
    auto moments_to_ring = [&](int kk, int zabs) {
        const double S = live ? win[3][kk] : 0.0;
        const double R = live ? sig * win[2][kk] : 0.0;
        const double D = live ? sig * (win[0][kk] - win[1][kk]) : 0.0;
        const double m6[6] = {S, vi * S, R, vi * R, D, vi * D};
        const double t = wave_sum6(m6, lane);
        const int idx = (kMomentLane >> (4 * (lane >> 3))) & 15;
        if ((lane & 7) == 0 && idx < 6)
            tring[zabs & 7][wit][idx] = t;
    };
    if (coll) {
        moments_to_ring(4, j0);
        moments_to_ring(5, j0 + 1);
        moments_to_ring(6, j0 + 2);
    }

    const double fac = -vi / A.dz;
    const double fdfac = -vi / (12.0 * A.dz);
    const bool ko5 = (A.flags & kFlagKO5) != 0;
    const double ko_eps = ko5 ? -A.ko / A.dz / 64.0 : -A.ko / A.dz / 16.0;
    const double mudv = A.mu * A.dv;
    const double st = A.st * A.pmo, ct = A.ct * A.pmo;
    const bool vac = (A.flags & kFlagVac) != 0, mat = (A.flags & kFlagMat) != 0;
    const bool advect = !(A.flags & kFlagAdvecOff);

    double Fprev[4] = {0.0, 0.0, 0.0, 0.0};
    double Fprev2[4] = {0.0, 0.0, 0.0, 0.0}; // open boundary only

    for (int j = j0 - 1; j < j0 + L; j++) {
        double nxt[4];
        load4(j + 5, nxt);
        const bool cell = j >= j0 && j < jend;
        const size_t q = (size_t)zrow(j) * zs + (size_t)(4 * h) * VP + v;
        // stage-value form: y (not at stage 1, where y is the stage input and
        // sits in the window), and Y2, Y3 at stage 4
        double yv[4], y2v[4], y3v[4];
        if (cell && (STAGE == 2 || STAGE == 4)) {
#pragma unroll
            for (int f = 0; f < 4; f++)
                yv[f] = A.y[q + (size_t)f * VP];
        }
        if (cell && STAGE == 4) {
#pragma unroll
            for (int f = 0; f < 4; f++) {
                y2v[f] = A.y2[q + (size_t)f * VP];
                y3v[f] = A.y3[q + (size_t)f * VP];
            }
        }
        if (coll)
            moments_to_ring(7, j + 4);

        double F[4] = {0.0, 0.0, 0.0, 0.0};
        if (SCHEME == 0 && advect) {
            if (dir > 0) {
#pragma unroll
                for (int f = 0; f < 4; f++)
                    F[f] = flux_dir<1>(win[f], pos);
            } else if (dir < 0) {
#pragma unroll
                for (int f = 0; f < 4; f++)
                    F[f] = flux_dir<-1>(win[f], pos);
            } else {
#pragma unroll
                for (int f = 0; f < 4; f++)
                    F[f] = flux_dir<0>(win[f], pos);
            }
        }

        __syncthreads();

        if (cell) {
            const double p = win[0][3], qq = win[1][3], r = win[2][3], s = win[3][3];
            double K[4] = {0.0, 0.0, 0.0, 0.0};
            if (advect) {
                if (SCHEME == 0) {
#pragma unroll
                    for (int f = 0; f < 4; f++) {
                        double Fr = F[f], Fl = Fprev[f];
                        if (OPENB && j == 0 && pos)
                            Fl = Fr; // rflux[-1] = rflux[0]
                        if (OPENB && j == nz - 1 && !pos)
                            Fr = Fprev[f] * 2 - Fprev2[f]; // lflux[nz] extrapolated
                        K[f] = fac * (Fr - Fl);
                    }
                } else {
#pragma unroll
                    for (int f = 0; f < 4; f++) {
                        const double *w = win[f];
                        K[f] = fdfac * ((w[1] - w[5]) - 8.0 * (w[2] - w[4]));
                    }
                }
            }
            if (vac) {
                K[0] += -2.0 * st * s;
                K[1] += 2.0 * st * s;
                K[2] += -2.0 * ct * s;
                K[3] += 2.0 * ct * r + st * (p - qq);
            }
            if (mat) {
                const double hm = sig * A.Hm[zrow(j)];
                K[2] += hm * s;
                K[3] -= hm * r;
            }
            if (coll) {
                double M[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
                for (int w = 0; w < NW; w++) {
                    const double *slot = tring[j & 7][w];
#pragma unroll
                    for (int m = 0; m < 6; m++)
                        M[m] += slot[m];
                }
                const double S = mudv * (M[0] - vi * M[1]);
                const double R = mudv * (M[2] - vi * M[3]);
                const double D = mudv * (M[4] - vi * M[5]);
                const double I1 = 2.0 * (r * S - sig * s * R);
                K[0] += I1;
                K[1] -= I1;
                K[2] += (qq - p) * S + sig * s * D;
                K[3] += sig * ((p - qq) * R - r * D);
            }
            if (SCHEME == 1) {
#pragma unroll
                for (int f = 0; f < 4; f++) {
                    const double *w = win[f]; // w[k] = u(j-3+k)
                    if (ko5)
                        K[f] += ko_eps * ((w[0] + w[6]) - 6.0 * (w[1] + w[5]) + 15.0 * (w[2] + w[4]) -
                                          20.0 * w[3]);
                    else
                        K[f] += ko_eps * ((w[1] + w[5]) - 4.0 * (w[2] + w[4]) + 6.0 * w[3]);
                }
            }
            if (store) {
#pragma unroll
                for (int f = 0; f < 4; f++) {
                    const size_t o = q + (size_t)f * VP;
                    if (STAGE == 0) {
                        A.yout[o] = K[f];
                    } else if (STAGE == 1) {
                        A.yout[o] = win[f][3] + A.a * K[f];
                    } else if (STAGE == 4) {
                        const double y0 = yv[f];
                        const double d = ((y2v[f] - y0) + 2.0 * (y3v[f] - y0)) + (win[f][3] - y0);
                        A.yout[o] = y0 + (d * (1.0 / 3.0) + A.a * K[f]);
                    } else {
                        A.yout[o] = yv[f] + A.a * K[f];
                    }
                }
            }
        }

#pragma unroll
        for (int f = 0; f < 4; f++) {
#pragma unroll
            for (int kk = 0; kk < kWin - 1; kk++)
                win[f][kk] = win[f][kk + 1];
            win[f][kWin - 1] = nxt[f];
            if (OPENB)
                Fprev2[f] = Fprev[f];
            Fprev[f] = F[f];
        }
    }
}

// nteams > 0: team g of nteams updates the contiguous range
// [g*nz/nteams, (g+1)*nz/nteams) (L = the longest range); nteams == 0: team g
// updates chunk g of length L (nchunks chunks).
template <int NW, int SCHEME, int STAGE, bool OPENB>
__global__ __launch_bounds__(NW >= 8 ? 512 : 256) void cosenu_fast2_stage(StageArgs A, int L,
                                                                           int nchunks, int nteams)
{
    constexpr int TPB = NW >= 4 ? 1 : 4 / NW; // teams per block
    __shared__ double ring[TPB][8][NW][8];
    const int team = (threadIdx.x >> 6) / NW;
    const int g = xcd_block_order() * TPB + team;
    if (nteams > 0) {
        const bool active = g < nteams;
        const int j0 = active ? (int)(((long)g * A.nz) / nteams) : 0;
        const int j1 = active ? (int)(((long)(g + 1) * A.nz) / nteams) : 0;
        fast2_body<NW, SCHEME, STAGE, OPENB>(A, L, j0, j1, active, ring[team]);
    } else {
        const bool active = g < nchunks;
        const int j0 = (active ? g : 0) * L;
        fast2_body<NW, SCHEME, STAGE, OPENB>(A, L, j0, j0 + L, active, ring[team]);
    }
}

// Batched launch: work item k = (run, first z cell); every item of one launch
// has the same chunk length L and the same padded width VP.
template <int NW, int SCHEME, int STAGE>
__global__ __launch_bounds__(NW >= 8 ? 512 : 256) void cosenu_fast2_batch(const BatchRun *runs,
                                                                           const int2 *items,
                                                                           int nitems, int L,
                                                                           int stage)
{
    constexpr int TPB = NW >= 4 ? 1 : 4 / NW;
    __shared__ double ring[TPB][8][NW][8];
    const int team = (threadIdx.x >> 6) / NW;
    const int item = xcd_block_order() * TPB + team;
    const bool active = item < nitems;
    const int2 it = items[active ? item : 0];
    const StageArgs A = batch_stage_args(runs[it.x], stage, true);
    // a run's last item may be shorter than L (its nz need not be a multiple)
    fast2_body<NW, SCHEME, STAGE, false>(A, L, it.y, min(it.y + L, A.nz), active, ring[team]);
}

template <int NW, int SCHEME, bool OPENB>
hipError_t launch_nw(const StageArgs &a, int stage, int L, int nteams, hipStream_t s)
{
    constexpr int TPB = NW >= 4 ? 1 : 4 / NW;
    int nchunks = 0, work;
    if (nteams > 0) {
        L = (a.nz + nteams - 1) / nteams;
        work = nteams;
    } else {
        nchunks = a.nz / L;
        work = nchunks;
    }
    dim3 grid((work + TPB - 1) / TPB), block(64 * NW * TPB);
    switch (stage) {
    case 0: hipLaunchKernelGGL((cosenu_fast2_stage<NW, SCHEME, 0, OPENB>), grid, block, 0, s, a, L, nchunks, nteams); break;
    case 1: hipLaunchKernelGGL((cosenu_fast2_stage<NW, SCHEME, 1, OPENB>), grid, block, 0, s, a, L, nchunks, nteams); break;
    case 2:
    case 3: // stages 2 and 3 run the same code with their own coefficient
        hipLaunchKernelGGL((cosenu_fast2_stage<NW, SCHEME, 2, OPENB>), grid, block, 0, s, a, L, nchunks, nteams); break;
    case 4: hipLaunchKernelGGL((cosenu_fast2_stage<NW, SCHEME, 4, OPENB>), grid, block, 0, s, a, L, nchunks, nteams); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Teams of the split-lane kernel that fit on the device at once (one
// instantiation stands for all: the stage variants differ by a few VGPRs).
template <int NW, int SCHEME, bool OPENB>
int resident_teams_nw()
{
    constexpr int TPB = NW >= 4 ? 1 : 4 / NW;
    int nb = 0, dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(&cosenu_fast2_stage<NW, SCHEME, 2, OPENB>),
            64 * NW * TPB, 0) != hipSuccess)
        return 0;
    return nb * prop.multiProcessorCount * TPB;
}

template <int SCHEME, bool OPENB>
hipError_t launch_scheme2(const StageArgs &a, int stage, int L, int nteams, hipStream_t s)
{
    switch (a.VP) {
    case 64: return launch_nw<2, SCHEME, OPENB>(a, stage, L, nteams, s);
    case 128: return launch_nw<4, SCHEME, OPENB>(a, stage, L, nteams, s);
    case 256: return launch_nw<8, SCHEME, OPENB>(a, stage, L, nteams, s);
    default: return hipErrorInvalidValue;
    }
}

template <int SCHEME, bool OPENB>
int resident_teams_scheme(int VP)
{
    switch (VP) {
    case 64: return resident_teams_nw<2, SCHEME, OPENB>();
    case 128: return resident_teams_nw<4, SCHEME, OPENB>();
    case 256: return resident_teams_nw<8, SCHEME, OPENB>();
    default: return 0;
    }
}

} // namespace

int fast_supported(int VP, int bc) { return (VP == 64 || VP == 128 || VP == 256) && (bc == 0 || bc == 1); }

hipError_t launch_stage_fast(const StageArgs &a, int stage, int L, int nteams, hipStream_t s)
{
    if (!fast_supported(a.VP, a.bc))
        return hipErrorInvalidValue;
    if (nteams > 0) {
        if (nteams > a.nz)
            return hipErrorInvalidValue;
    } else if (L < 1 || a.nz % L != 0) {
        return hipErrorInvalidValue;
    }
    if (a.bc == 1) // the open boundary's instances
        return a.scheme == 0 ? launch_scheme2<0, true>(a, stage, L, nteams, s)
                             : launch_scheme2<1, true>(a, stage, L, nteams, s);
    return a.scheme == 0 ? launch_scheme2<0, false>(a, stage, L, nteams, s)
                         : launch_scheme2<1, false>(a, stage, L, nteams, s);
}

int fast_resident_teams(int VP, int scheme, int bc)
{
    if (bc == 1)
        return scheme == 0 ? resident_teams_scheme<0, true>(VP) : resident_teams_scheme<1, true>(VP);
    return scheme == 0 ? resident_teams_scheme<0, false>(VP) : resident_teams_scheme<1, false>(VP);
}

hipError_t launch_rhs_fast(const RhsArgs &r, int L, int nteams, hipStream_t s)
{
    StageArgs a;
    a.yin = r.in; a.y = nullptr; a.acc = nullptr; a.yout = r.out;
    a.y2 = nullptr; a.y3 = nullptr; a.mom_in = nullptr; a.mom_out = nullptr;
    a.vz = r.vz; a.Hm = r.Hm;
    a.nz = r.nz; a.nvz = r.nvz; a.VP = r.VP; a.flags = r.flags; a.bc = r.bc; a.scheme = r.scheme;
    a.a = 0.0;
    a.dz = r.dz; a.dv = r.dv; a.ct = r.ct; a.st = r.st; a.pmo = r.pmo; a.mu = r.mu; a.ko = r.ko;
    return launch_stage_fast(a, 0, L, nteams, s);
}

namespace {
template <int NW, int SCHEME>
hipError_t launch_batch_nw(const BatchRun *runs, const int2 *items, int nitems, int L, int stage,
                           hipStream_t s)
{
    constexpr int TPB = NW >= 4 ? 1 : 4 / NW;
    dim3 grid((nitems + TPB - 1) / TPB), block(64 * NW * TPB);
    if (nitems <= 0)
        return hipSuccess;
    switch (stage) {
    case 1: hipLaunchKernelGGL((cosenu_fast2_batch<NW, SCHEME, 1>), grid, block, 0, s, runs, items, nitems, L, stage); break;
    case 2:
    case 3: hipLaunchKernelGGL((cosenu_fast2_batch<NW, SCHEME, 2>), grid, block, 0, s, runs, items, nitems, L, stage); break;
    case 4: hipLaunchKernelGGL((cosenu_fast2_batch<NW, SCHEME, 4>), grid, block, 0, s, runs, items, nitems, L, stage); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int SCHEME>
hipError_t launch_batch_scheme(const BatchRun *runs, const int2 *items, int nitems, int VP, int L,
                               int stage, hipStream_t s)
{
    switch (VP) {
    case 64: return launch_batch_nw<2, SCHEME>(runs, items, nitems, L, stage, s);
    case 128: return launch_batch_nw<4, SCHEME>(runs, items, nitems, L, stage, s);
    case 256: return launch_batch_nw<8, SCHEME>(runs, items, nitems, L, stage, s);
    default: return hipErrorInvalidValue;
    }
}
} // namespace

hipError_t launch_stage_fast_batch(const BatchRun *runs, const int2 *items, int nitems, int VP,
                                   int scheme, int L, int stage, hipStream_t s)
{
    if (!fast_supported(VP, 0) || L < 1)
        return hipErrorInvalidValue;
    return scheme == 0 ? launch_batch_scheme<0>(runs, items, nitems, VP, L, stage, s)
                       : launch_batch_scheme<1>(runs, items, nitems, VP, L, stage, s);
}

const char *stage_fast_kernel_name(int VP, int stage)
{
    static char buf[96];
    const int st = stage + 1 == 3 ? 2 : stage + 1;
    snprintf(buf, sizeof(buf), "cosenu_fast2_stage<%d, *, %d>", VP / 32, st);
    return buf;
}

} // namespace cosenu
