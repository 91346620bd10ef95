// FAST mode, barrier-free variant (variant 3): every wavefront is on its own.
//
// Work decomposition.  A wavefront owns 32 velocity bins (lanes 0-31 carry
// ee, xx, ex_re, ex_im of the bins, lanes 32-63 bee, bxx, bex_re, bex_im of
// the same bins) over a range of z cells, and marches along z holding, per
// field, a window of 8 rows in registers.  Nothing is shared between
// wavefronts inside a launch, so there is no barrier and no LDS.
//
// March direction.  A wavefront's bins have one sign of v, and it marches in
// the upwind direction: z increasing for v > 0, decreasing for v < 0.  In
// march order the upwind WENO7 stencil of the downstream face of the current
// cell is then always the 7 rows ending 3 rows ahead (rhs_fv.hpp:153-161
// select rflux for v > 0 and lflux for v < 0, the mirror image of each
// other), and the flux difference of the cell is -|v|/dz (F_downstream -
// F_upstream) in both directions; the FD derivative (rhs_fd.hpp:56-70) flips
// sign with the march direction the same way and its dissipation is
// symmetric.  One code path serves both signs.
//
// Collective term.  mu dv sum_k (1 - v v_k) f_k = mu dv (F0 - v F1) with the
// moments F0 = sum_k f_k, F1 = sum_k v_k f_k of S = exi + bexi, R = exr -
// bexr, D = ee - xx - bee + bxx (rhs_fv.hpp:188-233 summed analytically).
// The launch that writes a stage state also writes, per z row and bin group,
// the 6 partial moments of what it wrote (one wavefront reduction per row);
// the next launch adds the partials of the row's groups.  The initial state's
// partials come from cosenu_stream_moments.
//
// The window is a ring of 8 register slots: window position k = 0..6 (row
// r + s(k - 3), r the current cell, s the march direction) lives in slot
// (phase + k) & 7, and slot (phase + 7) & 7 receives the row that becomes
// position 6 at the next iteration.  The march loop is unrolled over the 8
// phases, so rows enter the ring without being shifted through it.
#include "fast_math.hpp"

#include <cstdio>
#include <type_traits>

namespace cosenu {
namespace {

using namespace fastk;

// 6 partial moments of the wavefront's 32 bins (fields p, q, r, s of this
// lane; sig = +1 for the nu lanes, -1 for the nubar lanes) into slot[0..5].
__device__ __forceinline__ void store_moments(double *slot, int lane, double vi, double sig,
                                              double p, double q, double r, double s)
{
    const double S = s, R = sig * r, D = sig * (p - q);
    const double m6[6] = {S, vi * S, R, vi * R, D, vi * D};
    const double t = wave_sum6(m6, lane);
    const int idx = (kMomentLane >> (4 * (lane >> 3))) & 15;
    if ((lane & 7) == 0 && idx < 6)
        slot[idx] = t;
}

// One wavefront per (z row, bin group); padding bins (v >= nvz) count zero.
__global__ __launch_bounds__(256) void cosenu_stream_moments(const double *y, const double *vz,
                                                             double *mom, int nz, int nvz, int VP)
{
    const int lane = threadIdx.x & 63;
    const int G = VP >> 5;
    const long w = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= (long)nz * G)
        return; // whole wavefronts
    const int r = (int)(w / G), g = (int)(w % G);
    const int h = lane >> 5, v = g * 32 + (lane & 31);
    const double *p = y + (size_t)r * 8 * VP + (size_t)(4 * h) * VP + v;
    const bool live = v < nvz;
    store_moments(mom + ((size_t)r * G + g) * kMomSlot, lane, vz[v], h ? -1.0 : 1.0,
                  live ? p[0] : 0.0, live ? p[VP] : 0.0, live ? p[2 * VP] : 0.0,
                  live ? p[3 * VP] : 0.0);
}

// OCC: wavefronts per SIMD the register allocation is held to.
template <int SCHEME, int STAGE, int OCC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC)))
void cosenu_stream_stage(StageArgs A, int L, int nwork, int nteams)
{
    const int lane = threadIdx.x & 63;
    const int VP = A.VP, nz = A.nz;
    const int G = VP >> 5;
    const int wv = __builtin_amdgcn_readfirstlane(xcd_block_order() * 4 + (threadIdx.x >> 6));
    const int g = wv % G, c = wv / G;
    int j0, j1;
    if (nteams > 0) { // contiguous range c of nteams
        if (c >= nteams)
            return;
        j0 = (int)(((long)c * nz) / nteams);
        j1 = (int)(((long)(c + 1) * nz) / nteams);
    } else { // chunk c of length L
        if (c >= nwork)
            return;
        j0 = c * L;
        j1 = j0 + L;
    }
    const int h = lane >> 5;
    const int v = g * 32 + (lane & 31);
    const size_t zs = (size_t)8 * VP;
    const int loff = 4 * h * VP + v; // this lane's first field within a z row
    const double sig = h ? -1.0 : 1.0;
    const double vi = A.vz[v];
    const int s = __any(vi > 0.0) ? 1 : -1; // march direction (one sign per group)
    const int n = j1 - j0;
    const int rfirst = s > 0 ? j0 : j1 - 1;
    // FV: iteration -1 reconstructs the upstream face of the first cell
    const int mstart = SCHEME == 0 ? -1 : 0;

    const int flags = A.flags;
    const bool coll = (flags & kFlagColl) != 0;
    const bool vac = (flags & kFlagVac) != 0, mat = (flags & kFlagMat) != 0;
    const bool advect = !(flags & kFlagAdvecOff);
    const bool ko5 = (flags & kFlagKO5) != 0;
    const double adv = -fabs(vi) / A.dz;          // FV: times (F_down - F_up)
    const double adv_fd = -fabs(vi) / (12.0 * A.dz); // FD: times the march-order stencil
    const double ko_eps = ko5 ? -A.ko / A.dz / 64.0 : -A.ko / A.dz / 16.0;
    const double mudv = A.mu * A.dv;
    const double st = A.st * A.pmo, ct = A.ct * A.pmo;

    auto wrap = [&](int r) -> int { return r < 0 ? r + nz : (r >= nz ? r - nz : r); };
    // Row addresses are wave-uniform (scalar base) plus this lane's 32-bit
    // offset, so the loads and stores take no per-lane address arithmetic.
    int off[4];
#pragma unroll
    for (int f = 0; f < 4; f++)
        off[f] = loff + f * VP;
    auto load4 = [&](int r, double o[4]) {
        const double *rp = A.yin + (size_t)wrap(r) * zs;
#pragma unroll
        for (int f = 0; f < 4; f++)
            o[f] = rp[off[f]];
    };
    // Partial moments of a row: lane gg*8 + i holds moment i of group gg.
    // More than 8 groups (VP > 256): lane gg*8 + i holds the sum of moment i
    // over groups gg, gg + 8, gg + 16, ... in that order.
    auto load_mom = [&](int r) -> double {
        const double *row = A.mom_in + (size_t)r * G * kMomSlot;
        if (G <= 8)
            return lane < G * kMomSlot ? row[lane] : 0.0;
        double t = row[lane];
        for (int c = 64; c < G * kMomSlot; c += 64)
            if (c + lane < G * kMomSlot) // G need not be a multiple of 8
                t += row[c + lane];
        return t;
    };
    // Totals over the groups (fixed order), moment i in lane i.
    auto sum_mom = [&](double t) -> double {
        t += dpp<0x128>(t); // row_ror:8 pairs groups 2k and 2k+1
        if (G > 2) {
            double u = t;
            swap16(t, u);
            t += u;
        }
        if (G > 4) {
            double u = t;
            swap32(t, u);
            t += u;
        }
        return t;
    };
    auto lane_value = [](double x, int l) -> double {
        return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l),
                                __builtin_amdgcn_readlane(__double2loint(x), l));
    };

    double w[4][8]; // window ring: position k in slot (phase + k) & 7
#pragma unroll
    for (int k = 0; k < 7; k++) {
        double t[4];
        load4(rfirst + s * (mstart + k - 3), t);
#pragma unroll
        for (int f = 0; f < 4; f++)
            w[f][k] = t[f];
    }
    double Fin[4] = {0.0, 0.0, 0.0, 0.0}; // FV: flux at the upstream face of the cell
    // partial moments of the next cell row, loaded one iteration ahead
    double mnext = coll ? load_mom(rfirst) : 0.0;
    double hnext = mat ? A.Hm[rfirst] : 0.0; // matter potential of the next cell row

    auto body = [&](auto P, int m) {
        constexpr int p = decltype(P)::value;
        constexpr int s0 = p & 7, s1 = (p + 1) & 7, s2 = (p + 2) & 7, s3 = (p + 3) & 7,
                      s4 = (p + 4) & 7, s5 = (p + 5) & 7, s6 = (p + 6) & 7;
        const bool more = m + 1 < n;
        if (more) { // position 6 of the next iteration, into the free slot
            double t[4];
            load4(rfirst + s * (m + 4), t);
#pragma unroll
            for (int f = 0; f < 4; f++)
                w[f][(p + 7) & 7] = t[f];
        }
        const bool cell = m >= 0;
        const int r = rfirst + s * m;
        const size_t q = (size_t)r * zs;
        double yv[4], y2v[4], y3v[4];
        double mcur = 0.0, hcur = 0.0;
        if (cell) {
            if (STAGE == 2 || STAGE == 4) {
                const double *yp = A.y + q;
#pragma unroll
                for (int f = 0; f < 4; f++)
                    yv[f] = yp[off[f]];
            }
            if (STAGE == 4) {
                const double *p2 = A.y2 + q, *p3 = A.y3 + q;
#pragma unroll
                for (int f = 0; f < 4; f++) {
                    y2v[f] = p2[off[f]];
                    y3v[f] = p3[off[f]];
                }
            }
            if (coll) {
                mcur = mnext;
                if (more)
                    mnext = load_mom(r + s);
            }
            if (mat) {
                hcur = hnext;
                if (more)
                    hnext = A.Hm[r + s];
            }
        }
        double K[4] = {0.0, 0.0, 0.0, 0.0};
        if (SCHEME == 0 && advect) {
            // downstream face; the flux difference is formed at once so that
            // only one face per field stays live
#pragma unroll
            for (int f = 0; f < 4; f++) {
                const double Fo = weno_up(w[f][s0], w[f][s1], w[f][s2], w[f][s3], w[f][s4],
                                          w[f][s5], w[f][s6]);
                K[f] = adv * (Fo - Fin[f]);
                Fin[f] = Fo;
            }
        }
        if (cell) {
            const double pe = w[0][s3], qe = w[1][s3], re = w[2][s3], se = w[3][s3];
            if (advect) {
                if (SCHEME == 1) {
#pragma unroll
                    for (int f = 0; f < 4; f++)
                        K[f] = adv_fd * ((w[f][s1] - w[f][s5]) - 8.0 * (w[f][s2] - w[f][s4]));
                }
            }
            if (vac) {
                K[0] += -2.0 * st * se;
                K[1] += 2.0 * st * se;
                K[2] += -2.0 * ct * se;
                K[3] += 2.0 * ct * re + st * (pe - qe);
            }
            if (mat) {
                const double hm = sig * hcur;
                K[2] += hm * se;
                K[3] -= hm * re;
            }
            if (coll) {
                const double t = sum_mom(mcur);
                double M[6];
#pragma unroll
                for (int i = 0; i < 6; i++)
                    M[i] = lane_value(t, i);
                const double S = mudv * (M[0] - vi * M[1]);
                const double R = mudv * (M[2] - vi * M[3]);
                const double D = mudv * (M[4] - vi * M[5]);
                const double I1 = 2.0 * (re * S - sig * se * R);
                K[0] += I1;
                K[1] -= I1;
                K[2] += (qe - pe) * S + sig * se * D;
                K[3] += sig * ((pe - qe) * R - re * D);
            }
            if (SCHEME == 1) {
#pragma unroll
                for (int f = 0; f < 4; f++) {
                    const double w0 = w[f][s0], w1 = w[f][s1], w2 = w[f][s2], w3 = w[f][s3],
                                 w4 = w[f][s4], w5 = w[f][s5], w6 = w[f][s6];
                    if (ko5)
                        K[f] += ko_eps * ((w0 + w6) - 6.0 * (w1 + w5) + 15.0 * (w2 + w4) -
                                          20.0 * w3);
                    else
                        K[f] += ko_eps * ((w1 + w5) - 4.0 * (w2 + w4) + 6.0 * w3);
                }
            }
            double o[4];
            double *op = A.yout + q;
#pragma unroll
            for (int f = 0; f < 4; f++) {
                if (STAGE == 0) {
                    o[f] = K[f];
                } else if (STAGE == 1) { // the stage input is y
                    o[f] = w[f][s3] + A.a * K[f];
                } else if (STAGE == 4) {
                    const double y0 = yv[f];
                    const double d = ((y2v[f] - y0) + 2.0 * (y3v[f] - y0)) + (w[f][s3] - y0);
                    o[f] = y0 + (d * (1.0 / 3.0) + A.a * K[f]);
                } else {
                    o[f] = yv[f] + A.a * K[f];
                }
                op[off[f]] = o[f];
            }
            if (STAGE != 0)
                store_moments(A.mom_out + ((size_t)r * G + g) * kMomSlot, lane, vi, sig, o[0],
                              o[1], o[2], o[3]);
        }
    };

    int m = mstart;
    for (;;) {
        if (m >= n) break;
        body(std::integral_constant<int, 0>{}, m++);
        if (m >= n) break;
        body(std::integral_constant<int, 1>{}, m++);
        if (m >= n) break;
        body(std::integral_constant<int, 2>{}, m++);
        if (m >= n) break;
        body(std::integral_constant<int, 3>{}, m++);
        if (m >= n) break;
        body(std::integral_constant<int, 4>{}, m++);
        if (m >= n) break;
        body(std::integral_constant<int, 5>{}, m++);
        if (m >= n) break;
        body(std::integral_constant<int, 6>{}, m++);
        if (m >= n) break;
        body(std::integral_constant<int, 7>{}, m++);
    }
}

// ---------------------------------------------------------------------------
// General form for the wide grids the kernel above does not take: bin groups
// that straddle v = 0, padding bins (nvz not a multiple of 64) and the open
// boundary.  Every wavefront marches up in z; its window holds rows r-3 ..
// r+4 of the current cell r, from which each lane reconstructs the face
// r+1/2 with its own upwind stencil (u(r-3..r+3) for v > 0, the mirror image
// u(r+4..r-2) for v < 0; a wave-uniform branch where the group has one sign)
// and forms K = -v/dz (F(r+1/2) - F(r-1/2)) (rhs_fv.hpp:153-161).  The open
// boundary (nuosc.hpp:247-277): rows below 0 copy row 0 and rows above nz-1
// extrapolate linearly, one after another, from the stage input; the flux
// ghosts follow the same rule (rflux[-1] = rflux[0], lflux[nz] = 2
// lflux[nz-1] - lflux[nz-2]).  Padding lanes contribute nothing to the
// moments and store nothing.  Window position k lives in ring slot
// (phase + k) & 7; the row that enters next is loaded into registers ahead
// of the compute and moved into the freed slot after it.
template <int SCHEME, int STAGE, bool OPENB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(STAGE == 4 ? 2 : 3)))
void cosenu_stream_general(StageArgs A, int L, int nwork, int nteams)
{
    const int lane = threadIdx.x & 63;
    const int VP = A.VP, nz = A.nz;
    const int G = VP >> 5;
    const int wv = __builtin_amdgcn_readfirstlane(xcd_block_order() * 4 + (threadIdx.x >> 6));
    const int g = wv % G, c = wv / G;
    int j0, j1;
    if (nteams > 0) {
        if (c >= nteams)
            return;
        j0 = (int)(((long)c * nz) / nteams);
        j1 = (int)(((long)(c + 1) * nz) / nteams);
    } else {
        if (c >= nwork)
            return;
        j0 = c * L;
        j1 = j0 + L;
    }
    const int h = lane >> 5;
    const int v = g * 32 + (lane & 31);
    const size_t zs = (size_t)8 * VP;
    const int loff = 4 * h * VP + v;
    const double sig = h ? -1.0 : 1.0;
    const double vi = A.vz[v];
    const bool live = v < A.nvz;
    const bool pos = vi > 0.0;
    // wave-uniform stencil order: +1 / -1 where every live bin has that sign
    const int dir = __all(!live || pos) ? 1 : (__all(!live || !pos) ? -1 : 0);
    const int n = j1 - j0;
    // FV: iteration -1 forms the face below the first cell
    const int mstart = SCHEME == 0 ? -1 : 0;

    const int flags = A.flags;
    const bool coll = (flags & kFlagColl) != 0;
    const bool vac = (flags & kFlagVac) != 0, mat = (flags & kFlagMat) != 0;
    const bool advect = !(flags & kFlagAdvecOff);
    const bool ko5 = (flags & kFlagKO5) != 0;
    const double fac = advect ? -vi / A.dz : 0.0;
    const double fdfac = advect ? -vi / (12.0 * A.dz) : 0.0;
    const double ko_eps = ko5 ? -A.ko / A.dz / 64.0 : -A.ko / A.dz / 16.0;
    const double mudv = A.mu * A.dv;
    const double st = A.st * A.pmo, ct = A.ct * A.pmo;

    int off[4];
#pragma unroll
    for (int f = 0; f < 4; f++)
        off[f] = loff + f * VP;
    // the 4 fields of this lane at row r of the stage input (r is wave-
    // uniform).  Periodic: wrapped.  Open: rows below 0 read row 0 (the
    // boundary rule); rows above nz-1 read row nz-1 as a placeholder and are
    // replaced by the linear extrapolation when they enter the window.
    auto load4 = [&](int r, double o[4]) {
        const int w = OPENB ? min(max(r, 0), nz - 1)
                            : (r < 0 ? r + nz : (r >= nz ? r - nz : r)); // -5 <= r < nz + 5, nz >= 8
        const double *rp = A.yin + (size_t)w * zs;
#pragma unroll
        for (int f = 0; f < 4; f++)
            o[f] = rp[off[f]];
    };
    auto load_mom = [&](int r) -> double {
        const double *row = A.mom_in + (size_t)r * G * kMomSlot;
        if (G <= 8)
            return lane < G * kMomSlot ? row[lane] : 0.0;
        double t = row[lane];
        for (int cc = 64; cc < G * kMomSlot; cc += 64)
            if (cc + lane < G * kMomSlot)
                t += row[cc + lane];
        return t;
    };
    auto sum_mom = [&](double t) -> double {
        t += dpp<0x128>(t);
        if (G > 2) {
            double u = t;
            swap16(t, u);
            t += u;
        }
        if (G > 4) {
            double u = t;
            swap32(t, u);
            t += u;
        }
        return t;
    };
    auto lane_value = [](double x, int l) -> double {
        return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l),
                                __builtin_amdgcn_readlane(__double2loint(x), l));
    };

    double w[4][8]; // position k (row r - 3 + k) in slot (phase + k) & 7
#pragma unroll
    for (int k = 0; k < 8; k++) {
        double t[4];
        load4(j0 + mstart - 3 + k, t);
        // open: a row above the grid extrapolates from the two before it
        // (the window starts at least 4 rows below row nz, so k >= 4 there)
        const bool ex = OPENB && k >= 2 && j0 + mstart - 3 + k >= nz;
#pragma unroll
        for (int f = 0; f < 4; f++)
            w[f][k] = ex ? w[f][k - 1] * 2 - w[f][k - 2] : t[f];
    }
    double Fin[4] = {0.0, 0.0, 0.0, 0.0}; // face r-1/2
    double mnext = coll ? load_mom(j0) : 0.0;
    double hnext = mat ? A.Hm[j0] : 0.0;
    __builtin_amdgcn_s_waitcnt(0);

    auto body = [&](auto P, int m) {
        constexpr int p = decltype(P)::value;
        constexpr int s0 = p & 7, s1 = (p + 1) & 7, s2 = (p + 2) & 7, s3 = (p + 3) & 7,
                      s4 = (p + 4) & 7, s5 = (p + 5) & 7, s6 = (p + 6) & 7, s7 = (p + 7) & 7;
        const bool more = m + 1 < n;
        const int r = j0 + m;
        // row r + 5, position 7 of the next iteration.  Loaded here, a full
        // iteration ahead of use, except by FV stages 1-3, which load it once
        // the faces are formed and the oldest row is dead: one row's
        // registers fewer across the WENO reconstruction gives them 3
        // wavefronts per SIMD instead of 2 (stage 4, HBM-bound at 2
        // wavefronts either way, keeps the earlier load: 0.65 against 0.72
        // ms at nz = 16384, nvz = 480)
        constexpr bool late = SCHEME == 0 && STAGE != 4;
        double nx[4];
        if (!late)
            load4(r + 5, nx);
        const bool cell = m >= 0;
        const size_t q = (size_t)(cell ? r : 0) * zs;
        double yv[4], y2v[4], y3v[4];
        double mcur = 0.0, hcur = 0.0;
        if (cell) {
            if (STAGE == 2 || STAGE == 4) {
                const double *yp = A.y + q;
#pragma unroll
                for (int f = 0; f < 4; f++)
                    yv[f] = yp[off[f]];
            }
            if (STAGE == 4) {
                const double *p2 = A.y2 + q, *p3 = A.y3 + q;
#pragma unroll
                for (int f = 0; f < 4; f++) {
                    y2v[f] = p2[off[f]];
                    y3v[f] = p3[off[f]];
                }
            }
            if (coll) {
                mcur = mnext;
                if (more)
                    mnext = load_mom(r + 1);
            }
            if (mat) {
                hcur = hnext;
                if (more)
                    hnext = A.Hm[r + 1];
            }
        }
        double K[4] = {0.0, 0.0, 0.0, 0.0};
        // open, top cell: the v < 0 face r-3/2 for the flux ghost, from the
        // window's rows r+2 .. r-3 and row r-4 read again (one wave-uniform
        // branch per range instead of a face kept through the march)
        double F2[4] = {0.0, 0.0, 0.0, 0.0};
        if (OPENB && SCHEME == 0 && advect && r == nz - 1) {
            double e[4];
            load4(r - 4, e);
#pragma unroll
            for (int f = 0; f < 4; f++)
                F2[f] = weno_up(w[f][s5], w[f][s4], w[f][s3], w[f][s2], w[f][s1], w[f][s0], e[f]);
        }
        if (SCHEME == 0 && advect) {
#pragma unroll
            for (int f = 0; f < 4; f++) {
                const double *u = w[f];
                double Fo;
                if (dir > 0)
                    Fo = weno_up(u[s0], u[s1], u[s2], u[s3], u[s4], u[s5], u[s6]);
                else if (dir < 0)
                    Fo = weno_up(u[s7], u[s6], u[s5], u[s4], u[s3], u[s2], u[s1]);
                else
                    Fo = weno_up(pos ? u[s0] : u[s7], pos ? u[s1] : u[s6], pos ? u[s2] : u[s5],
                                 pos ? u[s3] : u[s4], pos ? u[s4] : u[s3], pos ? u[s5] : u[s2],
                                 pos ? u[s6] : u[s1]);
                double Fl = Fin[f], Fr = Fo;
                if (OPENB && r == 0 && pos)
                    Fl = Fr; // rflux[-1] = rflux[0]
                if (OPENB && r == nz - 1 && !pos)
                    Fr = Fin[f] * 2 - F2[f]; // lflux[nz] extrapolated
                K[f] = fac * (Fr - Fl);
                Fin[f] = Fo;
            }
        }
        if (late)
            load4(r + 5, nx);
        if (cell) {
            const double pe = w[0][s3], qe = w[1][s3], re = w[2][s3], se = w[3][s3];
            if (SCHEME == 1) {
#pragma unroll
                for (int f = 0; f < 4; f++)
                    K[f] = fdfac * ((w[f][s1] - w[f][s5]) - 8.0 * (w[f][s2] - w[f][s4]));
            }
            if (vac) {
                K[0] += -2.0 * st * se;
                K[1] += 2.0 * st * se;
                K[2] += -2.0 * ct * se;
                K[3] += 2.0 * ct * re + st * (pe - qe);
            }
            if (mat) {
                const double hm = sig * hcur;
                K[2] += hm * se;
                K[3] -= hm * re;
            }
            if (coll) {
                const double t = sum_mom(mcur);
                double M[6];
#pragma unroll
                for (int i = 0; i < 6; i++)
                    M[i] = lane_value(t, i);
                const double S = mudv * (M[0] - vi * M[1]);
                const double R = mudv * (M[2] - vi * M[3]);
                const double D = mudv * (M[4] - vi * M[5]);
                const double I1 = 2.0 * (re * S - sig * se * R);
                K[0] += I1;
                K[1] -= I1;
                K[2] += (qe - pe) * S + sig * se * D;
                K[3] += sig * ((pe - qe) * R - re * D);
            }
            if (SCHEME == 1) {
#pragma unroll
                for (int f = 0; f < 4; f++) {
                    const double w0 = w[f][s0], w1 = w[f][s1], w2 = w[f][s2], w3 = w[f][s3],
                                 w4 = w[f][s4], w5 = w[f][s5], w6 = w[f][s6];
                    if (ko5)
                        K[f] += ko_eps * ((w0 + w6) - 6.0 * (w1 + w5) + 15.0 * (w2 + w4) -
                                          20.0 * w3);
                    else
                        K[f] += ko_eps * ((w1 + w5) - 4.0 * (w2 + w4) + 6.0 * w3);
                }
            }
            double o[4];
            double *op = A.yout + q;
#pragma unroll
            for (int f = 0; f < 4; f++) {
                if (STAGE == 0) {
                    o[f] = K[f];
                } else if (STAGE == 1) {
                    o[f] = w[f][s3] + A.a * K[f];
                } else if (STAGE == 4) {
                    const double y0 = yv[f];
                    const double d = ((y2v[f] - y0) + 2.0 * (y3v[f] - y0)) + (w[f][s3] - y0);
                    o[f] = y0 + (d * (1.0 / 3.0) + A.a * K[f]);
                } else {
                    o[f] = yv[f] + A.a * K[f];
                }
                if (live)
                    op[off[f]] = o[f];
            }
            if (STAGE != 0)
                store_moments(A.mom_out + ((size_t)r * G + g) * kMomSlot, lane, vi, sig,
                              live ? o[0] : 0.0, live ? o[1] : 0.0, live ? o[2] : 0.0,
                              live ? o[3] : 0.0);
        }
        const bool ex = OPENB && r + 5 >= nz; // rows r + 4, r + 3 in slots s7, s6
#pragma unroll
        for (int f = 0; f < 4; f++)
            w[f][s0] = ex ? w[f][s7] * 2 - w[f][s6] : nx[f]; // the slot of row r - 3 takes row r + 5
    };

    int m = mstart;
    for (;;) {
        if (m >= n) break;
        body(std::integral_constant<int, 0>{}, m++);
        if (m >= n) break;
        body(std::integral_constant<int, 1>{}, m++);
        if (m >= n) break;
        body(std::integral_constant<int, 2>{}, m++);
        if (m >= n) break;
        body(std::integral_constant<int, 3>{}, m++);
        if (m >= n) break;
        body(std::integral_constant<int, 4>{}, m++);
        if (m >= n) break;
        body(std::integral_constant<int, 5>{}, m++);
        if (m >= n) break;
        body(std::integral_constant<int, 6>{}, m++);
        if (m >= n) break;
        body(std::integral_constant<int, 7>{}, m++);
    }
}

template <int SCHEME, bool OPENB>
hipError_t launch_general(const StageArgs &a, int stage, int L, int nteams, hipStream_t s)
{
    const int G = a.VP / 32;
    int nwork = 0;
    long waves;
    if (nteams > 0) {
        L = (a.nz + nteams - 1) / nteams;
        waves = (long)nteams * G;
    } else {
        nwork = a.nz / L;
        waves = (long)nwork * G;
    }
    dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    switch (stage) {
    case 0: hipLaunchKernelGGL((cosenu_stream_general<SCHEME, 0, OPENB>), grid, block, 0, s, a, L, nwork, nteams); break;
    case 1: hipLaunchKernelGGL((cosenu_stream_general<SCHEME, 1, OPENB>), grid, block, 0, s, a, L, nwork, nteams); break;
    case 2:
    case 3:
        hipLaunchKernelGGL((cosenu_stream_general<SCHEME, 2, OPENB>), grid, block, 0, s, a, L, nwork, nteams); break;
    case 4: hipLaunchKernelGGL((cosenu_stream_general<SCHEME, 4, OPENB>), grid, block, 0, s, a, L, nwork, nteams); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int SCHEME, int OCC>
hipError_t launch_occ(const StageArgs &a, int stage, int L, int nteams, hipStream_t s)
{
    const int G = a.VP / 32;
    int nwork = 0;
    long waves;
    if (nteams > 0) {
        L = (a.nz + nteams - 1) / nteams;
        waves = (long)nteams * G;
    } else {
        nwork = a.nz / L;
        waves = (long)nwork * G;
    }
    dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    switch (stage) {
    case 0: hipLaunchKernelGGL((cosenu_stream_stage<SCHEME, 0, OCC>), grid, block, 0, s, a, L, nwork, nteams); break;
    case 1: hipLaunchKernelGGL((cosenu_stream_stage<SCHEME, 1, OCC>), grid, block, 0, s, a, L, nwork, nteams); break;
    case 2:
    case 3: // stages 2 and 3 run the same code with their own coefficient
        hipLaunchKernelGGL((cosenu_stream_stage<SCHEME, 2, OCC>), grid, block, 0, s, a, L, nwork, nteams); break;
    case 4: hipLaunchKernelGGL((cosenu_stream_stage<SCHEME, 4, 2>), grid, block, 0, s, a, L, nwork, nteams); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Three wavefronts per SIMD (four would spill the FV window), two at stage
// 4, which holds three more stage states per cell.
template <int SCHEME>
hipError_t launch_scheme(const StageArgs &a, int stage, int L, int nteams, hipStream_t s)
{
    return launch_occ<SCHEME, 3>(a, stage, L, nteams, s);
}

// Resident wavefronts of the stage-2/3 launch or (stage4) of the stage-4
// launch, which holds more registers and fewer wavefronts per SIMD.
template <int SCHEME>
int resident_waves(bool general, bool stage4)
{
    int nb = 0, dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return 0;
    const void *k =
        general ? (stage4 ? reinterpret_cast<const void *>(&cosenu_stream_general<SCHEME, 4, false>)
                          : reinterpret_cast<const void *>(&cosenu_stream_general<SCHEME, 2, false>))
                : (stage4 ? reinterpret_cast<const void *>(&cosenu_stream_stage<SCHEME, 4, 2>)
                          : reinterpret_cast<const void *>(&cosenu_stream_stage<SCHEME, 2, 3>));
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, 256, 0) != hipSuccess)
        return 0;
    return nb * prop.multiProcessorCount * 4;
}

} // namespace

int stream_supported(int nvz, int VP, int bc, const double *vz)
{
    if (bc != 0 || nvz % 64 != 0 || VP != nvz || VP < 64 || VP > 1024)
        return 0;
    for (int g = 0; g < VP / 32; g++) {
        bool pos = false, neg = false;
        for (int k = 0; k < 32; k++) {
            pos = pos || vz[g * 32 + k] > 0.0;
            neg = neg || vz[g * 32 + k] < 0.0;
        }
        if (pos && neg)
            return 0;
    }
    return 1;
}

hipError_t launch_stage_stream(const StageArgs &a, int stage, int L, int nteams, bool general,
                               hipStream_t s)
{
    if (a.VP % 64 != 0 || a.nvz > a.VP || a.nz < 8 || !a.mom_in || (stage != 0 && !a.mom_out))
        return hipErrorInvalidValue;
    if (!general && (a.VP != a.nvz || a.bc != 0))
        return hipErrorInvalidValue;
    if (nteams > 0) {
        if (nteams > a.nz)
            return hipErrorInvalidValue;
    } else if (L < 1 || a.nz % L != 0) {
        return hipErrorInvalidValue;
    }
    if (general) {
        if (a.bc == 1)
            return a.scheme == 0 ? launch_general<0, true>(a, stage, L, nteams, s)
                                 : launch_general<1, true>(a, stage, L, nteams, s);
        return a.scheme == 0 ? launch_general<0, false>(a, stage, L, nteams, s)
                             : launch_general<1, false>(a, stage, L, nteams, s);
    }
    return a.scheme == 0 ? launch_scheme<0>(a, stage, L, nteams, s)
                         : launch_scheme<1>(a, stage, L, nteams, s);
}

hipError_t launch_stream_moments(const double *y, const double *vz, double *mom, int nz, int nvz,
                                 int VP, hipStream_t s)
{
    if (VP % 64 != 0 || nvz > VP)
        return hipErrorInvalidValue;
    const long waves = (long)nz * (VP / 32);
    hipLaunchKernelGGL(cosenu_stream_moments, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, y,
                       vz, mom, nz, nvz, VP);
    return hipGetLastError();
}

int stream_resident_waves(int scheme, bool general, bool stage4)
{
    return scheme == 0 ? resident_waves<0>(general, stage4) : resident_waves<1>(general, stage4);
}

const char *stage_stream_kernel_name(int stage, bool general)
{
    static char buf[64];
    const int st = stage + 1 == 3 ? 2 : stage + 1;
    snprintf(buf, sizeof(buf), general ? "cosenu_stream_general<*, %d, *>" : "cosenu_stream_stage<*, %d>",
             st);
    return buf;
}

} // namespace cosenu
