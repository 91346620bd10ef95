// FAST mode, FD scheme: all four RK4 stages in one sweep (variant 5).
//
// The finite-difference right-hand side with 3rd-order Kreiss-Oliger
// dissipation (rhs_fd.hpp:6-148) reads z +- 2 only, so a stage's input
// window is 5 rows.  The split-lane team of the two-stage kernel (VP/32
// wavefronts covering all velocity bins of a z range, lanes 0-31 carrying ee,
// xx, ex_re, ex_im and lanes 32-63 the barred fields of 32 bins) marches
// along z with four register windows: y, read from HBM, and the outputs of
// stages 1, 2 and 3, produced in registers.  Stage s runs 2 cells behind
// stage s - 1 (iteration i: stage 1 at cell i, stage 2 at i - 2, stage 3 at
// i - 4, stage 4 at i - 6), so every window holds values already computed.
// Each stage covers the team's cells plus 2 (4 - s) on each side, which the
// later stages need for their stencils.
//
// RK4 (nuosc.hpp:320-348) in the stage-value form of the other FAST kernels:
//   Y2 = y + (dt/2) K1, Y3 = y + (dt/2) K2, Y4 = y + dt K3,
//   T = y + (Q + (Y4 - y))/3, Q = (Y2 - y) + 2 (Y3 - y),  y' = T + (dt/6) K4.
// Stage 2 holds y (the row leaving the first window), Y2 (its window's
// centre) and Y3 of its cell, forms Q and stores Q and y in wavefront-private
// LDS rings; stage 3 takes y from its ring and replaces Q by T, which is all
// stage 4 needs.  A ring slot lives from stage 2 (iteration c + 2) to stage 4
// (c + 6): 5 slots for Q / T, 3 for y (c + 2 .. c + 4), 16 KB per wavefront.
// HBM traffic per cell and step: read y, write y' (128 B, plus the halo
// rows) instead of 448 B for the two-stage launches.  (T rounds a
// state-sized value once more than the two-stage kernel's y + (d/3 +
// (dt/6) K4): unbiased, ~1e-14 relative after 20000 steps, and a zero
// right-hand side still leaves the state unchanged.)
//
// Moments: each stage's input moments are reduced when its input row is
// produced (stage 1: when a y row enters the window) and read 2 iterations
// later, when the row reaches the stage's centre; as no two consecutive
// iterations lack a barrier, one separates every store from its reads and
// every last read of a ring slot from its next store (6 iterations apart).
//
// The new state is written to a second buffer (neighbouring teams read y in
// their halos while the launch runs); the state buffers swap after every
// step.
//
// Open boundary (nuosc.hpp:247-277; OPENB): y rows below the grid read row 0
// (the load clamps its row index), rows above it enter the window as 2 x the
// previous row - the one before (the reference's ghost fill, row by row).
// Each stage's output at its ghost rows follows the same rule instead of
// being computed: its row-0 output is copied into the window slots of rows
// -2 and -1 when it is produced, and above the grid a ghost row is
// extrapolated from the two rows before it in the stage's window.
#include "fast_math.hpp"
#include "wide_exchange.hpp"

#include <cstdio>
#include <type_traits>

namespace cosenu {
namespace {

using namespace fastk;

// The z loop runs in groups of 5 iterations (the window length), so that a
// window shift is a change of which register holds which row.  Barriers at
// iterations 0, 2, 4 of each group: never two iterations in a row without
// one, which is what the moment rings need (a row's moments are stored 2
// iterations before they are read; a slot is reused kRing - 2 = 6 iterations
// after its last read).
constexpr int kRing = 8;
// the physics of the benchmark workloads: vacuum and collective terms with
// advection and 3rd-order dissipation, no matter term
constexpr int kStdPhys = kFlagVac | kFlagColl;
template <int T>
using Rot = std::integral_constant<int, T>;
constexpr int kTRing = 5, kYRing = 3; // ring slots per wavefront: Q / T and y
constexpr int kQRing = (kTRing + kYRing) * 4 * 64; // doubles per wavefront: [slot][field][lane]

// UNI: no padding bins (nvz == VP).  MAT: the matter term may be on (its
// potential is then loaded ahead of use, which costs registers that runs
// without it need not hold).  OPENB: the open boundary.  PHYS: the physics
// switches (kFlag* bits) fixed at compile time, or -1 to read them from the
// arguments (nz = 65536 x 128: 2% faster with them fixed).  ca[0..3]: dt/2,
// dt/2, dt, dt/6.  NB: blocks per team (1: the whole team in this block;
// > 1: this block holds bins [X.role * NW * 32, (X.role + 1) * NW * 32) of
// the team's rows, and the moments are summed over the team's blocks).
template <int NW, bool UNI, bool MAT, bool OPENB, int PHYS = -1, int NB = 1>
__device__ __forceinline__ void quad_body(const StageArgs &A, const double ca0, const double ca1,
                                          const double ca2, const double ca3, double *yout, int L,
                                          int j0, int jend, bool active, double (*rings)[kRing][NW][8],
                                          double *qring, const WideX &X = WideX{},
                                          double (*xst)[2][64] = nullptr)
// rings: [4][kRing][NW][8] (stage s input moments); qring: the team's NW x kQRing;
// xst (NB > 1): [4][2][64], the partner's partials staged in LDS
{
    static_assert(NB == 1 || NB == 2, "teams of one or two blocks");
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wit = wave % NW;
    const int h = lane >> 5;
    const int v = (NB > 1 ? X.role * NW * 32 : 0) + wit * 32 + (lane & 31);
    const int VP = A.VP, nz = A.nz;
    const size_t zs = (size_t)8 * VP;
    const double sig = h ? -1.0 : 1.0;
    j0 = __builtin_amdgcn_readfirstlane(j0);
    jend = __builtin_amdgcn_readfirstlane(jend);

    const double vi = A.vz[v];
    const bool live = UNI || v < A.nvz;
    const bool store = active && live;
    const int loff = 4 * h * VP + v; // this lane's first field within a z row
    double *const qr = qring + wit * kQRing;

    auto zrow = [&](int m) -> int {
        if (OPENB)
            return min(max(m, 0), nz - 1);
        int w = m < 0 ? m + nz : m; // m >= -12 > -nz (nz >= 16)
        while (w >= nz)
            w -= nz;
        return w;
    };
    auto load4 = [&](const double *base, int m, double o[4]) {
        const double *r = base + (size_t)zrow(m) * zs + loff;
#pragma unroll
        for (int f = 0; f < 4; f++)
            o[f] = r[(size_t)f * VP];
    };

    const int flags = PHYS >= 0 ? PHYS : A.flags;
    const bool coll = (flags & kFlagColl) != 0;
    const bool vac = (flags & kFlagVac) != 0, mat = MAT && (flags & kFlagMat) != 0;
    const bool advect = !(flags & kFlagAdvecOff);
    // without advection the stencil is multiplied by 0 (no per-value select)
    const double fdfac = advect ? -vi / (12.0 * A.dz) : 0.0;
    const double ko_eps = -A.ko / A.dz / 16.0;
    // Advection (rhs_fd.hpp:61-62), fdfac ((u-2 - u2) - 8 (u-1 - u1)), and
    // 3rd-order dissipation (rhs_fd.hpp:135-136), ko_eps (u-2 + u2 -
    // 4 (u-1 + u1) + 6 u0), are one linear 5-point stencil per bin:
    // sum_k c_k u(k-2), one multiply and four fused multiply-adds per field
    // instead of nine operations.  (The coefficients sum to zero; the
    // expanded sum rounds with ~eps sum_k |c_k u_k|, 1e-16 relative to the
    // state per step at the workloads' dz.)
    const double cs0 = fdfac + ko_eps, cs4 = ko_eps - fdfac, cs2 = 6.0 * ko_eps;
    const double cs1 = fma(-8.0, fdfac, -4.0 * ko_eps), cs3 = fma(8.0, fdfac, -4.0 * ko_eps);
    const double mudv = A.mu * A.dv;
    const double st = A.st * A.pmo, ct = A.ct * A.pmo, st2 = 2.0 * st, ct2 = 2.0 * ct;

    // NB > 1: the device-memory slots of ring rg (stage rg + 1's input
    // moments), block b, z row `row` (32-bit word offsets: the rings of a
    // launch hold at most a few million words)
    const unsigned xteam = (unsigned)X.team * (NB * kRing * 128);
    const unsigned xring = (unsigned)X.nteams * (NB * kRing * 128);
    auto xslot = [&](int rg, int b, int row) -> unsigned long long * {
        return X.slots + (xteam + (unsigned)rg * xring + (unsigned)b * (kRing * 128) +
                          (unsigned)(row & (kRing - 1)) * 128u);
    };
    bool gave_up = false;
    // partial moments of 4 field values of this lane into rings[rg][row %
    // kRing]; returns the lane's total (wave_sum6)
    auto moments = [&](int rg, int row, double p, double q, double r, double s) {
        const double S = live ? s : 0.0;
        const double R = live ? sig * r : 0.0;
        const double D = live ? sig * (p - q) : 0.0;
        const double m6[6] = {S, vi * S, R, vi * R, D, vi * D};
        const double t = wave_sum6(m6, lane);
        const int idx = (kMomentLane >> (4 * (lane >> 3))) & 15;
        if ((lane & 7) == 0 && idx < 6)
            rings[rg][row & (kRing - 1)][wit][idx] = t;
        return t;
    };
    // NB > 1: publish a row's partials (t: moments' return value) to this
    // block's device-memory ring rg.  Every lane stores, its 8-lane group's
    // total (the same value in all 8 lanes) to the group's element; the two
    // groups holding duplicates write element 7, which no block reads.  The
    // exchange's loads and stores are issued on every z step, whatever the
    // row (a row no stage needs publishes zeros, and its slot is rewritten
    // before a needed row's words are read from it): the compiler's
    // in-order count of outstanding memory operations then holds on every
    // path, and a wait for a partner word does not also wait for the
    // step's own window loads.
    auto xpub = [&](int rg, int row, double t) {
        if constexpr (NB > 1) {
            const int idx = (kMomentLane >> (4 * (lane >> 3))) & 15;
            xput(xslot(rg, X.role, row) + 2 * (wit * 8 + idx), xtag(X.epoch, row), t);
        }
    };
    // NB > 1: is row r's total needed by stage s + 1 (its cells, see below)
    auto xneed = [&](int s, int r) {
        const int m = 6 - 2 * s;
        return coll && active && r >= j0 - m && r < jend + m && !(OPENB && (r < 0 || r >= nz));
    };
    // NB > 1: the partner block's partials of the rows the stages take next
    // pass through an LDS staging buffer, xst[ring][row & 1][64 elements].
    // Wavefront r < 4 loads ring r's partner words of the row its stage needs
    // at the next z step, checks their tags at the end of this step (waiting,
    // bounded, for words not yet written) and writes the values to the
    // buffer.  After the next barrier (one on every z step when NB > 1) every
    // wavefront adds them to its own block's partials from the LDS rings:
    // own + partner, which both blocks of the team round alike (IEEE
    // addition commutes).  One wavefront per ring loads a row instead of
    // eight, and holds one ring's words instead of four (FD, nz = 16384,
    // nvz = 512: 0.444 against 0.474 ms per step; nvz = 480, open: 0.619
    // against 0.781).
    const bool xlane = (lane & 7) < 6; // elements 6, 7 of a wavefront's row are unused
    const int xr = wit;                // the ring this wavefront stages (< 4)
    XWords xw = {0ull, 0ull};
    auto xstage = [&](int row) {
        if (xr < 4 && xneed(xr, row) && xlane) {
            const double v = xvalue(xw, xslot(xr, X.role ^ 1, row) + 2 * lane,
                                    xtag(X.epoch, row) ^ X.tag_flip, X.ctl, X.wait_ticks, gave_up);
            xst[xr][row & 1][lane] = v;
        }
    };
    // right-hand side at the window centre of z row r: mx = this lane's
    // moment_partial of the row's ring slot, hmr = Hm[r] (MAT); the terms in
    // the order of the other FAST kernels.  The window's rows r-2 .. r+2
    // are in w[f][(RT + k) % 5], k = 0..4.  Every fused multiply-add is
    // written out (this file is built without contraction), so the 5
    // unrolled copies of the z step evaluate a cell identically whichever
    // copy it falls in.
    auto rhs = [&](const double (&w)[4][5], auto rot, double mx, double hmr,
                   double K[4]) __attribute__((always_inline)) {
        constexpr int RT = decltype(rot)::value;
        constexpr int k0 = RT % 5, k1 = (RT + 1) % 5, k2 = (RT + 2) % 5, k3 = (RT + 3) % 5,
                      k4 = (RT + 4) % 5;
        const double p = w[0][k2], qq = w[1][k2], r = w[2][k2], s = w[3][k2];
#pragma unroll
        for (int f = 0; f < 4; f++)
            K[f] = fma(cs4, w[f][k4], fma(cs0, w[f][k0],
                   fma(cs3, w[f][k3], fma(cs1, w[f][k1], cs2 * w[f][k2]))));
        if (vac) {
            K[0] = fma(-st2, s, K[0]);
            K[1] = fma(st2, s, K[1]);
            K[2] = fma(-ct2, s, K[2]);
            K[3] = fma(ct2, r, fma(st, p - qq, K[3]));
        }
        if (mat) {
            const double hm = sig * hmr;
            K[2] = fma(hm, s, K[2]);
            K[3] = fma(-hm, r, K[3]);
        }
        if (coll) {
            double M[6];
            moment_totals<NW>(mx, M);
            const double S = mudv * fma(-vi, M[1], M[0]);
            const double R = mudv * fma(-vi, M[3], M[2]);
            const double D = mudv * fma(-vi, M[5], M[4]);
            const double ss = sig * s;
            const double I1 = 2.0 * fma(r, S, -(ss * R));
            K[0] += I1;
            K[1] -= I1;
            K[2] = fma(qq - p, S, fma(ss, D, K[2]));
            K[3] = fma(sig, fma(p - qq, R, -(r * D)), K[3]);
        }
    };
    // ring slots by cell (cells >= ibeg - 6, see below)
    const int cbase = j0 - 12;
    auto tslot = [&](int cell, int f) -> double & {
        return qr[(((cell - cbase) % kTRing) * 4 + f) * 64 + lane];
    };
    auto yslot = [&](int cell, int f) -> double & {
        return qr[((kTRing + (cell - cbase) % kYRing) * 4 + f) * 64 + lane];
    };

    // Iteration i: stage 1 at cell i (cells [j0-6, jend+6)), stage 2 at i-2
    // ([j0-4, jend+4)), stage 3 at i-4 ([j0-2, jend+2)), stage 4 at i-6
    // ([j0, jend)).  The trip count depends on L only: the same iterations,
    // and barriers, in every team of a block.
    const int ibeg = j0 - 6, iend = j0 + L + 6;
    double w0[4][5], w1[4][5], w2[4][5], w3[4][5];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        double t[4];
        load4(A.yin, ibeg - 2 + k, t);
#pragma unroll
        for (int f = 0; f < 4; f++) {
            w0[f][k] = t[f];
            w1[f][k] = 0.0;
            w2[f][k] = 0.0;
            w3[f][k] = 0.0;
        }
    }
    {
        double t0 = 0.0, t1 = 0.0;
        if (coll) {
            t0 = moments(0, ibeg, w0[0][2], w0[1][2], w0[2][2], w0[3][2]);
            t1 = moments(0, ibeg + 1, w0[0][3], w0[1][3], w0[2][3], w0[3][3]);
        }
        xpub(0, ibeg, t0);
        xpub(0, ibeg + 1, t1);
    }
    // NB > 1: the rows the stages take at the first z step, staged before its
    // barrier
    if constexpr (NB > 1) {
        if (xr < 4) {
            const int row = ibeg - 2 * xr;
            xw = xload(xslot(xr, X.role ^ 1, row) + 2 * lane);
            xstage(row);
        }
    }
    double nxt[4];
    load4(A.yin, ibeg + 3, nxt); // the row that enters the window after iteration ibeg
    // The prologue's loads have landed before the loop (the wait-count
    // analysis would otherwise take them as in flight inside the loop).
    __builtin_amdgcn_s_waitcnt(0);

    // Iteration i = ibeg + 5 m + T.  Window rows: w0 holds y rows i-2 .. i+2
    // at w0[f][(T + k) % 5]; w1, w2, w3 hold stage outputs and are shifted
    // before their stage runs, after which their rows c-2 .. c+2 are at
    // [(T + 1 + k) % 5].  A shift overwrites slot T % 5 (the oldest row).
    auto step = [&](const int i, auto tc) __attribute__((always_inline)) {
        constexpr int T = decltype(tc)::value;
        constexpr int old = T % 5, cen0 = (T + 2) % 5, top0 = (T + 4) % 5, cen1 = (T + 3) % 5;
        if (NB > 1 || T % 2 == 0) // the same iterations in every team
            __syncthreads();
        const int c1 = i, c2 = i - 2, c3 = i - 4, c4 = i - 6;
        // the stages' input moments (rows stored 2 iterations ago, behind a
        // barrier), read before stage 1 needs them
        double mx[4] = {0.0, 0.0, 0.0, 0.0};
        if (coll) {
            mx[0] = moment_partial<NW>(rings[0][c1 & (kRing - 1)], lane);
            mx[1] = moment_partial<NW>(rings[1][c2 & (kRing - 1)], lane);
            mx[2] = moment_partial<NW>(rings[2][c3 & (kRing - 1)], lane);
            mx[3] = moment_partial<NW>(rings[3][c4 & (kRing - 1)], lane);
        }
        // NB > 1: the partner's partials of the four stages' rows, staged
        // behind the barrier; the staging wavefront's load for the next z
        // step, issued before this step's window load
        if constexpr (NB > 1) {
            const int cc[4] = {c1, c2, c3, c4};
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (xneed(r, cc[r]))
                    mx[r] = mx[r] + xst[r][cc[r] & 1][lane];
            if (xr < 4) // issued before the window load
                xw = xload(xslot(xr, X.role ^ 1, i + 1 - 2 * xr) + 2 * lane);
        }
        // the matter potential of the four stages' rows, loaded ahead of use
        double hmv[4] = {0.0, 0.0, 0.0, 0.0};
        if (MAT && mat) {
            hmv[0] = A.Hm[zrow(c1)];
            hmv[1] = A.Hm[zrow(c2)];
            hmv[2] = A.Hm[zrow(c3)];
            hmv[3] = A.Hm[zrow(c4)];
        }
        // the next window row
        double nn[4];
        load4(A.yin, i + 4, nn);
        {
            double t0 = 0.0;
            if (coll)
                t0 = moments(0, i + 2, w0[0][top0], w0[1][top0], w0[2][top0], w0[3][top0]);
            xpub(0, i + 2, t0);
        }

        // OPENB: a stage's output at a cell above the grid extrapolates from
        // the two rows before it in the stage's window (slots (T+4)%5 and
        // (T+3)%5 before the new row is written); at cell 0 it is also
        // copied into the slots of rows -2 and -1
        constexpr int prev1 = (T + 4) % 5, prev2 = (T + 3) % 5;
        auto ghost = [&](double (&w)[4][5], int c, const double o[4]) __attribute__((always_inline)) {
#pragma unroll
            for (int f = 0; f < 4; f++)
                w[f][old] = c >= nz ? w[f][prev1] * 2 - w[f][prev2] : o[f];
            if (c == 0) {
#pragma unroll
                for (int f = 0; f < 4; f++)
                    w[f][prev1] = w[f][prev2] = o[f];
            }
        };
        // stage 1 at c1: Y2 = y + (dt/2) K1
        double o[4] = {0.0, 0.0, 0.0, 0.0};
        double tp = 0.0; // the published total of the stage's output row
        if (c1 >= j0 - 6 && c1 < jend + 6 && !(OPENB && (c1 < 0 || c1 >= nz))) {
            double K[4];
            rhs(w0, Rot<T>{}, mx[0], hmv[0], K);
#pragma unroll
            for (int f = 0; f < 4; f++)
                o[f] = fma(ca0, K[f], w0[f][cen0]);
            if (coll)
                tp = moments(1, c1, o[0], o[1], o[2], o[3]);
        }
        xpub(1, c1, tp);
        if (OPENB) {
            ghost(w1, c1, o);
        } else {
#pragma unroll
            for (int f = 0; f < 4; f++)
                w1[f][old] = o[f];
        }
        // stage 2 at c2: Y3; Q and y to the rings.  y of this cell is the y
        // window's oldest row (shifted out at the end of the iteration).
        tp = 0.0;
        if (c2 >= j0 - 4 && c2 < jend + 4 && !(OPENB && (c2 < 0 || c2 >= nz))) {
            double K[4];
            rhs(w1, Rot<T + 1>{}, mx[1], hmv[1], K);
#pragma unroll
            for (int f = 0; f < 4; f++) {
                const double y2 = w0[f][old];
                o[f] = fma(ca1, K[f], y2);
                tslot(c2, f) = fma(2.0, o[f] - y2, w1[f][cen1] - y2);
                yslot(c2, f) = y2;
            }
            if (coll)
                tp = moments(2, c2, o[0], o[1], o[2], o[3]);
        } else {
#pragma unroll
            for (int f = 0; f < 4; f++)
                o[f] = 0.0;
        }
        xpub(2, c2, tp);
        if (OPENB) {
            ghost(w2, c2, o);
        } else {
#pragma unroll
            for (int f = 0; f < 4; f++)
                w2[f][old] = o[f];
        }
        // stage 3 at c3: Y4 = y + dt K3; T = y + (Q + (Y4 - y))/3 replaces Q
        tp = 0.0;
        if (c3 >= j0 - 2 && c3 < jend + 2 && !(OPENB && (c3 < 0 || c3 >= nz))) {
            double K[4];
            rhs(w2, Rot<T + 1>{}, mx[2], hmv[2], K);
#pragma unroll
            for (int f = 0; f < 4; f++) {
                const double yv = yslot(c3, f);
                o[f] = fma(ca2, K[f], yv);
                double &t = tslot(c3, f);
                t = fma(t + (o[f] - yv), 1.0 / 3.0, yv);
            }
            if (coll)
                tp = moments(3, c3, o[0], o[1], o[2], o[3]);
        } else {
#pragma unroll
            for (int f = 0; f < 4; f++)
                o[f] = 0.0;
        }
        xpub(3, c3, tp);
        if (OPENB) {
            ghost(w3, c3, o);
        } else {
#pragma unroll
            for (int f = 0; f < 4; f++)
                w3[f][old] = o[f];
        }
        // stage 4 at c4: y' = T + (dt/6) K4
        if (c4 >= j0 && c4 < jend) {
            double K[4];
            rhs(w3, Rot<T + 1>{}, mx[3], hmv[3], K);
            if (store) {
                double *out = yout + (size_t)zrow(c4) * zs + loff;
#pragma unroll
                for (int f = 0; f < 4; f++)
                    out[(size_t)f * VP] = fma(ca3, K[f], tslot(c4, f));
            }
        }
        if constexpr (NB > 1) { // (checked here: the load had the whole z step to land)
            if (xr < 4)
                xstage(i + 1 - 2 * xr);
        }
        // the y window advances (OPENB: row i+3 above the grid extrapolates
        // from rows i+2 and i+1)
        const bool ex = OPENB && i + 3 >= nz;
#pragma unroll
        for (int f = 0; f < 4; f++) {
            w0[f][old] = ex ? w0[f][top0] * 2 - w0[f][prev2] : nxt[f];
            nxt[f] = nn[f];
        }
    };

    for (int i = ibeg; i < iend; i += 5) {
        step(i, Rot<0>{});
        if (i + 1 >= iend)
            break;
        step(i + 1, Rot<1>{});
        if (i + 2 >= iend)
            break;
        step(i + 2, Rot<2>{});
        if (i + 3 >= iend)
            break;
        step(i + 3, Rot<3>{});
        if (i + 4 >= iend)
            break;
        step(i + 4, Rot<4>{});
    }
}

template <int NW, bool UNI, bool MAT, bool OPENB, int PHYS = -1>
__global__ __launch_bounds__(NW >= 8 ? 512 : 256) __attribute__((amdgpu_waves_per_eu(2)))
void cosenu_quad_stage(StageArgs A, double ca0, double ca1, double ca2, double ca3, double *yout,
                       int L, int nchunks, int nteams)
{
    constexpr int TPB = NW >= 4 ? 1 : 4 / NW; // teams per block
    __shared__ double rings[TPB][4][kRing][NW][8];
    __shared__ double qring[TPB * NW * kQRing];
    const int team = (threadIdx.x >> 6) / NW;
    const int g = xcd_block_order() * TPB + team;
    double *const qr = qring + team * NW * kQRing;
    if (nteams > 0) {
        const bool active = g < nteams;
        const int j0 = active ? (int)(((long)g * A.nz) / nteams) : 0;
        const int j1 = active ? (int)(((long)(g + 1) * A.nz) / nteams) : 0;
        quad_body<NW, UNI, MAT, OPENB, PHYS>(A, ca0, ca1, ca2, ca3, yout, L, j0, j1, active,
                                             rings[team], qr);
    } else {
        const bool active = g < nchunks;
        const int j0 = (active ? g : 0) * L;
        quad_body<NW, UNI, MAT, OPENB, PHYS>(A, ca0, ca1, ca2, ca3, yout, L, j0, j0 + L, active,
                                             rings[team], qr);
    }
}

// Rows of 512 bins: a team of NB = 2 blocks of 8 wavefronts per contiguous z
// range (range mode only: the grid is nteams * NB blocks, every team the same
// trip count).  Teams form by ticket (WideX); the partner's partials come in
// through xst, one loader wavefront per ring.  (Teams of 4 blocks, for rows
// of 1024 bins, in a form that held 3 partners' words of 4 rings in every
// wavefront's registers, spilled 192-300 B per lane and ran 1.9 ms per step
// against 1.05 ms for the two-stage kernel at nz = 8192: those rows take the
// two-stage kernel.)
template <bool UNI, bool MAT, bool OPENB, int NB, int PHYS = -1>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2)))
void cosenu_quad_wide(StageArgs A, double ca0, double ca1, double ca2, double ca3, double *yout,
                      int nteams, unsigned long long *slots, unsigned long long *ctl,
                      long long wait_ticks, unsigned tag_flip)
{
    constexpr int NW = 8;
    __shared__ double rings[4][kRing][NW][8];
    __shared__ double qring[NW * kQRing];
    __shared__ double xst[4][2][64];
    const WideX X = wide_join<NB>(slots, ctl, wait_ticks, tag_flip, nteams);
    const bool active = X.team < nteams;
    const int L = (A.nz + nteams - 1) / nteams;
    const int j0 = active ? (int)(((long)X.team * A.nz) / nteams) : 0;
    const int j1 = active ? (int)(((long)(X.team + 1) * A.nz) / nteams) : 0;
    quad_body<NW, UNI, MAT, OPENB, PHYS, NB>(A, ca0, ca1, ca2, ca3, yout, L, j0, j1, active, rings,
                                           qring, X, xst);
}

template <int NB>
hipError_t launch_wide_nb(const StageArgs &a, const double ca[4], double *yout, int nteams,
                          bool uni, unsigned long long *slots, unsigned long long *ctl,
                          long long ticks, unsigned flip, hipStream_t s)
{
    dim3 grid(nteams * NB), block(512);
    const bool mat = (a.flags & kFlagMat) != 0;
    if (a.bc == 1 && mat) // the open boundary: the general bin layout
        hipLaunchKernelGGL((cosenu_quad_wide<false, true, true, NB>), grid, block, 0, s, a, ca[0],
                           ca[1], ca[2], ca[3], yout, nteams, slots, ctl, ticks, flip);
    else if (a.bc == 1)
        hipLaunchKernelGGL((cosenu_quad_wide<false, false, true, NB>), grid, block, 0, s, a, ca[0],
                           ca[1], ca[2], ca[3], yout, nteams, slots, ctl, ticks, flip);
    else if (uni && a.flags == kStdPhys) // the benchmark physics fixed at compile time
        hipLaunchKernelGGL((cosenu_quad_wide<true, false, false, NB, kStdPhys>), grid, block, 0, s,
                           a, ca[0], ca[1], ca[2], ca[3], yout, nteams, slots, ctl, ticks, flip);
    else if (uni)
        hipLaunchKernelGGL((cosenu_quad_wide<true, true, false, NB>), grid, block, 0, s, a, ca[0],
                           ca[1], ca[2], ca[3], yout, nteams, slots, ctl, ticks, flip);
    else
        hipLaunchKernelGGL((cosenu_quad_wide<false, true, false, NB>), grid, block, 0, s, a, ca[0],
                           ca[1], ca[2], ca[3], yout, nteams, slots, ctl, ticks, flip);
    return hipGetLastError();
}

// Batched runs: work item g = (run, first z cell) of a packed list; every
// item of a launch has the same length L (a run's last item may hold fewer
// rows), so every team of a block runs the same iterations and barriers.
// The new state goes to the run's spare buffer (acc).
template <int NW, bool UNI, bool MAT, int PHYS = -1>
__global__ __launch_bounds__(NW >= 8 ? 512 : 256) __attribute__((amdgpu_waves_per_eu(2)))
void cosenu_quad_batch(const BatchRun *runs, const int2 *items, int nitems, int L)
{
    constexpr int TPB = NW >= 4 ? 1 : 4 / NW;
    __shared__ double rings[TPB][4][kRing][NW][8];
    __shared__ double qring[TPB * NW * kQRing];
    const int team = (threadIdx.x >> 6) / NW;
    const int item = xcd_block_order() * TPB + team;
    const bool active = item < nitems;
    const int2 it = items[active ? item : 0];
    const BatchRun &r = runs[it.x];
    const StageArgs A = batch_stage_args(r, 1, true);
    const int jend = min(it.y + L, r.nz);
    quad_body<NW, UNI, MAT, false, PHYS>(A, r.a[0], r.a[1], r.a[2], r.a[3], r.acc, L, it.y, jend,
                                         active, rings[team], qring + team * NW * kQRing);
}

template <int NW>
hipError_t launch_batch_nw(const BatchRun *runs, const int2 *items, int nitems, int L, bool uni,
                           bool mat, bool std_phys, hipStream_t s)
{
    constexpr int TPB = NW >= 4 ? 1 : 4 / NW;
    if (nitems <= 0)
        return hipSuccess;
    dim3 grid((nitems + TPB - 1) / TPB), block(64 * NW * TPB);
    if (uni && std_phys)
        hipLaunchKernelGGL((cosenu_quad_batch<NW, true, false, kStdPhys>), grid, block, 0, s, runs, items, nitems, L);
    else if (uni && mat)
        hipLaunchKernelGGL((cosenu_quad_batch<NW, true, true>), grid, block, 0, s, runs, items, nitems, L);
    else if (uni)
        hipLaunchKernelGGL((cosenu_quad_batch<NW, true, false>), grid, block, 0, s, runs, items, nitems, L);
    else if (mat)
        hipLaunchKernelGGL((cosenu_quad_batch<NW, false, true>), grid, block, 0, s, runs, items, nitems, L);
    else
        hipLaunchKernelGGL((cosenu_quad_batch<NW, false, false>), grid, block, 0, s, runs, items, nitems, L);
    return hipGetLastError();
}

template <int NW>
hipError_t launch_nw(const StageArgs &a, const double ca[4], double *yout, int L, int nteams,
                     bool uni, hipStream_t s)
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
    const bool mat = (a.flags & kFlagMat) != 0;
    if (a.bc == 1 && mat) // the open boundary: the general bin layout
        hipLaunchKernelGGL((cosenu_quad_stage<NW, false, true, true>), grid, block, 0, s, a, ca[0],
                           ca[1], ca[2], ca[3], yout, L, nchunks, nteams);
    else if (a.bc == 1)
        hipLaunchKernelGGL((cosenu_quad_stage<NW, false, false, true>), grid, block, 0, s, a, ca[0],
                           ca[1], ca[2], ca[3], yout, L, nchunks, nteams);
    else if (uni && mat)
        hipLaunchKernelGGL((cosenu_quad_stage<NW, true, true, false>), grid, block, 0, s, a, ca[0],
                           ca[1], ca[2], ca[3], yout, L, nchunks, nteams);
    else if (uni && a.flags == kStdPhys)
        hipLaunchKernelGGL((cosenu_quad_stage<NW, true, false, false, kStdPhys>), grid, block, 0, s,
                           a, ca[0], ca[1], ca[2], ca[3], yout, L, nchunks, nteams);
    else if (uni)
        hipLaunchKernelGGL((cosenu_quad_stage<NW, true, false, false>), grid, block, 0, s, a, ca[0],
                           ca[1], ca[2], ca[3], yout, L, nchunks, nteams);
    else if (mat)
        hipLaunchKernelGGL((cosenu_quad_stage<NW, false, true, false>), grid, block, 0, s, a, ca[0],
                           ca[1], ca[2], ca[3], yout, L, nchunks, nteams);
    else
        hipLaunchKernelGGL((cosenu_quad_stage<NW, false, false, false>), grid, block, 0, s, a,
                           ca[0], ca[1], ca[2], ca[3], yout, L, nchunks, nteams);
    return hipGetLastError();
}

template <int NW>
int resident_teams_nw()
{
    constexpr int TPB = NW >= 4 ? 1 : 4 / NW;
    int nb = 0, dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(&cosenu_quad_stage<NW, false, true, false>), 64 * NW * TPB, 0) !=
        hipSuccess)
        return 0;
    return nb * prop.multiProcessorCount * TPB;
}

} // namespace

int quad_supported(int VP, int bc, int scheme, int flags)
{
    return (VP == 64 || VP == 128 || VP == 256) && (bc == 0 || bc == 1) && scheme == 1 &&
           !(flags & kFlagKO5);
}

hipError_t launch_stage_quad(const StageArgs &a, const double ca[4], double *yout, int L,
                             int nteams, bool uniform, hipStream_t s)
{
    if (!quad_supported(a.VP, a.bc, a.scheme, a.flags) || a.nz < 16 || !yout)
        return hipErrorInvalidValue;
    if (nteams > 0) {
        if (nteams > a.nz)
            return hipErrorInvalidValue;
    } else if (L < 1 || a.nz % L != 0) {
        return hipErrorInvalidValue;
    }
    if (uniform && a.nvz != a.VP)
        return hipErrorInvalidValue;
    switch (a.VP) {
    case 64: return launch_nw<2>(a, ca, yout, L, nteams, uniform, s);
    case 128: return launch_nw<4>(a, ca, yout, L, nteams, uniform, s);
    case 256: return launch_nw<8>(a, ca, yout, L, nteams, uniform, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_stage_quad_batch(const BatchRun *runs, const int2 *items, int nitems, int VP,
                                   int L, bool uniform, bool mat, bool std_phys, hipStream_t s)
{
    if (L < 1)
        return hipErrorInvalidValue;
    switch (VP) {
    case 64: return launch_batch_nw<2>(runs, items, nitems, L, uniform, mat, std_phys, s);
    case 128: return launch_batch_nw<4>(runs, items, nitems, L, uniform, mat, std_phys, s);
    case 256: return launch_batch_nw<8>(runs, items, nitems, L, uniform, mat, std_phys, s);
    default: return hipErrorInvalidValue;
    }
}

int quad_resident_teams(int VP)
{
    switch (VP) {
    case 64: return resident_teams_nw<2>();
    case 128: return resident_teams_nw<4>();
    case 256: return resident_teams_nw<8>();
    default: return 0;
    }
}

const char *stage_quad_kernel_name()
{
    return "cosenu_quad_stage<*, *>";
}

int quad_wide_supported(int VP, int bc, int scheme, int flags)
{
    return VP == 512 && (bc == 0 || bc == 1) && scheme == 1 && !(flags & kFlagKO5);
}

size_t quad_wide_slot_words(int VP, int nteams)
{
    return (size_t)4 * nteams * (VP / 256) * kRing * 64 * 2;
}

int quad_wide_resident_blocks()
{
    int nb = 0, dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(&cosenu_quad_wide<false, true, false, 2>), 512, 0) !=
        hipSuccess)
        return 0;
    return nb * prop.multiProcessorCount;
}

hipError_t launch_stage_quad_wide(const StageArgs &a, const double ca[4], double *yout,
                                  int nteams, bool uniform, unsigned long long *slots,
                                  unsigned long long *ctl, bool give_up_test, hipStream_t s)
{
    if (!quad_wide_supported(a.VP, a.bc, a.scheme, a.flags) || a.nz < 16 || a.nz > (1 << 23) ||
        nteams < 1 || nteams > a.nz || !yout || !slots || !ctl)
        return hipErrorInvalidValue;
    if (uniform && a.nvz != a.VP)
        return hipErrorInvalidValue;
    long long ticks = 0;
    unsigned flip = 0;
    wide_wait_bound(give_up_test, &ticks, &flip);
    return launch_wide_nb<2>(a, ca, yout, nteams, uniform, slots, ctl, ticks, flip, s);
}

const char *stage_quad_wide_kernel_name()
{
    return "cosenu_quad_wide<*>";
}

} // namespace cosenu
