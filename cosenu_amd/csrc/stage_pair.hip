// FAST mode, two RK4 stages per sweep (variant 4): stages 1+2 in one launch,
// stages 3+4 in the next.
//
// The split-lane team of the stage kernel (VP/32 wavefronts covering all
// velocity bins of a z range, lanes 0-31 carrying ee, xx, ex_re, ex_im and
// lanes 32-63 the barred fields of 32 bins) marches along z with two register
// windows: the first stage's input, read from HBM, and the first stage's
// output, produced in registers.  The second stage runs 4 cells behind the
// first, so its window always holds first-stage values that are already
// computed; the intermediate state never goes through HBM unless a later
// stage needs it (Y2 for stage 4).  The first stage covers the team's cells
// plus 4 on each side (recomputed by both neighbours), which the second stage
// needs for its stencil.
//
// RK4 combination.  With Y2 = y + (dt/2) K1, Y3 = y + (dt/2) K2 and Y4 =
// y + dt K3 (nuosc.hpp:320-348), the new state is
//   y' = y + (dt/6)(K1 + 2 K2 + 2 K3 + K4) = y + ((Q + (Y4 - y))/3 + (dt/6) K4),
//   Q  = (Y2 - y) + 2 (Y3 - y),
// the stage-value form of the split-lane kernel.  Q is formed by the stages
// 1+2 launch, which holds y, Y2 and Y3 of the cell at once; the stages 3+4
// launch reads y at the first stage's cell (for Y4) and keeps it in a
// wavefront-private LDS ring until the second stage reaches that cell 4 z
// steps later.  Only the last addition rounds a state-sized value, as in the
// reference's update: a form that rounds state-sized products (such as
// (Y2 + 2 Y3 - y)/3 + Y4/3) drifts the conserved sums by ~1e-17 per step.
//
// Bytes per cell and step: stages 1+2 read y (the window) and write Q and Y3;
// stages 3+4 read Y3 (the window), y and Q and write the new state: 64 + 128
// + 192 + 64 = 448 B instead of 832 B for four separate stage launches (plus
// the halo rows of the windows).  The second stage of stages 1+2 takes y at
// its cell from the row that has just left the first window, kept in
// registers.
//
// The new state is written to a second buffer (the stages 3+4 launch reads y
// in neighbouring teams' halos while it runs), and the state buffers swap
// roles after every step.
//
// Open boundary (nuosc.hpp:247-277): stage-input rows below 0 read row 0 and
// rows above nz-1 extrapolate linearly, one after another, from the last two
// rows, where the window loads them.  The first stage's output at such ghost
// rows follows the same rule instead of being computed: below the grid, row
// 0's output is copied into the window positions of rows -4..-1 when it is
// produced; above it, each ghost row is 2 x the previous row - the one
// before.  The flux ghosts of the two boundary cells (rflux[-1] = rflux[0]
// for v > 0 at cell 0, lflux[nz] = 2 lflux[nz-1] - lflux[nz-2] for v < 0 at
// cell nz-1) keep the face before last of each stage in registers.
//
// Moments: the first stage's input moments are reduced when a row enters its
// window (as in the stage kernel); the second stage's input moments are
// reduced from the first stage's output row as it is produced.  Each goes
// through its own LDS ring of kRing rows and is read 4 iterations later.  A
// barrier every kSync iterations suffices: between the store of a row's
// moments and their use 4 iterations later lies at least one barrier, and so
// does between the last use of a ring slot and its next store (kRing - 5 >=
// kSync iterations apart), so the waves of a team drift up to kSync - 1
// iterations apart instead of meeting at every z step.
#include "fast_math.hpp"
#include "wide_exchange.hpp"

#include <cstdio>

namespace cosenu {
namespace {

using namespace fastk;

constexpr int kRing = 16, kSync = 4;
// the physics of the benchmark workloads: vacuum and collective terms with
// advection, no matter term (FV: the dissipation order does not apply)
constexpr int kStdPhys = kFlagVac | kFlagColl;
static_assert(kRing - 5 >= kSync && kSync <= 4, "moment ring too shallow for the barrier spacing");

// stages 3+4: y of the first stage's last 8 cells, per wavefront
// [slot = row % 8][field][lane]
constexpr int kYRing = 8 * 4 * 64;

// Teams of NB > 1 blocks (rows wider than 256 bins, cosenu_pair_wide): the
// per-row moment exchange between a team's blocks is in wide_exchange.hpp.

// A1: the first stage of the pair (1 or 3), A2: the second (2 or 4) of the
// same run.  PAIR 0: stages 1+2, PAIR 1: stages 3+4.  UNI: no padding bins
// and every wavefront's bins of one sign of v (the common case): no masks
// and no per-lane stencil selection.  OPENB: the open boundary.
// NB: blocks per team (1: the whole team in this block; > 1: this block holds
// bins [X.role * NW * 32, (X.role + 1) * NW * 32) of the team's rows).
// PHYS: the physics switches (kFlag* bits) fixed at compile time, or -1 to
// read them from the arguments (a fixed set drops the branches of the
// right-hand side and the values they carry: FV at nz = 65536 x 128 runs
// 3% faster so).
template <int NW, int SCHEME, int PAIR, bool UNI, bool OPENB, int NB = 1, int PHYS = -1>
__device__ __forceinline__ void pair_body(const StageArgs &A1, const StageArgs &A2, int L, int j0,
                                          int jend, bool active, double (*ra)[NW][8],
                                          double (*rb)[NW][8], double *yring,
                                          const WideX &X = WideX{}, double *xst = nullptr)
// ra, rb: [kRing][NW][8]; yring: the team's NW x kYRing doubles (stages 3+4);
// xst (NB > 1): [2 rings][NB - 1 partners][2 rows][64 elements], the partner
// blocks' partials staged in LDS
{
    // The right-hand side's terms and the RK4 combinations below write their
    // fused multiply-adds out and nothing else is fused here (the WENO7 code
    // keeps its own contraction): every instance of this body -- physics
    // switches fixed or read at run time, single runs or batches -- rounds a
    // cell alike, so batched runs stay bit-identical to single ones.
#pragma clang fp contract(off)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wit = wave % NW;
    const int h = lane >> 5;
    const int v = (NB > 1 ? X.role * NW * 32 : 0) + wit * 32 + (lane & 31);
    const int VP = A1.VP, nz = A1.nz;
    const size_t zs = (size_t)8 * VP;
    const double sig = h ? -1.0 : 1.0;
    j0 = __builtin_amdgcn_readfirstlane(j0);
    jend = __builtin_amdgcn_readfirstlane(jend);

    const double vi = A1.vz[v];
    const bool pos = vi > 0.0;
    const bool anyp = __any(vi > 0.0), anyn = __any(vi < 0.0);
    const int dir = UNI ? (anyp ? 1 : -1) : (anyn ? (anyp ? 0 : -1) : 1);
    const bool live = UNI || v < A1.nvz;
    const bool store = active && live;
    const int loff = 4 * h * VP + v; // this lane's first field within a z row
    double *const yr = yring + wit * kYRing; // this wavefront's part of the y ring

    auto zrow = [&](int m) -> int {
        int w = m < 0 ? m + nz : m; // m >= -9 > -nz
        while (w >= nz)
            w -= nz;
        return w;
    };
    // OPENB: rows below 0 read row 0 (the boundary rule); rows above nz-1
    // read row nz-1 as a placeholder, and the window replaces them by the
    // extrapolation when they enter it (ghost rows of y and Q are not used)
    auto load4 = [&](const double *base, int m, double o[4]) {
        const int row = OPENB ? min(max(m, 0), nz - 1) : zrow(m);
        const double *r = base + (size_t)row * zs + loff;
#pragma unroll
        for (int f = 0; f < 4; f++)
            o[f] = r[(size_t)f * VP];
    };

    const int flags = PHYS >= 0 ? PHYS : A1.flags;
    const bool coll = (flags & kFlagColl) != 0;
    const bool vac = (flags & kFlagVac) != 0, mat = (flags & kFlagMat) != 0;
    const bool advect = !(flags & kFlagAdvecOff);
    const bool ko5 = (flags & kFlagKO5) != 0;
    const double fac = -vi / A1.dz;
    const double fdfac = -vi / (12.0 * A1.dz);
    const double ko_eps = ko5 ? -A1.ko / A1.dz / 64.0 : -A1.ko / A1.dz / 16.0;
    const double mudv = A1.mu * A1.dv;
    const double st = A1.st * A1.pmo, ct = A1.ct * A1.pmo, st2 = 2.0 * st, ct2 = 2.0 * ct;

    // NB > 1: the device-memory slots of ring `rg` (0: first-stage input,
    // 1: second-stage input), block b, z row `row`
    // (32-bit word offsets: the rings of a launch hold at most a few million
    // words; the team's and the second ring's parts are formed once)
    const unsigned xteam = (unsigned)X.team * (NB * kRing * 128);
    const unsigned xring = (unsigned)X.nteams * (NB * kRing * 128);
    auto xslot = [&](int rg, int b, int row) -> unsigned long long * {
        return X.slots + (xteam + (rg ? xring : 0u) + (unsigned)b * (kRing * 128) +
                          (unsigned)(row & (kRing - 1)) * 128u);
    };
    bool gave_up = false;
    // partial moments of 4 field values of this lane into ring[row % kRing];
    // returns the lane's total (wave_sum6)
    auto moments = [&](double (*ring)[NW][8], int row, double p, double q, double r, double s) {
        const double S = live ? s : 0.0;
        const double R = live ? sig * r : 0.0;
        const double D = live ? sig * (p - q) : 0.0;
        const double m6[6] = {S, vi * S, R, vi * R, D, vi * D};
        const double t = wave_sum6(m6, lane);
        const int idx = (kMomentLane >> (4 * (lane >> 3))) & 15;
        if ((lane & 7) == 0 && idx < 6)
            ring[row & (kRing - 1)][wit][idx] = t;
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
    // NB > 1: the partner blocks' partials come in through LDS.  Wavefront
    // wit < 2 (NB - 1) stages ring lrg = wit / (NB - 1) from partner k =
    // wit % (NB - 1) + 1 (block role ^ k): it loads the words of the ring's
    // next row in the middle of a z step, checks their tags at its end and
    // writes the value to xst; the barrier that opens every z step orders
    // that write before the other wavefronts' reads.  The other wavefronts
    // issue no exchange loads, and no wavefront holds partner words in
    // registers across a step.
    constexpr int NX = NB > 1 ? NB - 1 : 1;
    const int lrg = wit / NX, lk = wit % NX + 1;
    const bool xld = NB > 1 && wit < 2 * NX;
    const bool xlane = (lane & 7) < 6; // elements 6, 7 of a wavefront's row are unused
    XWords xw = {0ull, 0ull};
    auto xst_at = [&](int rg, int k, int row) -> double & {
        return xst[((rg * NX + k - 1) * 2 + (row & 1)) * 64 + lane];
    };
    // this lane's element of a row's partials summed over the team's blocks,
    // its own block's from LDS (`own`), partner k's (k = 1 .. NB-1) staged.
    // The sum is the tree (own + p1) + (p2 + p3): as IEEE addition commutes,
    // every block of the team rounds it alike, (P0 + P1) + (P2 + P3) in
    // block numbers.
    auto xsum = [&](double own, int rg, int row) {
        double p[NB];
        p[0] = own;
#pragma unroll
        for (int k = 1; k < NB; k++)
            p[k] = xlane ? xst_at(rg, k, row) : 0.0;
        if constexpr (NB == 4)
            return (p[0] + p[1]) + (p[2] + p[3]);
        else if constexpr (NB == 2)
            return p[0] + p[1];
        else
            return p[0];
    };
    // face value from a window, by the wavefront's upwind direction
    auto face = [&](const double w[4][8], double F[4]) {
        if (dir > 0) {
#pragma unroll
            for (int f = 0; f < 4; f++)
                F[f] = flux_dir<1>(w[f], pos);
        } else if (UNI || dir < 0) {
#pragma unroll
            for (int f = 0; f < 4; f++)
                F[f] = flux_dir<-1>(w[f], pos);
        } else {
#pragma unroll
            for (int f = 0; f < 4; f++)
                F[f] = flux_dir<0>(w[f], pos);
        }
    };
    // right-hand side at the window centre (position 3): z row `row` (wrapped
    // into [0, nz)), mx = this lane's moment_partial of the row's ring slot
    auto rhs = [&](const double w[4][8], const double F[4], const double Fp[4], double mx,
                   int row, double K[4]) {
        const double p = w[0][3], qq = w[1][3], r = w[2][3], s = w[3][3];
#pragma unroll
        for (int f = 0; f < 4; f++)
            K[f] = !advect         ? 0.0
                   : SCHEME == 0 ? fac * (F[f] - Fp[f])
                                 : fdfac * fma(-8.0, w[f][2] - w[f][4], w[f][1] - w[f][5]);
        if (vac) {
            K[0] = fma(-st2, s, K[0]);
            K[1] = fma(st2, s, K[1]);
            K[2] = fma(-ct2, s, K[2]);
            K[3] = fma(ct2, r, fma(st, p - qq, K[3]));
        }
        if (mat) {
            const double hm = sig * A1.Hm[row];
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
        if (SCHEME == 1) {
#pragma unroll
            for (int f = 0; f < 4; f++) {
                const double *u = w[f];
                const double d = ko5 ? fma(-20.0, u[3], fma(15.0, u[2] + u[4],
                                                            fma(-6.0, u[1] + u[5], u[0] + u[6])))
                                     : fma(6.0, u[3], fma(-4.0, u[2] + u[4], u[1] + u[5]));
                K[f] = fma(ko_eps, d, K[f]);
            }
        }
    };

    // First stage: cells [j0-4, jend+4); second stage: cells [j0, jend),
    // 4 behind.  Iteration i: first stage at cell i, second at cell i-4;
    // i = j0-5 computes the first stage's upstream face only.
    const int ibeg = j0 - 5, iend = j0 + L + 4; // the same trip count in every team
    double wa[4][8], wb[4][8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        double t[4];
        load4(A1.yin, ibeg - 3 + k, t);
#pragma unroll
        for (int f = 0; f < 4; f++) {
            wa[f][k] = t[f];
            wb[f][k] = 0.0;
        }
    }
    {
        double t1 = 0.0, t2 = 0.0, t3 = 0.0;
        if (coll) {
            t1 = moments(ra, ibeg + 1, wa[0][4], wa[1][4], wa[2][4], wa[3][4]);
            t2 = moments(ra, ibeg + 2, wa[0][5], wa[1][5], wa[2][5], wa[3][5]);
            t3 = moments(ra, ibeg + 3, wa[0][6], wa[1][6], wa[2][6], wa[3][6]);
        }
        xpub(0, ibeg + 1, t1);
        xpub(0, ibeg + 2, t2);
        xpub(0, ibeg + 3, t3);
    }
    // The window's first rows have landed before the loop starts.  Without
    // this wait the compiler's wait-count analysis, which merges the loop
    // entry with the back edge, takes the last of them as possibly in flight
    // inside the loop and, as the y and Q loads of an iteration are
    // conditional, waits for all but 4 of the iteration's loads right after
    // issuing them (a memory latency per iteration in stages 3+4).
    // FD stages 1+2 (PF): the z step is short (~340 VALU instructions), so
    // the window row is loaded two iterations ahead (3.5% faster; stages 3+4,
    // with their y and Q loads, were 3.6% slower so).
    constexpr bool PF = SCHEME == 1 && PAIR == 0;
    double pre[4] = {0.0, 0.0, 0.0, 0.0}; // PF: the window row of the next iteration
    if (PF)
        load4(A1.yin, ibeg + 5, pre);
    __builtin_amdgcn_s_waitcnt(0);
    double Fa[4] = {0.0, 0.0, 0.0, 0.0}, Fb[4] = {0.0, 0.0, 0.0, 0.0};
    double Fa2[4] = {0.0, 0.0, 0.0, 0.0}, Fb2[4] = {0.0, 0.0, 0.0, 0.0}; // OPENB FV: faces before last
    double wprev[4] = {0.0, 0.0, 0.0, 0.0}; // the row before the window
    double yn[4] = {0.0, 0.0, 0.0, 0.0};    // stages 3+4: y of the next first-stage cell

    auto iteration = [&](const int i) {
        // NB > 1: is row r's total needed by the first (a) or second (b) stage
        auto xneed_a = [&](int r) {
            return coll && active && r >= j0 - 4 && r < jend + 4 && !(OPENB && (r < 0 || r >= nz));
        };
        auto xneed_b = [&](int r) { return coll && active && r >= j0 && r < jend; };
        // NB > 1: the row a loader wavefront stages during this iteration,
        // for the next: the first stage's row i + 1 (published by the
        // partner 3 iterations before), the second stage's row i - 3
        // (published by the partner's first stage 3 iterations before)
        const int xrow = lrg == 0 ? i + 1 : i - 3;
        double nxt[4];
        auto window_load = [&] {
            if (PF) {
#pragma unroll
                for (int f = 0; f < 4; f++)
                    nxt[f] = pre[f];
                load4(A1.yin, i + 6, pre);
            } else {
                load4(A1.yin, i + 5, nxt);
            }
        };
        if (NB == 1)
            window_load();
        const int jb = i - 4;
        const bool cell_a = i >= j0 - 4 && i < jend + 4;
        // OPENB: ghost rows of the first stage's output follow the boundary rule
        const bool ghost_a = OPENB && (i < 0 || i >= nz);
        // OPENB FV: one more face below the range, for the top cell's flux ghost
        const bool face_b = jb >= j0 - (OPENB ? 2 : 1) && jb < jend;
        const bool cell_b = jb >= j0 && jb < jend;
        // NB > 1: rows of partner words used by this iteration's stages
        const bool xa = NB > 1 && xneed_a(i), xb = NB > 1 && xneed_b(jb);
        if (NB > 1)
            window_load();
        const int ra_row = zrow(i), rb_row = zrow(jb);
        // pointwise operands: y at the first stage's cell for stage 3 (loaded
        // one iteration ahead, into yn, and kept in the LDS ring for stage
        // 4), Q at the second stage's cell for stage 4; stage 2 takes y at
        // its cell (row jb) from `wprev`, the row that left the first window
        // one iteration ago.
        // (qb is defined on every path: with an undefined value on the path
        // that skips its load, the optimiser folds the stage-4 combination
        // into the load's branch, and the wait for the load follows it at
        // once instead of at the end of the iteration)
        double ya[4], qb[4] = {0.0, 0.0, 0.0, 0.0};
        if (PAIR == 1) {
#pragma unroll
            for (int f = 0; f < 4; f++) {
                ya[f] = yn[f];
                yr[((i & 7) * 4 + f) * 64 + lane] = ya[f];
            }
            if (i + 1 >= j0 - 4 && i + 1 < jend + 4)
                load4(A1.y, i + 1, yn);
            if (cell_b)
                load4(A2.y2, jb, qb); // Q of the cell, written by stages 1+2
        }
        {
            double t0 = 0.0;
            if (coll)
                t0 = moments(ra, i + 4, wa[0][7], wa[1][7], wa[2][7], wa[3][7]);
            xpub(0, i + 4, t0);
        }
        double F[4] = {0.0, 0.0, 0.0, 0.0};
        if (SCHEME == 0 && advect)
            face(wa, F);

        // the same iterations in every team; every iteration with staging
        if (NB > 1 || ((i - ibeg) & (kSync - 1)) == 0)
            __syncthreads();
        // the two stages' input moments (rows stored 4 iterations ago, behind
        // a barrier), read before the first stage needs them
        double mxa = coll ? moment_partial<NW>(ra[i & (kRing - 1)], lane) : 0.0;
        double mxb = coll ? moment_partial<NW>(rb[jb & (kRing - 1)], lane) : 0.0;
        if constexpr (NB > 1) {
            if (xa)
                mxa = xsum(mxa, 0, i);
            if (xld) // issued before the window load
                xw = xload(xslot(lrg, X.role ^ lk, xrow) + 2 * lane);
        }

        // first stage at cell i: its output becomes the newest row of wb
        double oa[4] = {0.0, 0.0, 0.0, 0.0};
        double ta = 0.0; // the published total of the first stage's output row
        if (cell_a && !ghost_a) {
            double K[4], Fl[4], Fr[4];
#pragma unroll
            for (int f = 0; f < 4; f++) {
                Fl[f] = Fa[f];
                Fr[f] = F[f];
                if (OPENB && SCHEME == 0 && i == 0 && pos)
                    Fl[f] = F[f]; // rflux[-1] = rflux[0]
                if (OPENB && SCHEME == 0 && i == nz - 1 && !pos)
                    Fr[f] = Fa[f] * 2 - Fa2[f]; // lflux[nz], extrapolated
            }
            rhs(wa, Fr, Fl, mxa, ra_row, K);
#pragma unroll
            for (int f = 0; f < 4; f++)
                oa[f] = fma(A1.a, K[f], PAIR == 0 ? wa[f][3] : ya[f]);
            if (coll)
                ta = moments(rb, i, oa[0], oa[1], oa[2], oa[3]);
        } else if (cell_a && OPENB && i >= nz) {
#pragma unroll
            for (int f = 0; f < 4; f++)
                oa[f] = wb[f][7] * 2 - wb[f][6]; // rows i-1, i-2
        }
        xpub(1, i, ta);
#pragma unroll
        for (int f = 0; f < 4; f++) {
            if (OPENB)
                Fa2[f] = Fa[f];
            Fa[f] = F[f];
#pragma unroll
            for (int k = 0; k < 7; k++)
                wb[f][k] = wb[f][k + 1];
            wb[f][7] = oa[f];
        }
        if (OPENB && i == 0 && cell_a) {
            // rows -4..-1 of the first stage's output copy row 0
#pragma unroll
            for (int f = 0; f < 4; f++)
                wb[f][3] = wb[f][4] = wb[f][5] = wb[f][6] = oa[f];
        }
        // second stage at cell jb: window wb holds rows jb-3 .. jb+4
        if (face_b) {
            double G[4] = {0.0, 0.0, 0.0, 0.0};
            if (SCHEME == 0 && advect)
                face(wb, G);
            if constexpr (NB > 1) {
                if (xb)
                    mxb = xsum(mxb, 1, jb);
            }
            if (cell_b) {
                double K[4], Gl[4], Gr[4];
#pragma unroll
                for (int f = 0; f < 4; f++) {
                    Gl[f] = Fb[f];
                    Gr[f] = G[f];
                    if (OPENB && SCHEME == 0 && jb == 0 && pos)
                        Gl[f] = G[f];
                    if (OPENB && SCHEME == 0 && jb == nz - 1 && !pos)
                        Gr[f] = Fb[f] * 2 - Fb2[f];
                }
                rhs(wb, Gr, Gl, mxb, rb_row, K);
                if (store) {
                    double *o = A2.yout + (size_t)rb_row * zs + loff;
                    double *oq = A1.yout + (size_t)rb_row * zs + loff; // Q (stages 1+2)
#pragma unroll
                    for (int f = 0; f < 4; f++) {
                        if (PAIR == 0) {
                            // Y3 = y + (dt/2) K2; Q = (Y2 - y) + 2 (Y3 - y)
                            // with y from the departed window row, Y2 = the
                            // window centre
                            const double y0 = wprev[f];
                            const double y3 = fma(A2.a, K[f], y0);
                            o[(size_t)f * VP] = y3;
                            oq[(size_t)f * VP] = fma(2.0, y3 - y0, wb[f][3] - y0);
                        } else { // y' = y + ((Q + (Y4 - y))/3 + (dt/6) K4)
                            const double y0 = yr[((jb & 7) * 4 + f) * 64 + lane];
                            const double d = qb[f] + (wb[f][3] - y0);
                            o[(size_t)f * VP] = y0 + fma(A2.a, K[f], d * (1.0 / 3.0));
                        }
                    }
                }
            }
#pragma unroll
            for (int f = 0; f < 4; f++) {
                if (OPENB)
                    Fb2[f] = Fb[f];
                Fb[f] = G[f];
            }
        }
        if constexpr (NB > 1) { // (checked here: the load had half a z step to land)
            const int r = xrow;
            if (xld && (lrg == 0 ? xneed_a(r) : xneed_b(r)) && xlane)
                xst_at(lrg, lk, r) = xvalue(xw, xslot(lrg, X.role ^ lk, r) + 2 * lane,
                                            xtag(X.epoch, r) ^ X.tag_flip, X.ctl, X.wait_ticks,
                                            gave_up);
        }
        // OPENB: row i+5 above the grid extrapolates from rows i+4, i+3, one
        // row after another as the reference fills its ghost cells
        const bool extrap = OPENB && i + 5 >= nz;
#pragma unroll
        for (int f = 0; f < 4; f++) {
            const double in = extrap ? wa[f][7] * 2 - wa[f][6] : nxt[f];
            wprev[f] = wa[f][0];
#pragma unroll
            for (int k = 0; k < 7; k++)
                wa[f][k] = wa[f][k + 1];
            wa[f][7] = in;
        }
    };
    if (SCHEME == 1) {
        // Two iterations per trip: the window shifts of the first become
        // register renames (FD: 1.7% faster; FV: 1% slower, not used).  An
        // odd count ends with one idle iteration (i = iend: no cell of either
        // stage; its loads and moments are not used).
        const int trips = (iend - ibeg + 1) / 2;
        for (int t = 0; t < trips; t++) {
            iteration(ibeg + 2 * t);
            iteration(ibeg + 2 * t + 1);
        }
    } else {
        for (int i = ibeg; i < iend; i++)
            iteration(i);
    }
}

template <int NW, int SCHEME, int PAIR, bool UNI, bool OPENB, int PHYS = -1>
__global__ __launch_bounds__(NW >= 8 ? 512 : 256) __attribute__((amdgpu_waves_per_eu(2)))
void cosenu_pair_stage(StageArgs A1, StageArgs A2, int L, int nchunks, int nteams)
{
    constexpr int TPB = NW >= 4 ? 1 : 4 / NW; // teams per block
    __shared__ double ring_a[TPB][kRing][NW][8];
    __shared__ double ring_b[TPB][kRing][NW][8];
    __shared__ double yring[PAIR == 1 ? TPB * NW * kYRing : 1];
    const int team = (threadIdx.x >> 6) / NW;
    const int g = xcd_block_order() * TPB + team;
    double *const yr = yring + (PAIR == 1 ? team * NW * kYRing : 0);
    if (nteams > 0) {
        const bool active = g < nteams;
        const int j0 = active ? (int)(((long)g * A1.nz) / nteams) : 0;
        const int j1 = active ? (int)(((long)(g + 1) * A1.nz) / nteams) : 0;
        pair_body<NW, SCHEME, PAIR, UNI, OPENB, 1, PHYS>(A1, A2, L, j0, j1, active, ring_a[team],
                                                ring_b[team], yr);
    } else {
        const bool active = g < nchunks;
        const int j0 = (active ? g : 0) * L;
        pair_body<NW, SCHEME, PAIR, UNI, OPENB, 1, PHYS>(A1, A2, L, j0, j0 + L, active, ring_a[team],
                                                ring_b[team], yr);
    }
}

// Batched launch: work item k = (run, first z cell) of runs with the same VP
// and scheme, chunks of L rows (a run's last one takes what remains of its
// nz; every team runs the same L + 9 z steps).  `runs` holds each run's current state
// in .y and the spare in .acc (the host passes the array with the two
// swapped after every step).
template <int NW, int SCHEME, int PAIR, bool UNI, int PHYS = -1>
__global__ __launch_bounds__(NW >= 8 ? 512 : 256) __attribute__((amdgpu_waves_per_eu(2)))
void cosenu_pair_batch(const BatchRun *runs, const int2 *items, int nitems, int L)
{
    constexpr int TPB = NW >= 4 ? 1 : 4 / NW;
    __shared__ double ring_a[TPB][kRing][NW][8];
    __shared__ double ring_b[TPB][kRing][NW][8];
    __shared__ double yring[PAIR == 1 ? TPB * NW * kYRing : 1];
    const int team = (threadIdx.x >> 6) / NW;
    const int item = xcd_block_order() * TPB + team;
    const bool active = item < nitems;
    const int2 it = items[active ? item : 0];
    const BatchRun &r = runs[it.x];
    const StageArgs A1 = batch_stage_args(r, PAIR == 0 ? 1 : 3, true);
    StageArgs A2 = batch_stage_args(r, PAIR == 0 ? 2 : 4, true);
    if (PAIR == 1)
        A2.yout = r.acc;
    // a run's last item may be shorter than L (its nz need not be a multiple)
    const int jend = min(it.y + L, r.nz);
    pair_body<NW, SCHEME, PAIR, UNI, false, 1, PHYS>(A1, A2, L, it.y, jend, active, ring_a[team],
                                                     ring_b[team],
                                                     yring + (PAIR == 1 ? team * NW * kYRing : 0));
}

// Rows of 512 or 1024 bins: a team of NB = VP/256 blocks of 8 wavefronts per
// contiguous z range (range mode only: the grid is nteams * NB blocks, every
// team the same trip count).  Teams form by ticket (see WideX).
template <int SCHEME, int PAIR, bool UNI, bool OPENB, int NB, int PHYS = -1>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2)))
void cosenu_pair_wide(StageArgs A1, StageArgs A2, int nteams, unsigned long long *slots,
                      unsigned long long *ctl, long long wait_ticks, unsigned tag_flip)
{
    constexpr int NW = 8;
    __shared__ double ring_a[kRing][NW][8];
    __shared__ double ring_b[kRing][NW][8];
    __shared__ double yring[PAIR == 1 ? NW * kYRing : 1];
    __shared__ double xst[2 * (NB - 1) * 2 * 64];
    const WideX X = wide_join<NB>(slots, ctl, wait_ticks, tag_flip, nteams);
    const bool active = X.team < nteams;
    const int L = (A1.nz + nteams - 1) / nteams;
    const int j0 = active ? (int)(((long)X.team * A1.nz) / nteams) : 0;
    const int j1 = active ? (int)(((long)(X.team + 1) * A1.nz) / nteams) : 0;
    pair_body<NW, SCHEME, PAIR, UNI, OPENB, NB, PHYS>(A1, A2, L, j0, j1, active, ring_a, ring_b,
                                                      yring, X, xst);
}

template <int SCHEME, int NB>
hipError_t launch_wide_nb(const StageArgs &a1, const StageArgs &a2, int pair, int nteams,
                          bool uni, unsigned long long *slots, unsigned long long *ctl,
                          long long ticks, unsigned flip, hipStream_t s)
{
    dim3 grid(nteams * NB), block(512);
    if (a1.bc == 1 && pair == 0)
        hipLaunchKernelGGL((cosenu_pair_wide<SCHEME, 0, false, true, NB>), grid, block, 0, s, a1, a2, nteams, slots, ctl, ticks, flip);
    else if (a1.bc == 1)
        hipLaunchKernelGGL((cosenu_pair_wide<SCHEME, 1, false, true, NB>), grid, block, 0, s, a1, a2, nteams, slots, ctl, ticks, flip);
    else if (uni && SCHEME == 0 && a1.flags == kStdPhys && pair == 0)
        hipLaunchKernelGGL((cosenu_pair_wide<SCHEME, 0, true, false, NB, kStdPhys>), grid, block, 0, s, a1, a2, nteams, slots, ctl, ticks, flip);
    else if (uni && SCHEME == 0 && a1.flags == kStdPhys)
        hipLaunchKernelGGL((cosenu_pair_wide<SCHEME, 1, true, false, NB, kStdPhys>), grid, block, 0, s, a1, a2, nteams, slots, ctl, ticks, flip);
    else if (uni && pair == 0)
        hipLaunchKernelGGL((cosenu_pair_wide<SCHEME, 0, true, false, NB>), grid, block, 0, s, a1, a2, nteams, slots, ctl, ticks, flip);
    else if (uni)
        hipLaunchKernelGGL((cosenu_pair_wide<SCHEME, 1, true, false, NB>), grid, block, 0, s, a1, a2, nteams, slots, ctl, ticks, flip);
    else if (pair == 0)
        hipLaunchKernelGGL((cosenu_pair_wide<SCHEME, 0, false, false, NB>), grid, block, 0, s, a1, a2, nteams, slots, ctl, ticks, flip);
    else
        hipLaunchKernelGGL((cosenu_pair_wide<SCHEME, 1, false, false, NB>), grid, block, 0, s, a1, a2, nteams, slots, ctl, ticks, flip);
    return hipGetLastError();
}

template <int NW, int SCHEME>
hipError_t launch_batch_nw(const BatchRun *runs, const int2 *items, int nitems, int L, int pair,
                           bool uni, bool std_phys, hipStream_t s)
{
    constexpr int TPB = NW >= 4 ? 1 : 4 / NW;
    if (nitems <= 0)
        return hipSuccess;
    dim3 grid((nitems + TPB - 1) / TPB), block(64 * NW * TPB);
    if (uni && SCHEME == 0 && std_phys && pair == 0)
        hipLaunchKernelGGL((cosenu_pair_batch<NW, SCHEME, 0, true, kStdPhys>), grid, block, 0, s, runs, items, nitems, L);
    else if (uni && SCHEME == 0 && std_phys)
        hipLaunchKernelGGL((cosenu_pair_batch<NW, SCHEME, 1, true, kStdPhys>), grid, block, 0, s, runs, items, nitems, L);
    else if (uni && pair == 0)
        hipLaunchKernelGGL((cosenu_pair_batch<NW, SCHEME, 0, true>), grid, block, 0, s, runs, items, nitems, L);
    else if (uni)
        hipLaunchKernelGGL((cosenu_pair_batch<NW, SCHEME, 1, true>), grid, block, 0, s, runs, items, nitems, L);
    else if (pair == 0)
        hipLaunchKernelGGL((cosenu_pair_batch<NW, SCHEME, 0, false>), grid, block, 0, s, runs, items, nitems, L);
    else
        hipLaunchKernelGGL((cosenu_pair_batch<NW, SCHEME, 1, false>), grid, block, 0, s, runs, items, nitems, L);
    return hipGetLastError();
}

template <int NW, int SCHEME>
hipError_t launch_nw(const StageArgs &a1, const StageArgs &a2, int pair, int L, int nteams,
                     bool uni, hipStream_t s)
{
    constexpr int TPB = NW >= 4 ? 1 : 4 / NW;
    int nchunks = 0, work;
    if (nteams > 0) {
        L = (a1.nz + nteams - 1) / nteams;
        work = nteams;
    } else {
        nchunks = a1.nz / L;
        work = nchunks;
    }
    dim3 grid((work + TPB - 1) / TPB), block(64 * NW * TPB);
    // uniform: no padding bins and each group of 32 bins of one sign of v
    // (v0 = -v1 and nvz a multiple of 64 give that; `uniform` is the host's
    // check); the open boundary takes the general form
    if (a1.bc == 1 && pair == 0)
        hipLaunchKernelGGL((cosenu_pair_stage<NW, SCHEME, 0, false, true>), grid, block, 0, s, a1, a2, L, nchunks, nteams);
    else if (a1.bc == 1)
        hipLaunchKernelGGL((cosenu_pair_stage<NW, SCHEME, 1, false, true>), grid, block, 0, s, a1, a2, L, nchunks, nteams);
    else if (uni && SCHEME == 0 && a1.flags == kStdPhys && pair == 0)
        hipLaunchKernelGGL((cosenu_pair_stage<NW, SCHEME, 0, true, false, kStdPhys>), grid, block, 0, s, a1, a2, L, nchunks, nteams);
    else if (uni && SCHEME == 0 && a1.flags == kStdPhys)
        hipLaunchKernelGGL((cosenu_pair_stage<NW, SCHEME, 1, true, false, kStdPhys>), grid, block, 0, s, a1, a2, L, nchunks, nteams);
    else if (uni && pair == 0)
        hipLaunchKernelGGL((cosenu_pair_stage<NW, SCHEME, 0, true, false>), grid, block, 0, s, a1, a2, L, nchunks, nteams);
    else if (uni)
        hipLaunchKernelGGL((cosenu_pair_stage<NW, SCHEME, 1, true, false>), grid, block, 0, s, a1, a2, L, nchunks, nteams);
    else if (pair == 0)
        hipLaunchKernelGGL((cosenu_pair_stage<NW, SCHEME, 0, false, false>), grid, block, 0, s, a1, a2, L, nchunks, nteams);
    else
        hipLaunchKernelGGL((cosenu_pair_stage<NW, SCHEME, 1, false, false>), grid, block, 0, s, a1, a2, L, nchunks, nteams);
    return hipGetLastError();
}

template <int SCHEME>
hipError_t launch_scheme(const StageArgs &a1, const StageArgs &a2, int pair, int L, int nteams,
                         bool uni, hipStream_t s)
{
    switch (a1.VP) {
    case 64: return launch_nw<2, SCHEME>(a1, a2, pair, L, nteams, uni, s);
    case 128: return launch_nw<4, SCHEME>(a1, a2, pair, L, nteams, uni, s);
    case 256: return launch_nw<8, SCHEME>(a1, a2, pair, L, nteams, uni, s);
    default: return hipErrorInvalidValue;
    }
}

template <int NW, int SCHEME>
int resident_teams_nw()
{
    constexpr int TPB = NW >= 4 ? 1 : 4 / NW;
    int nb = 0, dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, reinterpret_cast<const void *>(&cosenu_pair_stage<NW, SCHEME, 1, false, false>),
            64 * NW * TPB, 0) != hipSuccess)
        return 0;
    return nb * prop.multiProcessorCount * TPB;
}

} // namespace

int pair_supported(int VP, int bc)
{
    return (VP == 64 || VP == 128 || VP == 256) && (bc == 0 || bc == 1);
}

hipError_t launch_stage_pair(const StageArgs &a1, const StageArgs &a2, int pair, int L,
                             int nteams, bool uniform, hipStream_t s)
{
    if (!pair_supported(a1.VP, a1.bc) || a1.nz < 16 || (pair != 0 && pair != 1))
        return hipErrorInvalidValue;
    if (nteams > 0) {
        if (nteams > a1.nz)
            return hipErrorInvalidValue;
    } else if (L < 1 || a1.nz % L != 0) {
        return hipErrorInvalidValue;
    }
    if (uniform && a1.nvz != a1.VP)
        return hipErrorInvalidValue;
    return a1.scheme == 0 ? launch_scheme<0>(a1, a2, pair, L, nteams, uniform, s)
                          : launch_scheme<1>(a1, a2, pair, L, nteams, uniform, s);
}

hipError_t launch_stage_pair_batch(const BatchRun *runs, const int2 *items, int nitems, int VP,
                                   int scheme, int L, int pair, bool uniform, bool std_phys,
                                   hipStream_t s)
{
    if (!pair_supported(VP, 0) || L < 1 || (pair != 0 && pair != 1))
        return hipErrorInvalidValue;
    const bool sp = std_phys;
    switch (VP) {
    case 64: return scheme == 0 ? launch_batch_nw<2, 0>(runs, items, nitems, L, pair, uniform, sp, s)
                                : launch_batch_nw<2, 1>(runs, items, nitems, L, pair, uniform, sp, s);
    case 128: return scheme == 0 ? launch_batch_nw<4, 0>(runs, items, nitems, L, pair, uniform, sp, s)
                                 : launch_batch_nw<4, 1>(runs, items, nitems, L, pair, uniform, sp, s);
    default: return scheme == 0 ? launch_batch_nw<8, 0>(runs, items, nitems, L, pair, uniform, sp, s)
                                : launch_batch_nw<8, 1>(runs, items, nitems, L, pair, uniform, sp, s);
    }
}

int pair_resident_teams(int VP, int scheme)
{
    switch (VP) {
    case 64: return scheme == 0 ? resident_teams_nw<2, 0>() : resident_teams_nw<2, 1>();
    case 128: return scheme == 0 ? resident_teams_nw<4, 0>() : resident_teams_nw<4, 1>();
    case 256: return scheme == 0 ? resident_teams_nw<8, 0>() : resident_teams_nw<8, 1>();
    default: return 0;
    }
}

int pair_wide_supported(int VP, int bc)
{
    return (VP == 512 || VP == 1024) && (bc == 0 || bc == 1);
}

size_t pair_wide_slot_words(int VP, int nteams)
{
    return (size_t)2 * nteams * (VP / 256) * kRing * 64 * 2;
}

int pair_wide_resident_blocks(int scheme)
{
    int nb = 0, dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return 0;
    const void *k = scheme == 0
                        ? reinterpret_cast<const void *>(&cosenu_pair_wide<0, 1, false, false, 2>)
                        : reinterpret_cast<const void *>(&cosenu_pair_wide<1, 1, false, false, 2>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, 512, 0) != hipSuccess)
        return 0;
    return nb * prop.multiProcessorCount;
}

hipError_t launch_stage_pair_wide(const StageArgs &a1, const StageArgs &a2, int pair, int nteams,
                                  bool uniform, unsigned long long *slots,
                                  unsigned long long *ctl, bool give_up_test, hipStream_t s)
{
    if (!pair_wide_supported(a1.VP, a1.bc) || a1.nz < 16 || a1.nz > (1 << 23) ||
        (pair != 0 && pair != 1) || nteams < 1 || nteams > a1.nz || !slots || !ctl)
        return hipErrorInvalidValue;
    if (uniform && a1.nvz != a1.VP)
        return hipErrorInvalidValue;
    long long ticks = 0;
    unsigned flip = 0;
    wide_wait_bound(give_up_test, &ticks, &flip);
    if (a1.VP == 512)
        return a1.scheme == 0
                   ? launch_wide_nb<0, 2>(a1, a2, pair, nteams, uniform, slots, ctl, ticks, flip, s)
                   : launch_wide_nb<1, 2>(a1, a2, pair, nteams, uniform, slots, ctl, ticks, flip, s);
    return a1.scheme == 0
               ? launch_wide_nb<0, 4>(a1, a2, pair, nteams, uniform, slots, ctl, ticks, flip, s)
               : launch_wide_nb<1, 4>(a1, a2, pair, nteams, uniform, slots, ctl, ticks, flip, s);
}

const char *stage_pair_wide_kernel_name(int pair)
{
    static char buf[64];
    snprintf(buf, sizeof(buf), "cosenu_pair_wide<*, %d>", pair);
    return buf;
}

const char *stage_pair_kernel_name(int pair)
{
    static char buf[64];
    snprintf(buf, sizeof(buf), "cosenu_pair_stage<*, *, %d>", pair);
    return buf;
}

} // namespace cosenu
