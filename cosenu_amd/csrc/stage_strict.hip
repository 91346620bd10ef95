// STRICT mode stage kernels: the reference's right-hand side evaluated in the
// reference's floating-point operation order, with true divisions and no
// fused multiply-add (this file is compiled with -ffp-contract=off and every
// function repeats `#pragma clang fp contract(off)`).  The result is
// bit-identical to the reference CPU build:
//   WENO7            src/rhs_fv.hpp:6-95
//   FV RHS           src/rhs_fv.hpp:115-236
//   FD RHS           src/rhs_fd.hpp:6-148 (ADVEC_CENTER_FD, KO_ORD_3 / 5th order)
//   ghost rules      src/nuosc.hpp:212-278 (periodic wrap, open extrapolation)
//   RK4 algebra      src/nuosc.hpp:280-348
//
// One thread per (z, v) cell.  A block holds ZB = max(1, 256/VP) z rows of VP
// lanes; the rows' 8-field columns are staged in LDS for the collective
// integral, which is the reference's sequential O(nvz) loop per cell.
#include "common.hpp"

namespace cosenu {
namespace {

__device__ inline int sgn_d(double x) { return (0.0 < x) - (x < 0.0); }

// u at z index m for the bin whose field row starts at `col` (stride 8*VP
// between z cells), applying the reference's ghost rule outside 0..nz-1.
__device__ inline double uval(const double *col, size_t zstride, int m, int nz, int bc)
{
#pragma clang fp contract(off)
    if (m >= 0 && m < nz)
        return col[(size_t)m * zstride];
    if (bc == 0) { // periodic copy (nuosc.hpp:223-242)
        int w = m % nz;
        if (w < 0)
            w += nz;
        return col[(size_t)w * zstride];
    }
    if (m < 0) // open: lower ghosts copy the first cell (nuosc.hpp:255-263)
        return col[0];
    // open: upper ghosts extrapolated one after another (nuosc.hpp:266-274)
    double a = col[(size_t)(nz - 2) * zstride], b = col[(size_t)(nz - 1) * zstride];
    for (int k = nz; k <= m; k++) {
        double c = b * 2 - a;
        a = b;
        b = c;
    }
    return b;
}

__device__ inline void smooth(const double *u, double si[4])
{
#pragma clang fp contract(off)
    si[0] = u[0] * (2107 * u[0] - 9402 * u[1] + 7042 * u[2] - 1854 * u[3])
          + u[1] * (11003 * u[1] - 17246 * u[2] + 4642 * u[3])
          + u[2] * (7043 * u[2] - 3882 * u[3]) + 547 * u[3] * u[3];
    si[1] = u[-1] * (547 * u[-1] - 2522 * u[0] + 1922 * u[1] - 494 * u[2])
          + u[0] * (3443 * u[0] - 5966 * u[1] + 1602 * u[2])
          + u[1] * (2843 * u[1] - 1642 * u[2]) + 267 * u[2] * u[2];
    si[2] = u[-2] * (267 * u[-2] - 1642 * u[-1] + 1602 * u[0] - 494 * u[1])
          + u[-1] * (2843 * u[-1] - 5966 * u[0] + 1922 * u[1])
          + u[0] * (3443 * u[0] - 2522 * u[1]) + 547 * u[1] * u[1];
    si[3] = u[-3] * (547 * u[-3] - 3882 * u[-2] + 4642 * u[-1] - 1854 * u[0])
          + u[-2] * (7043 * u[-2] - 17246 * u[-1] + 7042 * u[0])
          + u[-1] * (11003 * u[-1] - 9402 * u[0]) + 2107 * u[0] * u[0];
}

__device__ inline void weights(const double si[4], double g0, double g1, double g2, double g3,
                               double w[4])
{
#pragma clang fp contract(off)
    const double eps = 1E-6;
    double d0 = eps + si[0], d1 = eps + si[1], d2 = eps + si[2], d3 = eps + si[3];
    double a0 = g0 / (d0 * d0), a1 = g1 / (d1 * d1), a2 = g2 / (d2 * d2), a3 = g3 / (d3 * d3);
    double s = a0 + a1 + a2 + a3;
    w[0] = a0 / s;
    w[1] = a1 / s;
    w[2] = a2 / s;
    w[3] = a3 / s;
}

// rflux: left-biased value at m+1/2; u points at cell m, u[-3..3] valid.
__device__ inline double right_going(const double *u)
{
#pragma clang fp contract(off)
    double si[4], w[4];
    smooth(u, si);
    weights(si, 4. / 35., 18. / 35., 12. / 35., 1. / 35., w);
    double c0 = (1. / 4.) * u[0] + (13. / 12.) * u[1] - (5. / 12.) * u[2] + (1. / 12.) * u[3];
    double c1 = (-1. / 12.) * u[-1] + (7. / 12.) * u[0] + (7. / 12.) * u[1] - (1. / 12.) * u[2];
    double c2 = (1. / 12.) * u[-2] - (5. / 12.) * u[-1] + (13. / 12.) * u[0] + (1. / 4.) * u[1];
    double c3 = (-1. / 4.) * u[-3] + (13. / 12.) * u[-2] - (23. / 12.) * u[-1] + (25. / 12.) * u[0];
    return w[0] * c0 + w[1] * c1 + w[2] * c2 + w[3] * c3;
}

// lflux: right-biased value at m-1/2; u points at cell m, u[-3..3] valid.
__device__ inline double left_going(const double *u)
{
#pragma clang fp contract(off)
    double si[4], w[4];
    smooth(u, si);
    weights(si, 1. / 35., 12. / 35., 18. / 35., 4. / 35., w);
    double c0 = (-1. / 4.) * u[3] + (13. / 12.) * u[2] - (23. / 12.) * u[1] + (25. / 12.) * u[0];
    double c1 = (1. / 12.) * u[2] - (5. / 12.) * u[1] + (13. / 12.) * u[0] + (1. / 4.) * u[-1];
    double c2 = (-1. / 12.) * u[1] + (7. / 12.) * u[0] + (7. / 12.) * u[-1] - (1. / 12.) * u[-2];
    double c3 = (1. / 4.) * u[0] + (13. / 12.) * u[-1] - (5. / 12.) * u[-2] + (1. / 12.) * u[-3];
    return w[0] * c0 + w[1] * c1 + w[2] * c2 + w[3] * c3;
}

struct Cell {
    double ee, xx, exr, exi, bee, bxx, bexr, bexi;
};

// The reference's right-hand side at cell (z, v) into o[8].  `col0` points at
// field 0 of bin v at z = 0; sh is this z row's 8 columns in LDS.
__device__ void rhs_cell(const RhsArgs &A, const double *col0, const double *sh, int z, int v,
                         const Cell &c, double o[8])
{
#pragma clang fp contract(off)
    const int nz = A.nz, VP = A.VP, flags = A.flags;
    const size_t zs = (size_t)8 * VP;
    const double vi = A.vz[v];
    for (int f = 0; f < 8; f++)
        o[f] = 0;

    if (A.scheme == 0) {
        // ---------------- FV: out = 0 + adv + vac + mat + dv*I -------------
        if (!(flags & kFlagAdvecOff)) {
            const double fac = -vi / A.dz;
            const int s = sgn_d(vi);
            for (int f = 0; f < 8; f++) {
                const double *cf = col0 + (size_t)f * VP;
                double u[9]; // u[k] = value at z - 4 + k
                for (int k = 0; k < 9; k++)
                    u[k] = uval(cf, zs, z - 4 + k, nz, A.bc);
                double adv;
                if (s > 0) {
                    // fac*fabs(1+s)/2 * (rflux[j] - rflux[j-1]); the lflux term is +-0.
                    double rj = right_going(u + 4);
                    double rm = (A.bc == 1 && z == 0) ? rj : right_going(u + 3);
                    adv = fac * fabs((double)(1 + s)) / 2 * (rj - rm);
                } else if (s < 0) {
                    double lj = left_going(u + 4);
                    double lp;
                    if (A.bc == 1 && z == nz - 1) {
                        // lflux[nz] = lflux[nz-1]*2 - lflux[nz-2] (open ghost rule)
                        double lm = left_going(u + 3);
                        lp = lj * 2 - lm;
                    } else {
                        lp = left_going(u + 5);
                    }
                    adv = fac * fabs((double)(1 - s)) / 2 * (lp - lj);
                } else {
                    adv = 0.0; // v = 0: both selectors are zero
                }
                o[f] += adv;
            }
        }
        if (flags & kFlagVac) {
            const double st = A.st, ct = A.ct, pmo = A.pmo;
            o[0] += -2.0 * st * c.exi * pmo;
            o[1] += 2.0 * st * c.exi * pmo;
            o[2] += -2.0 * ct * c.exi * pmo;
            o[3] += (2.0 * ct * c.exr + st * (c.ee - c.xx)) * pmo;
            o[4] += -2.0 * st * c.bexi * pmo;
            o[5] += 2.0 * st * c.bexi * pmo;
            o[6] += -2.0 * ct * c.bexi * pmo;
            o[7] += (2.0 * ct * c.bexr + st * (c.bee - c.bxx)) * pmo;
        }
    } else {
        // ---------------- FD: out = vac; += mat; += adv; += dv*I; += KO -----
        if (flags & kFlagVac) {
            const double st = A.st, ct = A.ct, pmo = A.pmo;
            o[0] = -pmo * 2 * st * c.exi;
            o[1] = pmo * 2 * st * c.exi;
            o[2] = -pmo * 2 * ct * c.exi;
            o[3] = pmo * (2 * ct * c.exr + st * (c.ee - c.xx));
            o[4] = -pmo * 2 * st * c.bexi;
            o[5] = pmo * 2 * st * c.bexi;
            o[6] = -pmo * 2 * ct * c.bexi;
            o[7] = pmo * (2 * ct * c.bexr + st * (c.bee - c.bxx));
        }
    }
    if (flags & kFlagMat) {
        const double h = A.Hm[z];
        o[0] += 0.0;
        o[1] += 0.0;
        o[2] += h * c.exi;
        o[3] += -h * c.exr;
        o[4] += 0.0;
        o[5] += 0.0;
        o[6] += -h * c.bexi;
        o[7] += h * c.bexr;
    }
    if (A.scheme == 1 && !(flags & kFlagAdvecOff)) {
        const double factor = -vi / (12 * A.dz);
        for (int f = 0; f < 8; f++) {
            const double *cf = col0 + (size_t)f * VP;
            double xm2 = uval(cf, zs, z - 2, nz, A.bc), xm1 = uval(cf, zs, z - 1, nz, A.bc);
            double xp1 = uval(cf, zs, z + 1, nz, A.bc), xp2 = uval(cf, zs, z + 2, nz, A.bc);
            o[f] += factor * ((xm2 - xp2) - 8.0 * (xm1 - xp1));
        }
    }
    if (flags & kFlagColl) {
        // Sequential k loop in the reference's expression order
        // (rhs_fv.hpp:189-222; rhs_fd.hpp:86-116 is the same arithmetic).
        double Iee = 0.0, Ixx = 0.0, Iexr = 0.0, Iexi = 0.0;
        double Ibee = 0.0, Ibxx = 0.0, Ibexr = 0.0, Ibexi = 0.0;
        const double mu = A.mu;
        for (int k = 0; k < A.nvz; k++) {
            double eep = sh[0 * VP + k], xxp = sh[1 * VP + k];
            double expr = sh[2 * VP + k], expi = sh[3 * VP + k];
            double beep = sh[4 * VP + k], bxxp = sh[5 * VP + k];
            double bexpr = sh[6 * VP + k], bexpi = sh[7 * VP + k];
            double w = 1.0 - vi * A.vz[k];
            double S = expi + bexpi;
            double Rr = expr - bexpr;
            double D = eep - xxp - beep + bxxp;
            Iee += 2.0 * mu * w * (c.exr * S - c.exi * Rr);
            Ixx += -2.0 * mu * w * (c.exr * S - c.exi * Rr);
            Iexr += 1.0 * mu * w * ((c.xx - c.ee) * S + c.exi * D);
            Iexi += 1.0 * mu * w * (-(c.xx - c.ee) * Rr - c.exr * D);
            Ibee += 2.0 * mu * w * (-c.bexr * (-expi - bexpi) + c.bexi * Rr);
            Ibxx += -2.0 * mu * w * (-c.bexr * (-expi - bexpi) + c.bexi * Rr);
            Ibexr += 1.0 * mu * w * (-(c.bee - c.bxx) * S - c.bexi * D);
            Ibexi += 1.0 * mu * w * ((c.bee - c.bxx) * (bexpr - expr) + c.bexr * D);
        }
        const double dv = A.dv;
        o[0] += dv * Iee;
        o[1] += dv * Ixx;
        o[2] += dv * Iexr;
        o[3] += dv * Iexi;
        o[4] += dv * Ibee;
        o[5] += dv * Ibxx;
        o[6] += dv * Ibexr;
        o[7] += dv * Ibexi;
    }
    if (A.scheme == 1) {
        const bool ko5 = (flags & kFlagKO5) != 0;
        const double ko_eps = ko5 ? -A.ko / A.dz / 64.0 : -A.ko / A.dz / 16.0;
        for (int f = 0; f < 8; f++) {
            const double *cf = col0 + (size_t)f * VP;
            double x[7];
            for (int k = 0; k < 7; k++)
                x[k] = (ko5 || (k >= 1 && k <= 5)) ? uval(cf, zs, z - 3 + k, nz, A.bc) : 0.0;
            double d = ko5 ? (x[0] + x[6] - 6 * (x[1] + x[5]) + 15 * (x[2] + x[4]) - 20 * x[3])
                           : (x[1] + x[5] - 4 * (x[2] + x[4]) + 6 * x[3]);
            o[f] += ko_eps * d;
        }
    }
}

template <bool kStage>
__device__ void strict_body(const RhsArgs &A, const StageArgs *S, int stage)
{
#pragma clang fp contract(off)
    extern __shared__ double sh[]; // [ZB][8][VP]
    const int VP = A.VP;
    const int zb = blockDim.x / VP;
    const int zz = threadIdx.x / VP, v = threadIdx.x % VP;
    const int z = blockIdx.x * zb + zz;
    const bool live = z < A.nz;
    double *row = sh + (size_t)zz * 8 * VP;
    Cell c = {0, 0, 0, 0, 0, 0, 0, 0};
    if (live) {
        const double *src = A.in + row_index(z, 0, VP) + v;
        c.ee = src[0 * VP]; c.xx = src[1 * VP]; c.exr = src[2 * VP]; c.exi = src[3 * VP];
        c.bee = src[4 * VP]; c.bxx = src[5 * VP]; c.bexr = src[6 * VP]; c.bexi = src[7 * VP];
        row[0 * VP + v] = c.ee; row[1 * VP + v] = c.xx; row[2 * VP + v] = c.exr;
        row[3 * VP + v] = c.exi; row[4 * VP + v] = c.bee; row[5 * VP + v] = c.bxx;
        row[6 * VP + v] = c.bexr; row[7 * VP + v] = c.bexi;
    }
    __syncthreads();
    if (!live || v >= A.nvz)
        return;
    double o[8];
    rhs_cell(A, A.in + v, row, z, v, c, o);
    const size_t base = row_index(z, 0, VP) + v;
    if (!kStage) {
        for (int f = 0; f < 8; f++)
            A.out[base + (size_t)f * VP] = o[f];
        return;
    }
    const double a = S->a;
    for (int f = 0; f < 8; f++) {
        const size_t q = base + (size_t)f * VP;
        const double yv = S->y[q];
        if (stage == 1) {
            S->acc[q] = o[f];
            S->yout[q] = yv + a * o[f];
        } else if (stage == 4) {
            S->yout[q] = yv + a * (o[f] + S->acc[q]);
        } else {
            S->acc[q] = S->acc[q] + 2.0 * o[f];
            S->yout[q] = yv + a * o[f];
        }
    }
}

RhsArgs to_rhs(const StageArgs &s)
{
    RhsArgs r;
    r.in = s.yin;
    r.out = nullptr;
    r.vz = s.vz;
    r.Hm = s.Hm;
    r.nz = s.nz; r.nvz = s.nvz; r.VP = s.VP; r.flags = s.flags; r.bc = s.bc; r.scheme = s.scheme;
    r.dz = s.dz; r.dv = s.dv; r.ct = s.ct; r.st = s.st; r.pmo = s.pmo; r.mu = s.mu; r.ko = s.ko;
    return r;
}

template <int kStage>
__global__ __launch_bounds__(1024) void cosenu_strict_stage(RhsArgs A, StageArgs S)
{
    strict_body<true>(A, &S, kStage);
}

__global__ __launch_bounds__(1024) void cosenu_strict_rhs(RhsArgs A)
{
    strict_body<false>(A, nullptr, 0);
}

void shape(int VP, int nz, dim3 &grid, dim3 &block, size_t &lds)
{
    int zb = VP >= 256 ? 1 : 256 / VP;
    block = dim3(zb * VP);
    grid = dim3((nz + zb - 1) / zb);
    lds = (size_t)zb * 8 * VP * sizeof(double);
}

} // namespace

hipError_t launch_stage_strict(const StageArgs &s, int stage, hipStream_t st)
{
    if (s.VP > 1024 || s.VP % 64)
        return hipErrorInvalidValue;
    dim3 g, b;
    size_t lds;
    shape(s.VP, s.nz, g, b, lds);
    RhsArgs r = to_rhs(s);
    switch (stage) {
    case 1: hipLaunchKernelGGL(cosenu_strict_stage<1>, g, b, lds, st, r, s); break;
    case 2: hipLaunchKernelGGL(cosenu_strict_stage<2>, g, b, lds, st, r, s); break;
    case 3: hipLaunchKernelGGL(cosenu_strict_stage<2>, g, b, lds, st, r, s); break;
    case 4: hipLaunchKernelGGL(cosenu_strict_stage<4>, g, b, lds, st, r, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_rhs_strict(const RhsArgs &r, hipStream_t st)
{
    if (r.VP > 1024 || r.VP % 64)
        return hipErrorInvalidValue;
    dim3 g, b;
    size_t lds;
    shape(r.VP, r.nz, g, b, lds);
    hipLaunchKernelGGL(cosenu_strict_rhs, g, b, lds, st, r);
    return hipGetLastError();
}

const char *stage_strict_kernel_name(int stage)
{
    static const char *names[4] = {"cosenu_strict_stage<1>", "cosenu_strict_stage<2>",
                                   "cosenu_strict_stage<2>", "cosenu_strict_stage<4>"};
    return (stage >= 0 && stage < 4) ? names[stage] : "";
}

} // namespace cosenu
