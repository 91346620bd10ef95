// Device helpers shared by the FAST stage kernels: the upwind WENO7
// reconstruction, the wavefront reduction of the per-z moments and the
// XCD-aware block order.
#pragma once

#include "common.hpp"

namespace cosenu {
namespace fastk {

__device__ inline double wave_sum(double x)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
        x += __shfl_xor(x, o, 64);
    return x;
}

// eps + beta for a smoothness indicator written as a quadratic form in three
// consecutive differences d1, d2, d3 of its 4-cell stencil:
//   beta = c11 d1^2 + c12 d1 d2 + c13 d1 d3 + c22 d2^2 + c23 d2 d3 + c33 d3^2.
__device__ __forceinline__ double eps_beta(double c11, double c12, double c13, double c22,
                                           double c23, double c33, double eps, double d1,
                                           double d2, double d3)
{
    const double t1 = fma(c13, d3, fma(c12, d2, c11 * d1));
    const double t2 = fma(c23, d3, c22 * d2);
    return fma(d1, t1, fma(d2, t2, fma(d3, c33 * d3, eps)));
}

// Upwind WENO7 value from the upwind-ordered stencil a0..a6 (a3 = the cell
// on the upwind side of the interface).
//
// Same mathematics as rhs_fv.hpp:49-80, evaluated in difference form.  The
// reference's smoothness indicators are quadratic forms in the cell values
// whose coefficient rows sum to zero, so each is exactly a quadratic form in
// the three differences D_i = a(i+1) - a(i) of its stencil (coefficients
// obtained from the reference's by exact rational substitution); the
// candidate values are a3 plus a combination of differences with weights
// (1/12)(9,-4,1), (1/12)(1,6,-1), (1/12)(-1,4,3), (1/12)(3,-10,13).  Both forms
// are equal in exact arithmetic; this one needs fewer operations and avoids
// the cancellation of the expanded form (terms of size 1e4 u^2 summing to a
// beta of size (du)^2).
//
// Nonlinear weights w_k = (g_k / q_k) / sum_m (g_m / q_m), q_k = (eps +
// beta_k)^2, linear weights g = (4, 18, 12, 1)/35.  With r_k = q_k / g_k,
// w_k = prod_{m != k} r_m / sum_k prod_{m != k} r_m: one division.  r_k is
// formed as ((eps + beta_k) / sqrt(g_k))^2 with 1/sqrt(g_k) folded into the
// indicator's coefficients (the common factor 35 cancels).
__device__ inline double weno_up(double a0, double a1, double a2, double a3, double a4, double a5,
                                 double a6)
{
    constexpr double eps = 1e-6;
    constexpr double h1 = 0.23570226039551584; // 1/sqrt(18)
    constexpr double h2 = 0.28867513459481287; // 1/sqrt(12)
    const double D0 = a1 - a0, D1 = a2 - a1, D2 = a3 - a2, D3 = a4 - a3, D4 = a5 - a4,
                 D5 = a6 - a5;
    // (eps + beta_k) / sqrt(g_k); 1/sqrt(4) = 0.5 and 1/sqrt(1) are exact
    const double t0 = eps_beta(0.5 * 2107, 0.5 * -5188, 0.5 * 1854, 0.5 * 3708, 0.5 * -2788,
                               0.5 * 547, 0.5 * eps, D3, D4, D5);
    const double t1 = eps_beta(h1 * 547, h1 * -1428, h1 * 494, h1 * 1468, h1 * -1108, h1 * 267,
                               h1 * eps, D2, D3, D4);
    const double t2 = eps_beta(h2 * 267, h2 * -1108, h2 * 494, h2 * 1468, h2 * -1428, h2 * 547,
                               h2 * eps, D1, D2, D3);
    const double t3 = eps_beta(547, -2788, 1854, 3708, -5188, 2107, eps, D0, D1, D2);
    // 12 x (candidate value - a3)
    const double c0 = fma(9.0, D3, fma(-4.0, D4, D5));
    const double c1 = fma(6.0, D3, D2 - D4);
    const double c2 = fma(4.0, D2, fma(3.0, D3, -D1));
    const double c3 = fma(3.0, D0, fma(-10.0, D1, 13.0 * D2));
    const double r0 = t0 * t0, r1 = t1 * t1, r2 = t2 * t2, r3 = t3 * t3;
    const double p01 = r0 * r1, p23 = r2 * r3;
    const double e0 = r1 * p23, e1 = r0 * p23, e2 = r3 * p01, e3 = r2 * p01;
    const double num = e0 * c0 + e1 * c1 + e2 * c2 + e3 * c3;
    const double den = (e0 + e1) + (e2 + e3);
    // num/den: hardware reciprocal refined by one Newton step (den is a
    // positive normal number).  The quotient is a correction to a3 of the
    // size of the local differences, so its last-bit error is far below an
    // ulp of the result.
    double r = __builtin_amdgcn_rcp(den);
    r = fma(r, fma(-den, r, 1.0), r);
    return fma(num * r, 1.0 / 12.0, a3);
}

// Upwind face value from a window w[k] = u(j-3+k), k = 0..7: DIR > 0 the
// left-biased value at j+1/2 from u(j-3..j+3), DIR < 0 the right-biased one
// from the mirrored stencil u(j+4..j-2), DIR == 0 per lane by `pos`.
template <int DIR>
__device__ inline double flux_dir(const double w[8], bool pos)
{
    if (DIR > 0)
        return weno_up(w[0], w[1], w[2], w[3], w[4], w[5], w[6]);
    if (DIR < 0)
        return weno_up(w[7], w[6], w[5], w[4], w[3], w[2], w[1]);
    return weno_up(pos ? w[0] : w[7], pos ? w[1] : w[6], pos ? w[2] : w[5], pos ? w[3] : w[4],
                   pos ? w[4] : w[3], pos ? w[5] : w[2], pos ? w[6] : w[1]);
}

// Cross-lane moves of a 64-bit value, one 32-bit half at a time.
//   swap32(a, b): a <- [a(0..31) | b(0..31)], b <- [a(32..63) | b(32..63)]
//   swap16(a, b): rows of 16 lanes, a <- [a0 b0 a2 b2], b <- [a1 b1 a3 b3]
__device__ inline void swap32(double &a, double &b)
{
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(a),
                                                     (unsigned)__double2loint(b), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(a),
                                                     (unsigned)__double2hiint(b), false, false);
    a = __hiloint2double((int)hi[0], (int)lo[0]);
    b = __hiloint2double((int)hi[1], (int)lo[1]);
}

__device__ inline void swap16(double &a, double &b)
{
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(a),
                                                     (unsigned)__double2loint(b), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(a),
                                                     (unsigned)__double2hiint(b), false, false);
    a = __hiloint2double((int)hi[0], (int)lo[0]);
    b = __hiloint2double((int)hi[1], (int)lo[1]);
}

// x from another lane of the same row of 16 (DPP control CTRL).  The move is
// one 64-bit DPP instruction for the row broadcasts (row_newbcast, the one
// control gfx950 takes on 64-bit operands) and two 32-bit ones otherwise.
template <int CTRL>
__device__ inline double dpp(double x)
{
    long v = __builtin_bit_cast(long, x);
    long r = __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true);
    return __builtin_bit_cast(double, r);
}

// Sum of 6 per-lane values m[0..5] over the 64 lanes.  Each level halves the
// lanes a value is spread over and, for the first three, pairs two values in
// one register: lane swaps across the halves (32) and rows (16) of the
// wavefront, then row-local DPP moves (rotate by 8, half-row mirror, two quad
// permutations).  On return lane l holds the total of value
// kMomentLane[l >> 3] (values 6 and 7 mark duplicates, not to be stored).
constexpr unsigned kMomentLane = 0x74537120u; // nibble g: value in lanes 8g .. 8g+7

__device__ inline double wave_sum6(const double m[6], int lane)
{
    double p0 = m[0], p1 = m[1], p2 = m[2], p3 = m[3], p4 = m[4], p5 = m[5];
    swap32(p0, p3);
    p0 += p3; // lanes 0-31: value 0, lanes 32-63: value 3 (16 lanes apart still split)
    swap32(p1, p4);
    p1 += p4; // values 1 / 4
    swap32(p2, p5);
    p2 += p5; // values 2 / 5
    swap16(p0, p1);
    p0 += p1; // rows: values 0, 1, 3, 4
    double p2b = p2;
    swap16(p2, p2b);
    p2 += p2b; // rows: values 2, 2, 5, 5
    const bool hi8 = (lane >> 3) & 1;
    double x = (hi8 ? p2 : p0) + dpp<0x128>(hi8 ? p0 : p2); // row_ror:8
    x += dpp<0x141>(x);                                      // row_half_mirror
    x += dpp<0xB1>(x);                                       // quad_perm [1,0,3,2]
    x += dpp<0x4E>(x);                                       // quad_perm [2,3,0,1]
    return x;
}

// Totals of the 6 moments of one ring row over the NW wavefronts of a team:
// row[ww][0..5] holds wavefront ww's partial sums (slots 6 and 7 unused).
// One LDS read per lane (moment_partial: lane l reads element l of the row,
// repeated over the lanes beyond NW * 8; issued early, where the row is
// known to be complete), the wavefront partials of a moment summed in a
// fixed tree by lane moves (row rotate by 8, then row and half swaps:
// ((w0 + w1) + (w2 + w3)) + ((w4 + w5) + (w6 + w7))), after which every row
// of 16 lanes holds the totals in its lanes 0-5, broadcast to the row by
// DPP row_newbcast.  One LDS round trip instead of NW * 3 dependent ones.
template <int NW>
__device__ inline double moment_partial(const double (*row)[8], int lane)
{
    static_assert(NW == 2 || NW == 4 || NW == 8, "team of 2, 4 or 8 wavefronts");
    return (&row[0][0])[lane & (NW * 8 - 1)];
}

template <int NW>
__device__ inline void moment_totals(double x, double M[6])
{
    x += dpp<0x128>(x); // row_ror:8: lanes 0-7 of a row hold w(2r) + w(2r+1)
    if (NW >= 4) {
        double y = x;
        swap16(x, y);
        x += y;
    }
    if (NW >= 8) {
        double y = x;
        swap32(x, y);
        x += y;
    }
    M[0] = dpp<0x150>(x); // row_newbcast:m
    M[1] = dpp<0x151>(x);
    M[2] = dpp<0x152>(x);
    M[3] = dpp<0x153>(x);
    M[4] = dpp<0x154>(x);
    M[5] = dpp<0x155>(x);
}

// XCD-aware order: blocks are dealt round-robin to the 8 XCDs, so give each
// XCD a contiguous run of work (neighbouring chunks share halo rows in that
// XCD's L2).
__device__ inline int xcd_block_order()
{
    const int nb = gridDim.x, b = blockIdx.x;
    const int per = nb / 8, rem = nb % 8, x = b % 8, k = b / 8;
    return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + k;
}

} // namespace fastk
} // namespace cosenu
