"""The FAST kernels' WENO7 constants against the reference's scheme, exactly.

The FAST kernels (cosenu_amd/csrc/fast_math.hpp, `weno_up`) evaluate the
reference's WENO7 (src/rhs_fv.hpp:49-80) in difference form: each smoothness
indicator as a quadratic form in the three differences D_i = a(i+1) - a(i) of
its stencil, each candidate value as a3 + (1/12) * (combination of
differences), and the linear weights folded in as 1/sqrt(gamma_k) factors.
These tests read the constants out of the kernel source and check, in exact
rational arithmetic, that they define the reference's indicators and
candidate polynomials.  (The GPU tests check the kernels' numbers against the
oracle; these pin the constants themselves, so that a mistyped coefficient
fails here on the CPU.)
"""
import math
import os
import re
from fractions import Fraction as Fr

import pytest

SRC = os.path.join(os.path.dirname(__file__), "..", "cosenu_amd", "csrc", "fast_math.hpp")

# The reference's expanded indicators, src/rhs_fv.hpp:49-52, for the upwind-
# ordered stencil a0..a6 (beta_k uses a(3-k) .. a(6-k)): coefficient of
# a_i a_j (i <= j).
REF_BETA = [
    {(3, 3): 2107, (3, 4): -9402, (3, 5): 7042, (3, 6): -1854, (4, 4): 11003, (4, 5): -17246,
     (4, 6): 4642, (5, 5): 7043, (5, 6): -3882, (6, 6): 547},
    {(2, 2): 547, (2, 3): -2522, (2, 4): 1922, (2, 5): -494, (3, 3): 3443, (3, 4): -5966,
     (3, 5): 1602, (4, 4): 2843, (4, 5): -1642, (5, 5): 267},
    {(1, 1): 267, (1, 2): -1642, (1, 3): 1602, (1, 4): -494, (2, 2): 2843, (2, 3): -5966,
     (2, 4): 1922, (3, 3): 3443, (3, 4): -2522, (4, 4): 547},
    {(0, 0): 547, (0, 1): -3882, (0, 2): 4642, (0, 3): -1854, (1, 1): 7043, (1, 2): -17246,
     (1, 3): 7042, (2, 2): 11003, (2, 3): -9402, (3, 3): 2107},
]
# Candidate polynomials at the interface, src/rhs_fv.hpp:67-70.
REF_CAND = [
    {3: Fr(1, 4), 4: Fr(13, 12), 5: Fr(-5, 12), 6: Fr(1, 12)},
    {2: Fr(-1, 12), 3: Fr(7, 12), 4: Fr(7, 12), 5: Fr(-1, 12)},
    {1: Fr(1, 12), 2: Fr(-5, 12), 3: Fr(13, 12), 4: Fr(1, 4)},
    {0: Fr(-1, 4), 1: Fr(13, 12), 2: Fr(-23, 12), 3: Fr(25, 12)},
]
GAMMA = [4, 18, 12, 1]  # linear weights x 35, src/rhs_fv.hpp:20-28


def _source():
    with open(SRC) as f:
        return f.read()


def _kernel_indicators():
    """[(scale, [c11, c12, c13, c22, c23, c33], eps_scale, [d1, d2, d3])] per k."""
    src = _source()
    body = src[src.index("__device__ inline double weno_up"):]
    out = []
    for k in range(4):
        m = re.search(r"const double t%d = eps_beta\((.*?)\);" % k, body, re.S)
        args = [a.strip() for a in m.group(1).replace("\n", " ").split(",")]
        coef, scale = [], None
        for a in args[:6]:
            mm = re.fullmatch(r"(?:(0\.5|h1|h2) \* )?(-?\d+)", a)
            assert mm, a
            scale = mm.group(1)
            coef.append(int(mm.group(2)))
        ds = [int(re.fullmatch(r"D(\d)", a).group(1)) for a in args[7:10]]
        out.append((scale, coef, args[6], ds))
    return out


def _quadratic_form_in_differences(ref, base):
    """Exact 3x3 form N (d^T N d) equal to the 4-cell form `ref` on cells
    base..base+3, with d_p = a(base+p+1) - a(base+p)."""
    M = [[Fr(0)] * 4 for _ in range(4)]
    for (i, j), c in ref.items():
        i -= base
        j -= base
        if i == j:
            M[i][i] += c
        else:
            M[i][j] += Fr(c, 2)
            M[j][i] += Fr(c, 2)
    # the form vanishes on constants: rows sum to zero
    assert all(sum(r) == 0 for r in M)
    Lm = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [1, 1, 1]]  # a = a_base + L d
    return [[sum(Lm[a][p] * M[a][b] * Lm[b][q] for a in range(4) for b in range(4))
             for q in range(3)] for p in range(3)]


def test_indicator_coefficients_are_the_reference_forms():
    got = _kernel_indicators()
    for k, (scale, coef, eps_arg, ds) in enumerate(got):
        base = 3 - k
        # stencil of beta_k: cells base..base+3, differences D_base..D_base+2
        assert ds == [base, base + 1, base + 2]
        N = _quadratic_form_in_differences(REF_BETA[k], base)
        want = [N[0][0], 2 * N[0][1], 2 * N[0][2], N[1][1], 2 * N[1][2], N[2][2]]
        assert [Fr(c) for c in coef] == want, k


def test_weight_scales_are_inverse_square_roots_of_the_linear_weights():
    src = _source()
    h1 = float(re.search(r"constexpr double h1 = ([0-9.e-]+);", src).group(1))
    h2 = float(re.search(r"constexpr double h2 = ([0-9.e-]+);", src).group(1))
    assert h1 == pytest.approx(1 / math.sqrt(18), rel=1e-16)
    assert h2 == pytest.approx(1 / math.sqrt(12), rel=1e-16)
    scales = {None: 1.0, "0.5": 0.5, "h1": h1, "h2": h2}
    for k, (scale, _, eps_arg, _) in enumerate(_kernel_indicators()):
        s = scales[scale]
        assert s == pytest.approx(1 / math.sqrt(GAMMA[k]), rel=1e-15), k
        # eps is scaled like the indicator: (eps + beta_k) / sqrt(gamma_k)
        assert eps_arg == ("eps" if scale is None else f"{scale} * eps"), k


def _linear_in_differences(line):
    """{i: coefficient of D_i} of a `fma(c, Di, ...)` / `Di - Dj` / `c * Di`
    expression."""
    coef = {}
    for c, i in re.findall(r"fma\((-?\d+\.\d*), D(\d)", line):
        coef[int(i)] = coef.get(int(i), 0) + Fr(c)
    for c, i in re.findall(r"(-?\d+\.\d*) \* D(\d)", line):
        coef[int(i)] = coef.get(int(i), 0) + Fr(c)
    rest = re.sub(r"fma\(-?\d+\.\d*, D\d,|-?\d+\.\d* \* D\d", "", line)
    for sign, i in re.findall(r"(-?)\s*D(\d)", rest):
        coef[int(i)] = coef.get(int(i), 0) + (-1 if sign else 1)
    return coef


def test_candidate_combinations_are_the_reference_polynomials():
    src = _source()
    for k in range(4):
        line = re.search(r"const double c%d = (.*?);" % k, src).group(1)
        got = _linear_in_differences(line)
        # reference candidate minus a3, written in differences: a_j - a3 is
        # the sum of D_3..D_(j-1) (j > 3) or minus D_j..D_2 (j < 3)
        want = {}
        for j, w in REF_CAND[k].items():
            if j > 3:
                for i in range(3, j):
                    want[i] = want.get(i, 0) + w
            elif j < 3:
                for i in range(j, 3):
                    want[i] = want.get(i, 0) - w
        assert sum(REF_CAND[k].values()) == 1  # so the a3 coefficient is 1
        want = {i: 12 * c for i, c in want.items() if c != 0}
        assert got == want, (k, got, want)
