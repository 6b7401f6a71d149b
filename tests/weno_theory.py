"""WENO reconstruction derived from its definition in exact rational arithmetic.

Jiang & Shu (1996), Balsara & Shu (2000): on unit cells centred at integer
positions, candidate stencil k of r cells (k = 0..r-1, cells k-r+1 .. k) holds
the polynomial p_k of degree r-1 whose cell averages are the stencil values;
the interface value is p_k(1/2).  The optimal weights d_k make the combination
of candidates equal the (2r-1)-cell reconstruction.  The smoothness indicator
is beta_k = sum_{l=1}^{r-1} int_{-1/2}^{1/2} (d^l p_k / dx^l)^2 dx, a quadratic
form in the stencil values.  Nothing here is taken from an implementation: all
coefficients come out of the definitions, so a wrong coefficient in the oracle
fails the comparison.
"""
from __future__ import annotations

from fractions import Fraction as Fr
from functools import lru_cache

import numpy as np


def _solve(A, b):
    n = len(A)
    M = [list(row) + [bb] for row, bb in zip(A, b)]
    for c in range(n):
        p = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[p] = M[p], M[c]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [x - f * y for x, y in zip(M[r], M[c])]
    return [M[i][n] / M[i][i] for i in range(n)]


def _recon_poly(cells, m):
    """Ascending coefficients of the polynomial whose averages over the unit
    cells centred at ``cells`` are the unit vector e_m (via the primitive)."""
    r = len(cells)
    xs = [Fr(2 * cells[0] - 1, 2) + i for i in range(r + 1)]
    prim = [Fr(0)]
    for i in range(r):
        prim.append(prim[-1] + (1 if i == m else 0))
    a = _solve([[x ** e for e in range(r + 1)] for x in xs], prim)
    return [a[e] * e for e in range(1, r + 1)]


def _ev(c, x):
    return sum(ci * x ** i for i, ci in enumerate(c))


def _der(c):
    return [c[i] * i for i in range(1, len(c))]


def _int(c, lo, hi):
    return sum(ci * (hi ** (i + 1) - lo ** (i + 1)) / (i + 1) for i, ci in enumerate(c))


def _mul(a, b):
    out = [Fr(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            out[i + j] += x * y
    return out


@lru_cache(maxsize=None)
def weno_coefficients(r: int):
    """(c, d, Q): c[k][m] candidate interface coefficients over the candidate's
    own r cells, d[k] optimal weights, Q[k] the smoothness quadratic form."""
    half = Fr(1, 2)
    c, Q = [], []
    for k in range(r):
        cells = list(range(k - (r - 1), k + 1))
        polys = [_recon_poly(cells, m) for m in range(r)]
        c.append([_ev(p, half) for p in polys])
        q = [[Fr(0)] * r for _ in range(r)]
        for l in range(1, r):
            ds = []
            for p in polys:
                for _ in range(l):
                    p = _der(p)
                ds.append(p)
            for a in range(r):
                for b in range(r):
                    if ds[a] and ds[b]:
                        q[a][b] += _int(_mul(ds[a], ds[b]), -half, half)
        Q.append(q)
    big_cells = list(range(-(r - 1), r))
    big = [_ev(_recon_poly(big_cells, m), half) for m in range(2 * r - 1)]
    # sum_k d_k c_k (placed at big-stencil offset k) = big: lower triangular in the first r rows
    d = []
    for row in range(r):
        acc = sum(d[k] * c[k][row - k] for k in range(row))
        d.append((big[row] - acc) / c[row][0])
    for row in range(2 * r - 1):
        tot = sum(d[k] * c[k][row - k] for k in range(r) if 0 <= row - k < r)
        assert tot == big[row], "optimal weights inconsistent"
    return c, d, Q


#: Normalisation of the smoothness indicators: Jiang-Shu's form for r = 3 and
#: the integer-coefficient forms of Balsara & Shu (2000) for r = 4, 5.
BETA_SCALE = {3: 1, 4: 240, 5: 5040}
EPS = 1e-10


def weno_from_definition(q, order: int) -> float:
    """Interface value at i+1/2 from q[0..2r-2] = phi_{i-r+1} .. phi_{i+r-1},
    evaluated in double precision from the derived coefficients, with the
    weights d_k / (eps + beta_k)^2 normalised."""
    r = (order + 1) // 2
    c, d, Q = weno_coefficients(r)
    q = np.asarray(q, dtype=np.float64)
    s = BETA_SCALE[r]
    num = 0.0
    den = 0.0
    for k in range(r):
        w = q[k:k + r]
        p = float(sum(float(c[k][m]) * w[m] for m in range(r)))
        Qf = np.array([[float(s * Q[k][a][b]) for b in range(r)] for a in range(r)])
        beta = float(w @ Qf @ w)
        alpha = float(d[k]) / ((beta + EPS) * (beta + EPS))
        num += alpha * p
        den += alpha
    return num / den


def linear_reconstruction(q, order: int) -> float:
    """The (2r-1)-cell linear reconstruction (the WENO limit on smooth data)."""
    r = (order + 1) // 2
    big_cells = list(range(-(r - 1), r))
    coef = [float(_ev(_recon_poly(big_cells, m), Fr(1, 2))) for m in range(2 * r - 1)]
    return float(np.dot(coef, np.asarray(q, dtype=np.float64)))
