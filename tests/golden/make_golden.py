"""Derive the exact coefficient sets the oracle is pinned against (rational arithmetic).

Nothing here comes from the reference: every vector is the unique solution of a textbook
accuracy condition, computed with fractions and written to linear_schemes.json:

  central_d1[w]   first derivative on offsets -w..w, exact for polynomials of degree <= 2w
  central_d2[w]   second derivative on offsets -w..w, exact for degree <= 2w+1
  upwind[r]       finite-difference WENO linear flux f^_{i+1/2} on offsets -r..r (upwind-biased),
                  exact for fluxes that are cell averages of degree-2r polynomials
  is_forms[r]     Jiang-Shu smoothness indicator of each r-point sub-stencil, as a symmetric
                  matrix A_k with IS_k = f^T A_k f, IS_k = sum_l int_{-1/2}^{1/2} (p_k^(l))^2 dx

Run: python tests/golden/make_golden.py   (writes tests/golden/linear_schemes.json)
"""
import json
import os
from fractions import Fraction as Fr

HERE = os.path.dirname(os.path.abspath(__file__))


def solve(A, b):
    """Exact Gaussian elimination over the rationals."""
    n = len(A)
    M = [row[:] + [bb] for row, bb in zip(A, b)]
    for c in range(n):
        p = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[p] = M[p], M[c]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [x - f * y for x, y in zip(M[r], M[c])]
    return [M[r][n] / M[r][r] for r in range(n)]


def central(w, deriv):
    offs = list(range(-w, w + 1))
    A = [[Fr(o) ** p for o in offs] for p in range(len(offs))]
    b = [Fr(0)] * len(offs)
    b[deriv] = Fr(1) if deriv == 1 else Fr(2)
    return solve(A, b)


def prim_poly_avg(j, p):
    """Average over cell [j-1/2, j+1/2] of x^(p-1) * p, i.e. H(j+1/2) - H(j-1/2) for H = x^p."""
    return Fr(2 * j + 1, 2) ** p - Fr(2 * j - 1, 2) ** p


def upwind(r):
    """Coefficients c_m, m = -r..r, with sum_m c_m h_bar(m) = h(1/2) for polynomial h of degree 2r."""
    offs = list(range(-r, r + 1))
    A, b = [], []
    for p in range(1, 2 * r + 2):      # H = x^p  ->  h = p x^(p-1), degrees 0..2r
        A.append([prim_poly_avg(m, p) for m in offs])
        b.append(p * Fr(1, 2) ** (p - 1))
    return solve(A, b)


def poly_from_averages(cells):
    """Coefficients a_q of p(x) = sum a_q x^q whose averages over the given unit cells are e_i."""
    n = len(cells)
    A = [[(Fr(2 * c + 1, 2) ** (q + 1) - Fr(2 * c - 1, 2) ** (q + 1)) / (q + 1) for q in range(n)] for c in cells]
    cols = []
    for i in range(n):
        e = [Fr(0)] * n
        e[i] = Fr(1)
        cols.append(solve(A, e))          # polynomial for unit data on cell i
    return cols


def deriv(a, l):
    out = a[:]
    for _ in range(l):
        out = [q * out[q] for q in range(1, len(out))]
    return out


def int_sq(a):
    """int_{-1/2}^{1/2} (sum a_q x^q)^2 dx."""
    tot = Fr(0)
    for i, ai in enumerate(a):
        for j, aj in enumerate(a):
            e = i + j + 1
            tot += ai * aj * (Fr(1, 2) ** e - Fr(-1, 2) ** e) / e
    return tot


def is_forms(r):
    """Smoothness-indicator matrices of the r+1 sub-stencils of a (2r+1)-point WENO stencil,
    indices relative to the stencil start (values f_0 .. f_2r)."""
    forms = []
    for k in range(r + 1):
        cells = [c - r for c in range(k, k + r + 1)]       # sub-stencil k, target cell 0
        basis = poly_from_averages(cells)
        n = len(cells)
        Amat = [[Fr(0)] * (2 * r + 1) for _ in range(2 * r + 1)]
        for i in range(n):
            for j in range(n):
                v = Fr(0)
                for l in range(1, n):
                    di, dj = deriv(basis[i], l), deriv(basis[j], l)
                    m = max(len(di), len(dj))
                    di = di + [Fr(0)] * (m - len(di))
                    dj = dj + [Fr(0)] * (m - len(dj))
                    # int (p_i^(l) p_j^(l)) = polarisation of int_sq
                    v += (int_sq([x + y for x, y in zip(di, dj)]) - int_sq(di) - int_sq(dj)) / 2
                Amat[k + i][k + j] = v
        forms.append(Amat)
    return forms


def fs(v):
    return [str(x) for x in v]


def main():
    out = {
        "central_d1": {str(w): fs(central(w, 1)) for w in (3, 5)},
        "central_d2": {str(w): fs(central(w, 2)) for w in (3, 5)},
        "upwind": {str(r): fs(upwind(r)) for r in (2, 4)},
        "is_forms": {str(r): [[fs(row) for row in A] for A in is_forms(r)] for r in (2, 4)},
    }
    path = os.path.join(HERE, "linear_schemes.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
