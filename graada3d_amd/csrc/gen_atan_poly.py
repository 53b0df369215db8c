"""Writes atan_poly.hpp: coefficients of P with atan(t) = t P(t^2) on
0 <= t <= 1, the degree-20 polynomial interpolating atan(sqrt(s))/sqrt(s) at
the 21 Chebyshev points of s in [0, 1], computed in 70-digit decimal
arithmetic (Vandermonde solve) and rounded to double.  Checked: within 1 ulp
of atan on [0, 1] when evaluated in double by Horner's rule.

    python graada3d_amd/csrc/gen_atan_poly.py
"""
import math
import os
from decimal import Decimal as D, getcontext

getcontext().prec = 70
DEG = 20


def datan(x):
    """atan by argument halving (atan x = 2 atan(x / (1 + sqrt(1 + x^2))))
    and its Taylor series."""
    x, k = D(x), 0
    while abs(x) > D("0.05"):
        x, k = x / (1 + (1 + x * x).sqrt()), k + 1
    s, term, n, x2 = D(0), x, 0, x * x
    while True:
        t = term / (2 * n + 1)
        if abs(t) < D(10) ** -65:
            break
        s += t if n % 2 == 0 else -t
        term, n = term * x2, n + 1
    return s * (2 ** k)


def f(s):
    s = D(s)
    if s == 0:
        return D(1)
    r = s.sqrt()
    return datan(r) / r


def fit(n):
    nodes = [D("0.5") + D("0.5") * D(math.cos(math.pi * (k + 0.5) / (n + 1))) for k in range(n + 1)]
    N = n + 1
    A = [[x ** j for j in range(N)] for x in nodes]
    b = [f(x) for x in nodes]
    for i in range(N):
        p = max(range(i, N), key=lambda r: abs(A[r][i]))
        A[i], A[p], b[i], b[p] = A[p], A[i], b[p], b[i]
        for r in range(i + 1, N):
            m = A[r][i] / A[i][i]
            for c in range(i, N):
                A[r][c] -= m * A[i][c]
            b[r] -= m * b[i]
    c = [D(0)] * N
    for i in reversed(range(N)):
        c[i] = (b[i] - sum(A[i][k] * c[k] for k in range(i + 1, N))) / A[i][i]
    return [float(v) for v in c]


def main():
    c = fit(DEG)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "atan_poly.hpp")
    with open(out, "w") as fh:
        fh.write("// Generated by gen_atan_poly.py: atan(t) = t * P(t^2) on [0, 1],\n")
        fh.write(f"// P(s) = sum_k kAtanP[k] s^k, degree {DEG}.\n#pragma once\n\nnamespace gra {{\n\n")
        fh.write(f"constexpr int kAtanDeg = {DEG};\n")
        fh.write("constexpr double kAtanP[kAtanDeg + 1] = {\n")
        fh.write(",\n".join(f"    {v!r}" for v in c) + "};\n\n}  // namespace gra\n")
    print(out)


if __name__ == "__main__":
    main()
