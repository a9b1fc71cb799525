"""Writes rp_erf_poly.h: erf(x) = x P(x^2) on [0, pi], P one polynomial of degree DEG.

The real-space sum needs erfc(sqrt(alpha) r) / r only for r <= Rc, and alpha = (pi / Rc)^2
(repeat.py:1087), so u = alpha r^2 lies in [0, pi^2] for every cutoff, and

    erfc(sqrt(alpha) r) / r = 1/r - sqrt(alpha) P(alpha r^2),

which needs no square root, no division and no table lookup.  P interpolates erf(x)/x at
Chebyshev nodes in u (mpmath, 40 digits) and is stored as monomial coefficients in
s = u - U0, U0 = pi^2 / 2; the header states the measured error of the fp64 Horner evaluation.
Run from this directory: python gen_erf_poly.py
"""
import mpmath
import numpy as np

DEG = 21


def fit():
    mpmath.mp.dps = 40
    U = mpmath.pi ** 2
    u0 = U / 2
    n = DEG + 1
    nodes = [u0 + u0 * mpmath.cos(mpmath.pi * (k + mpmath.mpf(1) / 2) / n) for k in range(n)]
    vals = [mpmath.erf(mpmath.sqrt(u)) / mpmath.sqrt(u) for u in nodes]
    # interpolating polynomial in s = u - u0, solved at 40 digits
    A = mpmath.matrix([[(u - u0) ** d for d in range(n)] for u in nodes])
    coef = mpmath.lu_solve(A, mpmath.matrix(vals))
    c64 = [float(coef[d]) for d in range(n)]
    # error of erfc(x) = 1 - x P(x^2) with fp64 Horner, against 40-digit erfc
    worst = 0.0
    for x in np.linspace(0.0, float(mpmath.pi), 4001):
        s = x * x - float(u0)
        p = c64[-1]
        for d in range(DEG - 1, -1, -1):
            p = p * s + c64[d]
        worst = max(worst, abs(float(1 - x * p) - float(mpmath.erfc(x))))
    return float(u0), c64, worst


def main():
    u0, c64, worst = fit()
    out = ["// Generated by gen_erf_poly.py: erf(x) = x P(x^2) for 0 <= x <= pi, P of degree ERF_DEG in",
           f"// s = x^2 - ERF_U0 (fp64 Horner; max abs error of 1 - x P(x^2) against erfc: {worst:.1e}).",
           "#pragma once", "", "namespace rp {", "",
           f"constexpr int ERF_DEG = {DEG};",
           f"constexpr double ERF_U0 = {u0!r};",
           "// coefficients of s^0 .. s^ERF_DEG",
           "constexpr double kErfPoly[ERF_DEG + 1] = {"]
    for d in range(0, DEG + 1, 3):
        out.append("    " + " ".join(f"{v!r}," for v in c64[d:d + 3]))
    out += ["};", "", "}  // namespace rp", ""]
    with open("rp_erf_poly.h", "w") as f:
        f.write("\n".join(out))
    print("max abs error", worst)


if __name__ == "__main__":
    main()
