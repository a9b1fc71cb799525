"""Writes rp_erfc_table.h: erfc on [0, pi] as 16 polynomials of degree 10.

The real-space sum needs erfc(sqrt(alpha) r) only for r <= Rc, and alpha = (pi / Rc)^2
(repeat.py:1087), so the argument lies in [0, pi] for every cutoff.  Each segment
[k pi/16, (k+1) pi/16] is fitted in t = x - centre by least squares on Chebyshev nodes
against 30-digit erfc (mpmath); the fit error is below 1e-15 absolute (1e-14 relative
at x = pi).  Run from this directory: python gen_erfc_table.py
"""
import mpmath
import numpy as np

NSEG, DEG = 16, 10


def fit():
    mpmath.mp.dps = 30
    w = np.pi / NSEG
    rows, worst = [], 0.0
    for k in range(NSEG):
        a, b = k * w, (k + 1) * w
        c = 0.5 * (a + b)
        n = 4 * (DEG + 1)
        xs = c + 0.5 * w * np.cos(np.pi * (np.arange(n) + 0.5) / n)
        ys = np.array([float(mpmath.erfc(mpmath.mpf(x))) for x in xs])
        coef, *_ = np.linalg.lstsq(np.vander(xs - c, DEG + 1, increasing=True), ys, rcond=None)
        xt = np.linspace(a, b, 257)
        yt = np.array([float(mpmath.erfc(mpmath.mpf(x))) for x in xt])
        worst = max(worst, float(np.abs(np.polynomial.polynomial.polyval(xt - c, coef) - yt).max()))
        rows.append((c, coef))
    return rows, worst


def main():
    rows, worst = fit()
    out = ["// Generated by gen_erfc_table.py: erfc(x) on [0, pi] as NSEG polynomials of degree DEG in",
           f"// t = x - centre_k (max abs error {worst:.1e}).",
           "#pragma once", "", "namespace rp {", "",
           f"constexpr int ERFC_NSEG = {NSEG}, ERFC_DEG = {DEG};",
           "// per segment: centre, then coefficients of t^0 .. t^DEG",
           f"__constant__ double kErfcTable[ERFC_NSEG][ERFC_DEG + 2] = {{"]
    for c, coef in rows:
        vals = [float(c)] + [float(v) for v in coef]
        out.append("    {" + ", ".join(f"{v!r}" for v in vals) + "},")
    out += ["};", "", "}  // namespace rp", ""]
    with open("rp_erfc_table.h", "w") as f:
        f.write("\n".join(out))
    print("max abs error", worst)


if __name__ == "__main__":
    main()
