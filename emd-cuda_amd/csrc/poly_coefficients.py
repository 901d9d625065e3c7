"""Derives the polynomial tables of the Hilbert kernel's own atan2 and sin/cos of pi x
(emd_hilbert.h: EMDH_ATAN_POLY, EMDH_TRIG_POLY).

Every table is a Chebyshev fit (mpmath.chebyfit at 60 digits, near-minimax) printed
highest degree first and rounded to double:

* atan: atan(t) = t + t^3 P(t^2) on t in [0, 1]; P is the degree-20 fit of
  (atan(sqrt(s)) - sqrt(s)) / s^(3/2) on s in [0, 1] (max error 5.0e-18).  With one correctly
  rounded division t = min(|x|,|y|) / max(|x|,|y|), Horner with fused multiply-adds and
  pi/2, pi as two-double sums, the atan2 is within 1.5 ulp of the exact value (20 million
  random arguments against long double atan2l).
* sin(pi r) = r S(r^2) and cos(pi r) = 1 + r^2 C(r^2) on r in [0, 1/4]: S and C are the
  degree-7 fits of sin(pi sqrt(u)) / sqrt(u) and (cos(pi sqrt(u)) - 1) / u on u in
  [0, 1/16] (max errors 5.6e-21 and 9.8e-22).

Run: python emd-cuda_amd/csrc/poly_coefficients.py
"""
import mpmath as mp


def fit(f, hi, terms):
    poly, err = mp.chebyfit(f, [0, hi], terms, error=True)
    return [float(c) for c in poly], err


def main():
    mp.mp.dps = 60

    def f_atan(s):
        if s == 0:
            return mp.mpf(-1) / 3
        t = mp.sqrt(s)
        return (mp.atan(t) - t) / (s * t)

    def f_sin(u):
        if u == 0:
            return mp.pi
        r = mp.sqrt(u)
        return mp.sin(mp.pi * r) / r

    def f_cos(u):
        if u == 0:
            return -mp.pi ** 2 / 2
        r = mp.sqrt(u)
        return (mp.cos(mp.pi * r) - 1) / u

    for name, f, hi, terms in (("atan P", f_atan, 1, 21), ("sin S", f_sin, mp.mpf(1) / 16, 8),
                               ("cos C", f_cos, mp.mpf(1) / 16, 8)):
        c, err = fit(f, hi, terms)
        print(f"// {name}: max error {mp.nstr(err, 3)}")
        print(",\n".join(f"  {v!r}" for v in c))
    for v, name in ((mp.pi / 2, "pi/2"), (mp.pi, "pi")):
        hi = float(v)
        print(f"// {name}: {hi!r} + {float(v - mp.mpf(hi))!r}")


if __name__ == "__main__":
    main()
