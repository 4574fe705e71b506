"""Write tests/golden/sincos_cr.npz: correctly rounded sin and cos of FP64
arguments, the checker of the device prologue's sin / cos.

Each value is computed in exact decimal arithmetic: the argument (a binary64
number, hence an exact decimal) is reduced by pi/2 with pi to 110 digits, the
Taylor series of sin and cos of the reduced argument is summed at 60
significant digits, and the sum is rounded once to the nearest binary64
(Python's Decimal -> float conversion is correctly rounded).

Arguments: the cell and vertex longitudes / latitudes of a small synthetic
ICON domain (the values the prologue sees), uniform samples of [-pi, pi],
doubles next to multiples of pi/2, and edge cases (signed zeros, tiny and
larger magnitudes).  Run:  python tests/golden/make_sincos_fixture.py
"""

import decimal
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

PI = decimal.Decimal(
    "3.14159265358979323846264338327950288419716939937510582097494459230781640628620899862803"
    "482534211706798214808651")


def cr_sin_cos(x: float):
    """(sin x, cos x) rounded to nearest binary64, from exact decimals."""
    if x == 0.0:
        return x, 1.0
    with decimal.localcontext() as ctx:
        ctx.prec = 130
        X = decimal.Decimal(x)  # exact
        half = PI / 2
        k = int((X / half).to_integral_value(rounding=decimal.ROUND_HALF_EVEN))
        r = X - k * half
        ctx.prec = 60
        r2 = r * r
        eps = decimal.Decimal(10) ** -80
        s, term, n = r, r, 1
        while abs(term) > eps:
            term = -term * r2 / ((2 * n) * (2 * n + 1))
            s += term
            n += 1
        c, term, n = decimal.Decimal(1), decimal.Decimal(1), 1
        while abs(term) > eps:
            term = -term * r2 / ((2 * n - 1) * (2 * n))
            c += term
            n += 1
        q = k % 4
        sv, cv = [(s, c), (c, -s), (-s, -c), (-c, s)][q]
        return float(sv), float(cv)


def arguments() -> np.ndarray:
    from horayzon_extpar_amd import synthetic

    d = synthetic.lattice_domain(30.0, 30.0, 1.0)
    rng = np.random.default_rng(20240601)
    xs = [d.clon[::7], d.clat[::7], d.vlon[::5], d.vlat[::5],
          rng.uniform(-np.pi, np.pi, 1500)]
    near = []
    for k in range(-8, 9):
        c = k * np.pi / 2
        for m in range(-3, 4):
            near.append(np.nextafter(c, c + m) if m else c)
            near.append(c + m * 1e-9)
    edge = [0.0, -0.0, 1e-300, -1e-300, 2.0 ** -30, -(2.0 ** -30), 1e-8, 0.5, 1.0, 2.0, 3.0,
            100.0, -1000.5, 12345.678, 9.9e4, np.pi / 6, np.pi / 4, np.pi / 3]
    x = np.concatenate([np.concatenate(xs), np.asarray(near), np.asarray(edge)])
    return x.astype(np.float64)


def main():
    x = arguments()
    sc = np.array([cr_sin_cos(float(v)) for v in x], dtype=np.float64)
    out = os.path.join(HERE, "sincos_cr.npz")
    np.savez(out, x=x, sin=sc[:, 0], cos=sc[:, 1])
    print(f"wrote {out}: {x.size} arguments")


if __name__ == "__main__":
    main()
