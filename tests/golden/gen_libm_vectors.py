"""Write tests/golden/libm_cr_vectors.json: correctly rounded log, exp, sin, cos.

Reference values come from Python's decimal module at 45 significant digits
(Decimal.ln / Decimal.exp are correctly rounded at that precision), rounded
once to the nearest double.  Inputs cover what the sweep feeds these
functions: Open01 values for the normal tail's log, and -x^2/2 for the
ziggurat wedge's exp, plus edge cases.  sin and cos (which decimal lacks)
are Taylor series summed at 60 digits, on [0, pi/2] where the rotation
sampler evaluates them (at half the rotation angle).
"""
import json
import pathlib
import random
from decimal import Decimal, getcontext

getcontext().prec = 45
OUT = pathlib.Path(__file__).with_name("libm_cr_vectors.json")


def _sin_cos(x):
    ctx = getcontext().copy()
    ctx.prec = 60
    d = Decimal(x)
    x2 = ctx.multiply(d, d)
    s_term, c_term = d, Decimal(1)
    s_sum, c_sum = d, Decimal(1)
    for n in range(1, 40):
        s_term = ctx.divide(ctx.multiply(-s_term, x2), Decimal((2 * n) * (2 * n + 1)))
        c_term = ctx.divide(ctx.multiply(-c_term, x2), Decimal((2 * n - 1) * (2 * n)))
        s_sum = ctx.add(s_sum, s_term)
        c_sum = ctx.add(c_sum, c_term)
    return float(s_sum), float(c_sum)


def main():
    rnd = random.Random(12345)
    xs = [(rnd.getrandbits(52) + 0.5) / 2**52 for _ in range(3000)]
    xs += [2.0**-53, 0.5, 0.75, 1.0 - 2.0**-53, 1.0, 2.0, 1e-300, 3.0, 0.7071067811865476, 1.4142135623730951]
    xs += [1.0 - k * 2.0**-40 for k in range(1, 50)] + [1.0 + k * 2.0**-40 for k in range(1, 50)]
    ys = [-(rnd.random() * 3.91) ** 2 / 2 for _ in range(3000)]
    ys += [rnd.uniform(-700.0, 700.0) for _ in range(500)]
    ys += [0.0, -1e-20, 1e-20, -0.5, -1.0, -700.0, 700.0, 0.34657359027997264, -0.34657359027997264]
    log = [[x.hex(), float(Decimal(x).ln()).hex()] for x in xs]
    exp = [[y.hex(), float(Decimal(y).exp()).hex()] for y in ys]
    zs = [rnd.random() * 1.5707963267948966 for _ in range(3000)]
    zs += [abs(rnd.gauss(0.0, 0.05)) for _ in range(1000)]
    zs += [0.0, 1e-300, 2.0**-30, 1e-8, 0.5, 1.0, 1.5707963267948966, 1.5707963267948963, 0.7853981633974483]
    sc = [_sin_cos(z) for z in zs]
    sin = [[z.hex(), v[0].hex()] for z, v in zip(zs, sc)]
    cos = [[z.hex(), v[1].hex()] for z, v in zip(zs, sc)]
    OUT.write_text(json.dumps({"source": "decimal prec=45 (log, exp), Taylor at 60 digits (sin, cos)",
                               "log": log, "exp": exp, "sin": sin, "cos": cos}))


if __name__ == "__main__":
    main()
