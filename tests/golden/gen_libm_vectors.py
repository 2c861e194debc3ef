"""Write tests/golden/libm_cr_vectors.json: correctly rounded log and exp.

Reference values come from Python's decimal module at 45 significant digits
(Decimal.ln / Decimal.exp are correctly rounded at that precision), rounded
once to the nearest double.  Inputs cover what the sweep feeds these
functions: Open01 values for the normal tail's log, and -x^2/2 for the
ziggurat wedge's exp, plus edge cases.
"""
import json
import pathlib
import random
from decimal import Decimal, getcontext

getcontext().prec = 45
OUT = pathlib.Path(__file__).with_name("libm_cr_vectors.json")


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
    OUT.write_text(json.dumps({"source": "decimal prec=45", "log": log, "exp": exp}))


if __name__ == "__main__":
    main()
