"""Writes tests/golden/su2_cg.json: exact Clebsch-Gordan coefficients
<j1 m1; j2 m2 | J M> (Condon-Shortley) from sympy.physics.wigner, as
two-j/two-m integer rows [tj1, tm1, tj2, tm2, tJ, tM, value].  The
reference ships no golden vectors for this path; these closed-form values pin
the SU(2) coefficient engines (oracle and product) instead."""
import json
import os
import random

from sympy import Rational as R
from sympy.physics.wigner import clebsch_gordan


def main():
    random.seed(2024)
    rows = []
    while len(rows) < 400:
        tj1 = random.randint(0, 12)
        tj2 = random.randint(0, 12)
        tJ = random.choice(range(abs(tj1 - tj2), tj1 + tj2 + 1, 2))
        tm1 = random.choice(range(-tj1, tj1 + 1, 2))
        tm2 = random.choice(range(-tj2, tj2 + 1, 2))
        tM = tm1 + tm2
        if abs(tM) > tJ:
            continue
        v = float(clebsch_gordan(R(tj1, 2), R(tj2, 2), R(tJ, 2), R(tm1, 2), R(tm2, 2), R(tM, 2)))
        rows.append([tj1, tm1, tj2, tm2, tJ, tM, v])
    # a few large-spin values with a spin-1/2 or spin-1 partner
    for tj1, tj2, tJ, tm1, tm2 in [(101, 1, 100, 1, -1), (101, 1, 102, -51, 1), (140, 2, 140, 0, 2),
                                   (140, 2, 142, 10, -2), (200, 2, 198, 100, 0)]:
        v = float(clebsch_gordan(R(tj1, 2), R(tj2, 2), R(tJ, 2), R(tm1, 2), R(tm2, 2), R(tm1 + tm2, 2)))
        rows.append([tj1, tm1, tj2, tm2, tJ, tm1 + tm2, v])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "su2_cg.json")
    with open(out, "w") as f:
        json.dump({"source": "sympy.physics.wigner.clebsch_gordan (exact, Condon-Shortley)",
                   "columns": ["tj1", "tm1", "tj2", "tm2", "tJ", "tM", "value"], "rows": rows}, f)
    print(f"wrote {len(rows)} rows to {out}")


if __name__ == "__main__":
    main()
