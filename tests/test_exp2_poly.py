"""The kernels' exp2 polynomials (templates/cg_math.j2), read out of the template and evaluated
in float64 Horner form on [-1/2, 1/2] against 2^f: the degree-10 interpolant within 6e-16
(3.1e-16 plus the rounding of the evaluation) and the degree-9 option ("d9") within 2e-14
relative, both exactly 1 at f = 0 (so 2^n is exact for integer arguments).  A wrong or dropped
coefficient fails these bounds by orders of magnitude."""
import os
import re

import numpy as np
import pytest

TPL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "chemgen_amd", "templates",
                   "cg_math.j2")


def _horner_block(label):
    s = open(TPL).read()
    i = s.index(label)
    j = s.index("{%", i + len(label))
    body = s[i:j]
    first = float(re.search(r"double p = ([0-9.eE+-]+);", body).group(1))
    rest = [float(x) for x in re.findall(r"p = __builtin_fma\(p, f, ([0-9.eE+-]+)\);", body)]
    return [first] + rest          # highest degree first


@pytest.mark.parametrize("label,degree,bound", [('{% elif poly == "d9" %}', 9, 2e-14),
                                                 ("{% else %}\n  double p = 7.07", 10, 6e-16)])
def test_exp2_polynomial_accuracy(label, degree, bound):
    c = _horner_block(label)
    assert len(c) == degree + 1 and c[-1] == 1.0
    f = np.linspace(-0.5, 0.5, 200001)
    p = np.full_like(f, c[0])
    for a in c[1:]:
        p = p * f + a
    rel = np.abs(p / np.exp2(f) - 1.0)
    assert rel.max() < bound, rel.max()
    assert p[100000] == 1.0           # f = 0
