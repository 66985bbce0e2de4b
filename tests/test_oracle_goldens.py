"""The oracle's expansion algebra against pi-BEM's own known-answer tests
(tests/multipole_expansion_*.output, tests/local_expansion_*.output, order 6,
five charges around (3,3,3), evaluation at the origin).  The printed absolute
error pins the approximation to about 6 significant digits of the error, far
tighter than the printed potential."""
import json
import os

import numpy as np
import pytest

G = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.json")))


def _rel6(a, b):
    return abs(a - b) <= 5.1e-6 * abs(b)


@pytest.mark.parametrize("name", list(G["cases"]))
def test_reference_golden(name, oracle_mod):
    O = oracle_mod
    cs = G["cases"][name]
    p = G["order"]
    c = np.array(G["center"], float)
    x0 = np.array(G["eval_point"], float)
    y = np.array(cs["charges"], float)
    q = np.array(cs["intensities"], float)
    if cs["normals"]:
        nr = np.array(cs["normals"], float)
        M = O.p2m_dipole(p, c, q, y, nr)
        exact = sum(-qi * (yi @ ni) / np.linalg.norm(yi) ** 3 for qi, yi, ni in zip(q, y, nr))
    else:
        M = O.p2m(p, c, q, y)
        exact = sum(qi / np.linalg.norm(yi) for qi, yi in zip(q, y))
    if name in ("p2m_m2p", "dipole_p2m_m2p"):
        v = O.me_eval(p, c, M, x0)
    elif name == "m2m_m2p":
        v = O.me_eval(p, [4, 4, 4], O.m2m(p, c, M, np.array([4.0, 4, 4])), x0)
    elif name == "m2l_l2p":
        v = O.le_eval(p, [-2, -2, -2], O.m2l(p, c, M, np.array([-2.0, -2, -2])), x0)
    else:
        Lc = O.m2l(p, c, M, np.array([-2.0, -2, -2]))
        v = O.le_eval(p, [-1, -1, -1], O.l2l(p, [-2, -2, -2], Lc, np.array([-1.0, -1, -1])), x0)
    assert _rel6(exact, cs["exact"])
    assert _rel6(v, cs["approx"])
    assert _rel6(abs(exact - v), cs["abs_err"])
