"""Extract the known-answer values of pi-BEM's expansion tests into a fixture.

Reads the reference's own test outputs (tests/*.output, deal.II deallog
format) and writes tests/golden/reference_goldens.json.  The charge layouts
of the corresponding tests/*.cc are recorded by hand as data next to the
values.  Run in the build container only (the reference is not present on the
GPU host); the JSON is what the tests load.
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests"
CASES = {
    # name: (output file, charges, intensities, normals or None, ops)
    "p2m_m2p": ("multipole_expansion_creation_and_evaluation.output",
                [[2.8, 3, 3], [3, 2.8, 3], [3, 3, 2.8], [3.2, 3, 3], [3, 3, 3.2]],
                [0.1, 0.4, 1.0, 0.2, 0.7], None, ["p2m", "m2p"]),
    "dipole_p2m_m2p": ("multipole_expansion_creation_and_evaluation_neumann_kernel.output",
                       [[2.8, 3, 3], [3, 2.8, 3], [3.2, 3, 2.8], [3.2, 3, 3], [2.7, 3, 3.2]],
                       [0.1, 0.4, 1.0, 0.2, 0.7], [[1, 0, 0]] * 5, ["p2m_dipole", "m2p"]),
    "m2m_m2p": ("multipole_expansion_translation.output",
                [[2.8, 3, 3], [3, 2.8, 3], [3, 3, 2.8], [3.2, 3, 3], [3, 3, 3.2]],
                [0.1, 0.4, 1.0, 0.2, 0.7], None, ["p2m", "m2m(4,4,4)", "m2p"]),
    "m2l_l2p": ("local_expansion_creation_and_evaluation.output",
                [[2.8, 3, 3], [3, 2.8, 3], [3, 3, 2.8], [3.2, 3, 3], [3, 3, 3.2]],
                [0.1, 0.4, 1.0, 0.2, 0.7], None, ["p2m", "m2l(-2,-2,-2)", "l2p"]),
    "m2l_l2l_l2p": ("local_expansion_translation.output",
                    [[2.8, 3, 3], [3, 2.8, 3], [3, 3, 2.8], [3.2, 3, 3], [3, 3, 3.2]],
                    [0.1, 0.4, 1.0, 0.2, 0.7], None, ["p2m", "m2l(-2,-2,-2)", "l2l(-1,-1,-1)", "l2p"]),
}


def parse(path):
    vals = {}
    for line in open(path):
        m = re.match(r"DEAL::(Exact|Approx)[^:]*:\s*(\S+)", line)
        if m:
            vals[m.group(1).lower()] = float(m.group(2))
        m = re.match(r"DEAL::The absolute .*: (\S+)", line)
        if m:
            vals["abs_err"] = float(m.group(1))
        m = re.match(r"DEAL::The relative .*: (\S+)", line)
        if m:
            vals["rel_err"] = float(m.group(1))
    return vals


out = {"source": "pi-BEM tests/*.output (deal.II deallog, 6 significant digits)",
       "order": 6, "center": [3, 3, 3], "eval_point": [0, 0, 0], "cases": {}}
for name, (f, ch, q, nrm, ops) in CASES.items():
    v = parse(os.path.join(REF, f))
    out["cases"][name] = dict(file=f, charges=ch, intensities=q, normals=nrm, ops=ops, **v)
json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_goldens.json"), "w"),
          indent=1)
print(json.dumps(out["cases"], indent=1))
