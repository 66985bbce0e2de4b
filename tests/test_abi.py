"""The C ABI library loads and exports every entry point that include/pbfmm.h
declares; argument and call-order errors come back as status codes with a
message (no GPU needed for these paths)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "pbfmm.h")).read()
    return sorted(set(re.findall(r"\b(pbfmm_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    from pibem_amd import _lib
    L = _lib.load()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES)


def test_invalid_parameters_rejected():
    from pibem_amd import _lib
    L = _lib.load()
    prm = _lib.Params()
    L.pbfmm_default_params(C.byref(prm))
    assert (prm.num_octree_levels, prm.max_nodes_per_leaf, prm.trunc_order) == (10, 20, 6)
    prm.trunc_order = 11  # the reference's static tables stop at p = 10
    h = C.c_void_p()
    assert L.pbfmm_create(C.byref(prm), C.byref(h)) == -1


def test_call_order_and_arguments():
    import pibem_amd as P
    from pibem_amd import mesh as M
    f = P.BEMFMA(trunc_order=4)
    with pytest.raises(P.PbfmmError) as e:
        f.generate_octree_blocking()
    assert e.value.code == -2
    pr = M.config_sphere(4, 2)
    bad = pr.mesh.cell_dofs.copy()
    bad[0, 0] = pr.mesh.n_dofs + 5
    m2 = M.Mesh(pr.mesh.support_xyz, bad, 4, pr.mesh.normals, pr.mesh.cell_face, pr.mesh.dbl_ptr, pr.mesh.dbl_idx)
    with pytest.raises(P.PbfmmError) as e:
        f.init_fma(m2, pr.dirichlet)
    assert e.value.code == -1 and "out of range" in str(e.value)
