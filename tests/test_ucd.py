"""UCD ingest (pi-bem_amd/pibem_amd/ucd.py) on the grids pi-BEM ships
(tests/golden/meshes/, copied by tests/golden/make_reference_meshes.py):
parsing, deal.II vertex order and outward orientation, double nodes of
coincident vertices, refinement that keeps closed bodies closed and double
nodes double, and the product octree against the oracle's on refined grids.
CPU only."""
import os

import numpy as np
import pytest

from pibem_amd import ucd

HERE = os.path.dirname(os.path.abspath(__file__))
MESHES = os.path.join(HERE, "golden", "meshes")
CLOSED = ["coarse_cube", "coarse_cube_double_nodes", "coarse_sphere", "coarse_sphere_double_nodes"]


def _area_vectors(m):
    x = m.support_xyz[m.cell_dofs]
    xi = 0.5 * ((x[:, 1] - x[:, 0]) + (x[:, 3] - x[:, 2]))
    eta = 0.5 * ((x[:, 2] - x[:, 0]) + (x[:, 3] - x[:, 1]))
    return x.mean(axis=1), np.cross(xi, eta)


@pytest.mark.parametrize("name", CLOSED)
def test_closed_grids(name):
    m = ucd.load(os.path.join(MESHES, name + ".inp"))
    c, a = _area_vectors(m)
    # closed surface: the area vectors sum to zero; every cell faces outwards
    assert np.abs(a.sum(axis=0)).max() < 1e-12
    ctr = m.support_xyz.mean(axis=0)
    assert np.all(((c - ctr) * a).sum(axis=1) > 0)
    # refining keeps it closed and multiplies the cells by 4 per level
    m2 = ucd.load(os.path.join(MESHES, name + ".inp"), levels=2)
    assert m2.n_cells == 16 * m.n_cells
    assert np.abs(_area_vectors(m2)[1].sum(axis=0)).max() < 1e-12


def test_double_nodes():
    m = ucd.load(os.path.join(MESHES, "coarse_cube_double_nodes.inp"))
    # six separate faces: every corner DoF has two doubles
    assert m.n_dofs == 24 and np.all(np.diff(m.dbl_ptr) == 3)
    m2 = ucd.load(os.path.join(MESHES, "coarse_cube_double_nodes.inp"), levels=2)
    k = 4  # cells per face edge after two refinements
    sizes = np.bincount(np.diff(m2.dbl_ptr))
    assert m2.n_dofs == 6 * (k + 1) ** 2
    assert sizes[3] == 24 and sizes[2] == 12 * (k - 1) * 2 and sizes[1] == 6 * (k - 1) ** 2
    # a face shared by two grids with coincident vertices: pairs
    s = ucd.load(os.path.join(MESHES, "coarse_sphere_double_nodes.inp"))
    assert np.array_equal(np.bincount(np.diff(s.dbl_ptr)), [0, 2, 12])


def test_sphere_projection():
    m = ucd.load(os.path.join(MESHES, "coarse_sphere.inp"), levels=3, project=ucd.sphere_projection())
    assert m.n_cells == 6 * 64
    assert np.allclose(np.linalg.norm(m.support_xyz, axis=1), 1.0, atol=1e-9)
    # vertex normals of a sphere point along the position
    assert np.min((m.normals * m.support_xyz).sum(axis=1)) > 0.99


def test_open_grids_parse():
    for name, nq in (("hemisphere", 10), ("dtc_ship", 28)):
        v, q, mt = ucd.read_ucd(os.path.join(MESHES, name + ".inp"))
        assert q.shape == (nq, 4) and set(mt.tolist()) == {1, 2}


@pytest.mark.parametrize("name,levels", [("coarse_cube_double_nodes", 3), ("coarse_sphere", 3), ("dtc_ship", 2)])
def test_tree_matches_oracle_on_ingested(name, levels, oracle_mod):
    import pibem_amd as P
    proj = ucd.sphere_projection() if name == "coarse_sphere" else None
    m = ucd.load(os.path.join(MESHES, name + ".inp"), levels=levels, project=proj)
    pr = ucd.problem(m, dirichlet_by_z=True, p=4)
    o = oracle_mod.OracleFMM(pr)
    o.generate_octree_blocking()
    to = o.tree()
    f = P.BEMFMA(trunc_order=pr.p, num_octree_levels=pr.levels, max_nodes_per_leaf=pr.max_nodes_per_leaf)
    f.init_fma(pr.mesh, pr.dirichlet)
    f.generate_octree_blocking()
    tp = f.tree()
    for k in to:
        if isinstance(to[k], np.ndarray):
            assert np.array_equal(to[k], tp[k]), k


def test_line_records_and_refinement():
    """Line records keep their manifold ids on the cell lines they lie on,
    refinement splits each recorded line in two at the shared edge midpoint
    (both halves keep the id), and the boundary lines of the DTC ship are the
    open edges of its patches."""
    from pibem_amd import mesh as M
    path = os.path.join(MESHES, "dtc_ship.inp")
    v, q, mt, sg, si = ucd.read_ucd(path, with_lines=True)
    assert len(sg) == len(si) > 0
    m0 = ucd.load(path)
    # every recorded line that is an edge of some quad is found on a cell line
    ids0, counts0 = np.unique(m0.line_ids[m0.line_ids >= 0], return_counts=True)
    assert set(ids0.tolist()) <= set(si.tolist())
    v1, q1, mt1, sg1, si1 = ucd.refine(v, q, mt, 1, None, sg, si)
    assert len(sg1) == 2 * len(sg) and np.array_equal(np.repeat(si, 2), si1)
    # the two halves meet at the old line's midpoint
    mids = v1[sg1[0::2, 1]]
    assert np.allclose(mids, 0.5 * (v[sg[:, 0]] + v[sg[:, 1]]))
    m1 = ucd.load(path, 1)
    ids1, counts1 = np.unique(m1.line_ids[m1.line_ids >= 0], return_counts=True)
    assert np.array_equal(ids0, ids1) and np.array_equal(2 * counts0, counts1)
    # lines at the boundary of the surface: each belongs to one cell only
    bl = M.boundary_lines(m1)
    cd = m1.cell_dofs
    for c, j in zip(*np.nonzero(bl)):
        a, b = cd[c, M.CELL_LINES[j][0]], cd[c, M.CELL_LINES[j][1]]
        share = sum(1 for cc in range(len(cd)) for jj, (u, w) in enumerate(M.CELL_LINES)
                    if {cd[cc, u], cd[cc, w]} == {a, b})
        assert share == 1


def test_mark_lines_waterline():
    """mark_lines gives an id to the boundary lines whose end points both lie
    on a predicate's set (the z = 0 waterline of the half Wigley hull)."""
    from pibem_amd import mesh as M
    pr = M.config_wigley_half(12, 4, 4, 1)
    m = M.mark_lines(pr.mesh, lambda x: np.abs(x[:, 2]) < 1e-12, 5)
    c, j = np.nonzero(m.line_ids == 5)
    assert c.size == 2 * 12  # one waterline line per column of cells, two sides
    for cc, jj in zip(c, j):
        ends = m.support_xyz[m.cell_dofs[cc, list(M.CELL_LINES[jj])]]
        assert np.all(np.abs(ends[:, 2]) < 1e-12)
