"""UCD (.inp) surface meshes for the BEM-FMM path.

pi-BEM reads its grids with deal.II ``GridIn::read_ucd(in, true)``
(source/computational_domain.cc:258-300), refines them, distributes Q1 DoFs and
finds the double nodes (source/bem_problem.cc:413-454).  This module does the
same for quadrilateral surface meshes without deal.II:

* ``read_ucd``   the AVS UCD text format: a header ``n_vertices n_cells 0 0 0``,
  ``n_vertices`` lines ``id x y z`` and ``n_cells`` lines ``id material type
  v1 ...``; ``quad`` records are surface cells, ``line`` records carry the
  manifold ids of lines (read with apply_all_indicators_to_manifolds, as the
  reference does) and no DoFs.
* ``refine``     uniform refinement of every quad into four (edge midpoints
  shared by the cells that share the edge, so double nodes stay double),
  optionally followed by a projection of the new points (the reference
  projects on its CAD manifolds instead).
* ``to_mesh``    Q1 DoFs on the vertices, in deal.II FE_Q order (UCD lists a
  quad's vertices counter-clockwise; deal.II's ``ucd_to_deal`` numbering is
  {0, 1, 3, 2}), double-node sets of coincident vertices, and per-DoF normals
  (the area-weighted mean of the adjacent cell normals of the DoF's own
  cells).  The cell's material id is kept in ``Mesh.cell_face`` and the line
  records' ids in ``Mesh.line_ids`` (used to find waterline lines).
"""
from __future__ import annotations

import numpy as np

from .mesh import Mesh, _double_nodes

UCD_TO_DEAL_QUAD = (0, 1, 3, 2)


def read_ucd(source, with_lines=False):
    """Vertices (V, 3), quads (C, 4) in UCD vertex order (0-based) and the
    material id of each quad; with_lines also returns the line records'
    vertex pairs (L, 2) and ids (L,).  ``source`` is a path or the file's text."""
    if "\n" not in source:
        with open(source) as fh:
            source = fh.read()
    lines = [ln.split() for ln in source.splitlines() if ln.strip() and not ln.lstrip().startswith("#")]
    nv, nc = int(lines[0][0]), int(lines[0][1])
    ids = {}
    verts = np.empty((nv, 3))
    for k in range(nv):
        t = lines[1 + k]
        ids[int(t[0])] = k
        verts[k] = [float(t[1]), float(t[2]), float(t[3])]
    quads, mats, segs, seg_ids = [], [], [], []
    for k in range(nc):
        t = lines[1 + nv + k]
        if t[2] == "quad":
            quads.append([ids[int(v)] for v in t[3:7]])
            mats.append(int(t[1]))
        elif t[2] == "line":
            segs.append([ids[int(v)] for v in t[3:5]])
            seg_ids.append(int(t[1]))
        else:
            raise ValueError(f"UCD cell type {t[2]!r} is not a surface quad")
    out = (verts, np.asarray(quads, dtype=np.int64), np.asarray(mats, dtype=np.int32))
    if with_lines:
        out += (np.asarray(segs, dtype=np.int64).reshape(-1, 2), np.asarray(seg_ids, dtype=np.int32))
    return out


def refine(verts, quads, mats, levels=1, project=None, segs=None, seg_ids=None):
    """Uniform refinement (each quad into four, children in UCD order), new
    points optionally mapped through ``project`` (an (n, 3) -> (n, 3)
    function).  Line records, when given, are split at the same midpoints
    and keep their ids; they are then returned after the quads' materials."""
    verts = np.asarray(verts, dtype=np.float64)
    for _ in range(levels):
        pts = [p for p in verts]
        edge_mid = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            m = edge_mid.get(key)
            if m is None:
                m = len(pts)
                pts.append(0.5 * (verts[a] + verts[b]))
                edge_mid[key] = m
            return m

        new_q, new_m = [], []
        for q, mat in zip(quads, mats):
            v0, v1, v2, v3 = (int(v) for v in q)
            e01, e12, e23, e30 = mid(v0, v1), mid(v1, v2), mid(v2, v3), mid(v3, v0)
            c = len(pts)
            pts.append(0.25 * (verts[v0] + verts[v1] + verts[v2] + verts[v3]))
            new_q += [[v0, e01, c, e30], [e01, v1, e12, c], [c, e12, v2, e23], [e30, c, e23, v3]]
            new_m += [mat] * 4
        if segs is not None:
            new_s, new_i = [], []
            for (a, b), i in zip(segs, seg_ids):
                m = mid(int(a), int(b))
                new_s += [[int(a), m], [m, int(b)]]
                new_i += [i, i]
            segs = np.asarray(new_s, dtype=np.int64).reshape(-1, 2)
            seg_ids = np.asarray(new_i, dtype=np.int32)
        n_old = verts.shape[0]
        verts = np.asarray(pts)
        if project is not None:
            verts[n_old:] = project(verts[n_old:])
        quads = np.asarray(new_q, dtype=np.int64)
        mats = np.asarray(new_m, dtype=np.int32)
    if segs is not None:
        return verts, quads, mats, segs, seg_ids
    return verts, quads, mats


def to_mesh(verts, quads, mats, name="ucd", segs=None, seg_ids=None):
    """Q1 mesh on the vertices that quads use (renumbered in order of first
    use), cell DoFs in deal.II FE_Q order, double nodes of coincident
    vertices, area-weighted vertex normals, and the ids of the line records
    on each cell's lines."""
    used = np.unique(quads.ravel())
    remap = -np.ones(verts.shape[0], dtype=np.int64)
    remap[used] = np.arange(used.size)
    xyz = verts[used]
    cd = remap[quads][:, list(UCD_TO_DEAL_QUAD)]
    # cell normal x_xi cross x_eta at the centre, deal.II vertex order
    x = xyz[cd]
    xi = 0.5 * ((x[:, 1] - x[:, 0]) + (x[:, 3] - x[:, 2]))
    eta = 0.5 * ((x[:, 2] - x[:, 0]) + (x[:, 3] - x[:, 1]))
    cn = np.cross(xi, eta)
    nrm = np.zeros_like(xyz)
    for j in range(4):
        np.add.at(nrm, cd[:, j], cn)
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    m = Mesh(xyz, cd, 4, nrm, np.asarray(mats, dtype=np.int32), name=name)
    m.dbl_ptr, m.dbl_idx = _double_nodes(xyz)
    if segs is not None and len(segs):
        from .mesh import CELL_LINES
        tag = {}
        for (a, b), i in zip(remap[np.asarray(segs)], seg_ids):
            if a >= 0 and b >= 0:
                tag[(min(a, b), max(a, b))] = int(i)
        ids = -np.ones((cd.shape[0], 4), dtype=np.int32)
        for c in range(cd.shape[0]):
            for j, (u, v) in enumerate(CELL_LINES):
                a, b = int(cd[c, u]), int(cd[c, v])
                ids[c, j] = tag.get((min(a, b), max(a, b)), -1)
        m.line_ids = ids
    return m


def load(source, levels=0, project=None, name=None):
    """read_ucd + refine + to_mesh, line records included."""
    v, q, mt, sg, si = read_ucd(source, with_lines=True)
    if levels:
        v, q, mt, sg, si = refine(v, q, mt, levels, project, sg, si)
    return to_mesh(v, q, mt, name or "ucd", sg, si)


def sphere_projection(radius=1.0):
    """Projection of new points onto the sphere of the given radius about the
    origin (refining the reference's coarse sphere grids)."""
    def project(p):
        return radius * p / np.linalg.norm(p, axis=1)[:, None]
    return project


def problem(mesh, dirichlet_materials=(), p=6, seed=11, dirichlet_by_z=False):
    """A Problem on an ingested mesh: Dirichlet on the cells whose material id
    is listed (or, with dirichlet_by_z, on cells with centroid z above the
    mesh's mid height), and a seeded 8-source harmonic field whose sources sit
    at 2.5x the mesh's half extent around its centre."""
    from .mesh import PointSources, Problem, bc_vector, cell_centroids, dirichlet_mask
    c = cell_centroids(mesh)
    if dirichlet_by_z:
        zmid = 0.5 * (mesh.support_xyz[:, 2].min() + mesh.support_xyz[:, 2].max())
        cell_dir = c[:, 2] > zmid
    else:
        cell_dir = np.isin(mesh.cell_face, np.asarray(dirichlet_materials, dtype=np.int32))
    dmask = dirichlet_mask(mesh, cell_dir)
    centre = 0.5 * (mesh.support_xyz.min(axis=0) + mesh.support_xyz.max(axis=0))
    half = 0.5 * (mesh.support_xyz.max(axis=0) - mesh.support_xyz.min(axis=0)).max()
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(8, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    sol = PointSources(centre + 2.5 * half * v, rng.normal(size=8))
    return Problem(mesh, dmask, sol, bc_vector(mesh, dmask, sol), p)
