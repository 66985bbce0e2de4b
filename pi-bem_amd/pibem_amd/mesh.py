"""Synthetic surface meshes and boundary data for the BEM-FMM path.

pi-BEM reads its meshes through deal.II ``GridIn`` and builds DoF numbering,
double-node sets and Dirichlet masks in ``BEMProblem::reinit``
(source/bem_problem.cc:91-211, 314-454).  This module produces the same
quantities for closed synthetic bodies so that the C-ABI can be driven without
deal.II:

* ``support_xyz``  (N, 3)   DoF support points
* ``cell_dofs``    (C, dpc) DoFs of each cell in deal.II FE_Q numbering
                            (vertices (0,0),(1,0),(0,1),(1,1); then the lines
                            x=0, x=1, y=0, y=1; then the interior), oriented so
                            that x_xi cross x_eta points out of the body
* ``dbl_ptr/idx``  CSR of the double-node sets, each set holding the DoF itself
                   (compute_double_nodes_set, bem_problem.cc:413-454, tol 1e-10)
* ``dirichlet``    uint8 mask: a DoF is Dirichlet when any cell holding it is
                   Dirichlet (compute_dirichlet_and_neumann_dofs_vectors,
                   bem_problem.cc:314-411)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# outward normal, tangent u, tangent v with u x v = normal
_FACES = [
    ((1, 0, 0), (0, 1, 0), (0, 0, 1)),
    ((-1, 0, 0), (0, 0, 1), (0, 1, 0)),
    ((0, 1, 0), (0, 0, 1), (1, 0, 0)),
    ((0, -1, 0), (1, 0, 0), (0, 0, 1)),
    ((0, 0, 1), (1, 0, 0), (0, 1, 0)),
    ((0, 0, -1), (0, 1, 0), (1, 0, 0)),
]
# FE_Q local node -> (i, j) offsets on a (deg+1)^2 lattice patch
_Q1 = [(0, 0), (1, 0), (0, 1), (1, 1)]
_Q2 = [(0, 0), (2, 0), (0, 2), (2, 2), (0, 1), (2, 1), (1, 0), (1, 2), (1, 1)]


@dataclass
class Mesh:
    support_xyz: np.ndarray
    cell_dofs: np.ndarray
    dpc: int
    normals: np.ndarray  # exact outward unit normal at each DoF (of its own face)
    cell_face: np.ndarray  # face/patch id of each cell
    dbl_ptr: np.ndarray = field(default=None)
    dbl_idx: np.ndarray = field(default=None)
    name: str = ""
    # manifold id of each cell's four lines (deal.II line numbering: x = 0,
    # x = 1, y = 0, y = 1 of the unit cell), -1 where none is set
    line_ids: np.ndarray = field(default=None)

    @property
    def n_dofs(self) -> int:
        return int(self.support_xyz.shape[0])

    @property
    def n_cells(self) -> int:
        return int(self.cell_dofs.shape[0])

    def double_sets(self):
        return [set(self.dbl_idx[self.dbl_ptr[i]:self.dbl_ptr[i + 1]].tolist()) for i in range(self.n_dofs)]


# vertex DoFs (FE_Q local numbering) at the ends of the four lines of a cell
CELL_LINES = ((0, 2), (1, 3), (0, 1), (2, 3))


def boundary_lines(mesh) -> np.ndarray:
    """(n_cells, 4) flags of lines that belong to one cell only (deal.II's
    line->at_boundary() on the surface mesh; double nodes are distinct
    DoFs, so lines along sharp edges count as boundary lines too)."""
    cd = mesh.cell_dofs
    a = np.stack([cd[:, i] for i, _ in CELL_LINES], axis=1)
    b = np.stack([cd[:, j] for _, j in CELL_LINES], axis=1)
    lo, hi = np.minimum(a, b).ravel(), np.maximum(a, b).ravel()
    key = lo.astype(np.int64) * (mesh.n_dofs + 1) + hi
    _, inv, cnt = np.unique(key, return_inverse=True, return_counts=True)
    return (cnt[inv] == 1).reshape(-1, 4)


def merge_constraints(cons_list, offsets):
    """Constraint CSRs (ptr, idx, val, inhom, is_constrained) of independent
    problems as one CSR over the concatenated DoFs (part s from offsets[s])."""
    ptr, idx, val, inh, isc = [np.zeros(1, dtype=np.int64)], [], [], [], []
    for (p, i, v, h, c), off in zip(cons_list, offsets):
        ptr.append(np.asarray(p[1:], dtype=np.int64) + ptr[-1][-1])
        idx.append(np.asarray(i, dtype=np.int64) + int(off))
        val.append(np.asarray(v, dtype=np.float64))
        inh.append(np.asarray(h, dtype=np.float64))
        isc.append(np.asarray(c, dtype=np.uint8))
    return (np.concatenate(ptr), np.concatenate(idx) if idx else np.zeros(0, np.int64),
            np.concatenate(val) if val else np.zeros(0), np.concatenate(inh), np.concatenate(isc))


def mark_lines(mesh, on_line, line_id: int):
    """Give line_id to every boundary line whose two end points satisfy
    on_line ((m, 3) points -> (m,) bools), e.g. a waterline z = z0."""
    if mesh.line_ids is None:
        mesh.line_ids = -np.ones((mesh.n_cells, 4), dtype=np.int32)
    bl = boundary_lines(mesh)
    for j, (u, v) in enumerate(CELL_LINES):
        ok = on_line(mesh.support_xyz[mesh.cell_dofs[:, u]]) & on_line(mesh.support_xyz[mesh.cell_dofs[:, v]])
        mesh.line_ids[bl[:, j] & ok, j] = line_id
    return mesh


def _face_lattice(k: int, face: int, i: np.ndarray, j: np.ndarray) -> np.ndarray:
    """Integer cube-surface lattice coordinates (0..k) of face point (i, j)."""
    n, u, v = (np.array(a) for a in _FACES[face])
    base = np.where(n > 0, k, 0) * np.abs(n)
    # along u and v the coordinate runs 0..k in the positive direction
    return base[None, :] + i[:, None] * np.abs(u)[None, :] + j[:, None] * np.abs(v)[None, :]


def _build(k: int, deg: int, shared: bool, mapper, normal_fn):
    """Cube-surface lattice mesh with k lattice intervals per face edge."""
    assert k % deg == 0
    nc = k // deg
    key_to_id: dict = {}
    pts, nrm = [], []
    cells, cface = [], []
    local = _Q1 if deg == 1 else _Q2
    for f in range(6):
        ii, jj = np.meshgrid(np.arange(k + 1), np.arange(k + 1), indexing="ij")
        lat = _face_lattice(k, f, ii.ravel(), jj.ravel())
        ids = np.empty(lat.shape[0], dtype=np.int64)
        for t, key in enumerate(map(tuple, lat)):
            kk = key if shared else (f,) + key
            d = key_to_id.get(kk)
            if d is None:
                d = len(pts)
                key_to_id[kk] = d
                c = 2.0 * np.array(key, dtype=np.float64) / k - 1.0
                x = mapper(c)
                pts.append(x)
                nrm.append(normal_fn(x, f))
            ids[t] = d
        grid = ids.reshape(k + 1, k + 1)
        for a in range(nc):
            for b in range(nc):
                cells.append([grid[a * deg + di, b * deg + dj] for (di, dj) in local])
                cface.append(f)
    return np.array(pts), np.array(cells, dtype=np.int64), np.array(nrm), np.array(cface, dtype=np.int32)


def _double_nodes(xyz: np.ndarray, tol: float = 1e-10):
    """Double-node CSR: DoFs whose support points coincide within tol."""
    n = xyz.shape[0]
    q = np.round(xyz / 1e-9).astype(np.int64)
    order = np.lexsort((q[:, 2], q[:, 1], q[:, 0]))
    groups = [[i] for i in range(n)]
    s = q[order]
    start = 0
    for t in range(1, n + 1):
        if t == n or np.any(s[t] != s[start]):
            if t - start > 1:
                members = order[start:t]
                for m in members:
                    groups[m] = sorted(int(x) for x in members
                                       if np.linalg.norm(xyz[x] - xyz[m]) < tol)
            start = t
    ptr = np.zeros(n + 1, dtype=np.int64)
    for i in range(n):
        ptr[i + 1] = ptr[i] + len(groups[i])
    idx = np.array([x for g in groups for x in g], dtype=np.int64)
    return ptr, idx


def cube_sphere(n_cells_per_edge: int, deg: int = 1, radii=(1.0, 1.0, 1.0), equiangular: bool = False) -> Mesh:
    """Sphere (or ellipsoid with semi-axes ``radii``) from a cube surface grid
    whose lattice points are normalised onto the unit sphere and then scaled.

    ``equiangular`` applies c -> tan(pi c / 4) on the cube face coordinates
    before normalising, the fixed node-stretching map used for the spheroid.
    """
    k = n_cells_per_edge * deg
    R = np.array(radii, dtype=np.float64)

    def mapper(c):
        if equiangular:
            c = np.tan(np.pi * c / 4.0)
        s = c / np.linalg.norm(c)
        return s * R

    def normal_fn(x, f):
        g = x / (R * R)
        return g / np.linalg.norm(g)

    xyz, cells, nrm, cf = _build(k, deg, True, mapper, normal_fn)
    m = Mesh(xyz, cells, 4 if deg == 1 else 9, nrm, cf, name=f"sphere{n_cells_per_edge}q{deg}")
    m.dbl_ptr, m.dbl_idx = _double_nodes(xyz)
    return m


def double_node_cube(n_cells_per_edge: int, half: float = 1.0) -> Mesh:
    """Cube [-half, half]^3 whose six faces are separate Q1 patches: DoFs on
    the edges are doubled (tripled at corners)."""
    k = n_cells_per_edge

    def mapper(c):
        return c * half

    def normal_fn(x, f):
        return np.array(_FACES[f][0], dtype=np.float64)

    xyz, cells, nrm, cf = _build(k, 1, False, mapper, normal_fn)
    m = Mesh(xyz, cells, 4, nrm, cf, name=f"dcube{n_cells_per_edge}")
    m.dbl_ptr, m.dbl_idx = _double_nodes(xyz)
    return m


def linear_map(mesh: Mesh, A) -> Mesh:
    """The mesh mapped by the linear map x -> A x: flat faces stay flat, the
    normals become A^-T n, normalised; the DoFs and double sets are kept."""
    A = np.asarray(A, dtype=np.float64)
    nrm = mesh.normals @ np.linalg.inv(A)
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    m = Mesh(mesh.support_xyz @ A.T, mesh.cell_dofs, mesh.dpc, nrm, mesh.cell_face, mesh.dbl_ptr, mesh.dbl_idx,
             name=f"{mesh.name}-mapped", line_ids=mesh.line_ids)
    return m


def wigley_hull(nx: int = 400, nz: int = 100, L: float = 1.0, B: float = 0.1, T: float = 0.0625,
                z_top: float = None) -> Mesh:
    """Closed Wigley-type double body y = +-(B/2)(1 - (2x/L)^2)(1 - (z/T)^2),
    |x| <= L/2, |z| <= T (the hull of draft T mirrored about z = 0): two Q1
    sheets of nx x nz cells, one per sign of y, meeting at the stem and stern
    (x = +-L/2) and at the keel lines (z = +-T), where each sheet keeps its own
    DoFs, so the DoFs there are double nodes.  Each DoF carries the exact
    outward normal of its own sheet.  z_top = 0 keeps only the submerged half
    z in [-T, 0] (open at the waterline), the surface a double-body model
    mirrors about z = 0."""
    xs = np.linspace(-L / 2, L / 2, nx + 1)
    zs = np.linspace(-T, T if z_top is None else z_top, nz + 1)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    fx = 1.0 - (2.0 * X / L) ** 2
    fz = 1.0 - (Z / T) ** 2
    F = 0.5 * B * fx * fz
    dFx = 0.5 * B * (-8.0 * X / L ** 2) * fz
    dFz = 0.5 * B * fx * (-2.0 * Z / T ** 2)
    pts, nrm, cells, cface = [], [], [], []
    nper = (nx + 1) * (nz + 1)
    for sheet, sgn in ((0, 1.0), (1, -1.0)):
        P = np.stack([X, sgn * F, Z], axis=-1).reshape(-1, 3)
        Nv = np.stack([-dFx, sgn * np.ones_like(F), -dFz], axis=-1).reshape(-1, 3)
        Nv /= np.linalg.norm(Nv, axis=1)[:, None]
        pts.append(P)
        nrm.append(Nv)
        ids = sheet * nper + np.arange(nper).reshape(nx + 1, nz + 1)
        a00, a10, a01, a11 = ids[:-1, :-1], ids[1:, :-1], ids[:-1, 1:], ids[1:, 1:]
        if sgn > 0:  # xi along z, eta along x: x_xi cross x_eta points to +y
            c = np.stack([a00, a01, a10, a11], axis=-1)
        else:        # xi along x, eta along z: points to -y
            c = np.stack([a00, a10, a01, a11], axis=-1)
        cells.append(c.reshape(-1, 4))
        cface.append(np.full(nx * nz, sheet, dtype=np.int32))
    xyz = np.concatenate(pts)
    m = Mesh(xyz, np.concatenate(cells).astype(np.int64), 4, np.concatenate(nrm), np.concatenate(cface),
             name=f"wigley{nx}x{nz}")
    m.dbl_ptr, m.dbl_idx = _double_nodes(xyz)
    return m


def unit_directions(seed: int, n: int) -> np.ndarray:
    """n seeded directions uniform on the unit sphere."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, 3))
    return u / np.linalg.norm(u, axis=1)[:, None]


def dirichlet_mask(mesh: Mesh, cell_is_dirichlet: np.ndarray) -> np.ndarray:
    mask = np.zeros(mesh.n_dofs, dtype=np.uint8)
    mask[mesh.cell_dofs[cell_is_dirichlet].ravel()] = 1
    return mask


def cell_centroids(mesh: Mesh) -> np.ndarray:
    return mesh.support_xyz[mesh.cell_dofs].mean(axis=1)


# ---------------------------------------------------------------------------
# Harmonic exact solutions (the BEM is the interior formulation of
# BEMProblem: alpha = -N.1, alpha phi + N phi - D dphi/dn = 0).
# ---------------------------------------------------------------------------
class PointSources:
    """phi(x) = c . x + sum_k s_k / |x - y_k|, harmonic away from the y_k."""

    def __init__(self, centres, strengths, linear=(0.0, 0.0, 0.0)):
        self.y = np.atleast_2d(np.asarray(centres, dtype=np.float64))
        self.s = np.asarray(strengths, dtype=np.float64).ravel()
        self.c = np.asarray(linear, dtype=np.float64)

    def phi(self, x):
        d = x[:, None, :] - self.y[None, :, :]
        r = np.linalg.norm(d, axis=2)
        return (self.s[None, :] / r).sum(axis=1) + x @ self.c

    def grad(self, x):
        d = x[:, None, :] - self.y[None, :, :]
        r = np.linalg.norm(d, axis=2)
        return -(self.s[None, :, None] * d / r[:, :, None] ** 3).sum(axis=1) + self.c[None, :]


def bc_vector(mesh: Mesh, mask: np.ndarray, sol: PointSources) -> np.ndarray:
    """tmp_rhs of BoundaryConditions::prepare_bem_vectors
    (boundary_conditions.cc:163-259): phi at Dirichlet DoFs, grad phi . n at
    the others, with the exact outward normals of each DoF's own face (the
    reference uses L2-projected normals there)."""
    x = mesh.support_xyz
    g = np.where(mask == 1, sol.phi(x), (sol.grad(x) * mesh.normals).sum(axis=1))
    return g


def seeded_sources(seed: int, n: int, radii, scale: float = 1.5):
    """n point sources at scale*radii on seeded uniform sphere directions with
    N(0,1) strengths (std::mt19937_64 is not used: numpy's PCG64 is the seed
    source here, fixed per configuration)."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    y = scale * v * np.asarray(radii)[None, :]
    s = rng.normal(size=n)
    return y, s


@dataclass
class Problem:
    mesh: Mesh
    dirichlet: np.ndarray
    exact: PointSources
    tmp_rhs: np.ndarray
    p: int
    max_nodes_per_leaf: int = 20
    levels: int = 10
    quad_order: int = 4
    sing_quad_order: int = 5
    # singular rule: 1 the Duffy split of the reference's dense solve
    # (bem_problem.cc:213-229), 0 the Telles rule of its BEMFMA (bem_fma.cc:127-131)
    sing_rule: int = 1
    # several right-hand sides of one operator: (n_rhs, N) boundary data and
    # the exact solution of each
    rhs_set: np.ndarray = None
    exact_set: list = None
    # double body: the mirror plane z = mirror_z0 (None: one body)
    mirror_z0: float = None


def config_sphere(n: int, p: int, deg: int = 1) -> Problem:
    """Configurations 1 (n=16, p=4) and 2 (n=64, p=6): unit sphere from a
    6-face cube with n x n cells per face; Dirichlet where the cell centroid
    has z > 0; phi = 1/|x - (2,1,1)|."""
    m = cube_sphere(n, deg)
    dmask = dirichlet_mask(m, cell_centroids(m)[:, 2] > 0)
    sol = PointSources([[2.0, 1.0, 1.0]], [1.0])
    return Problem(m, dmask, sol, bc_vector(m, dmask, sol), p)


def config_spheroid(n: int = 128, p: int = 8, seed: int = 3) -> Problem:
    """Configuration 3: prolate spheroid (4,1,1), Q2, equiangular cube grid
    (n x n cells per face), phi = x + 8 seeded exterior sources at 1.5x."""
    radii = (4.0, 1.0, 1.0)
    m = cube_sphere(n, 2, radii, equiangular=True)
    dmask = dirichlet_mask(m, cell_centroids(m)[:, 2] > 0)
    y, s = seeded_sources(seed, 8, radii)
    sol = PointSources(y, s, linear=(1.0, 0.0, 0.0))
    return Problem(m, dmask, sol, bc_vector(m, dmask, sol), p)


def config_dcube(n: int, p: int, seed: int = 4) -> Problem:
    """Double-node cube with a seeded 8-source exterior solution, Dirichlet on
    z > 0 cells."""
    m = double_node_cube(n)
    dmask = dirichlet_mask(m, cell_centroids(m)[:, 2] > 0)
    y, s = seeded_sources(seed, 8, (1.0, 1.0, 1.0), scale=2.5)
    sol = PointSources(y, s)
    return Problem(m, dmask, sol, bc_vector(m, dmask, sol), p)


def config_wigley(nx: int = 400, nz: int = 100, p: int = 8, n_rhs: int = 32, seed: int = 5) -> Problem:
    """Configuration 5: the Wigley double body (wigley_hull) with n_rhs
    uniform inflows U_k (seeded directions on the unit sphere): pure Neumann
    data dphi/dn = -U_k . n, exact interior solution phi_k = -U_k . x + C.
    tmp_rhs / exact are the first right-hand side; rhs_set / exact_set all."""
    m = wigley_hull(nx, nz)
    dmask = np.zeros(m.n_dofs, dtype=np.uint8)
    dirs = unit_directions(seed, n_rhs)
    exact_set = [PointSources(np.zeros((0, 3)), [], linear=-u) for u in dirs]
    rhs_set = np.stack([bc_vector(m, dmask, e) for e in exact_set])
    return Problem(m, dmask, exact_set[0], rhs_set[0], p, rhs_set=rhs_set, exact_set=exact_set)


def config_batch(count: int = 64, p: int = 6, seed: int = 4, n_min: int = 1500, n_max: int = 25000):
    """Configuration 4: count independent problems of uneven size from one
    seeded stream: each a sphere, a prolate spheroid of aspect 1..6 or a
    double-node cube with N drawn uniformly in [n_min, n_max] and rounded to
    the nearest grid size, Q1, Dirichlet on cells with centroid z > 0, and its
    own 8-source exterior solution (sources at 1.5x the radii, 2.5x for cubes).
    Returns a list of Problems."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        kind = int(rng.integers(3))
        target = rng.uniform(n_min, n_max)
        src_seed = int(rng.integers(1 << 31))
        if kind == 2:
            k = max(2, int(round(math.sqrt(target / 6.0))) - 1)
            m = double_node_cube(k)
            radii, scale = (1.0, 1.0, 1.0), 2.5
        else:
            k = max(2, int(round(math.sqrt((target - 2.0) / 6.0))))
            aspect = 1.0 if kind == 0 else float(rng.uniform(1.0, 6.0))
            radii, scale = (aspect, 1.0, 1.0), 1.5
            m = cube_sphere(k, 1, radii, equiangular=kind == 1)
        m.name = f"batch{b}-{m.name}"
        dmask = dirichlet_mask(m, cell_centroids(m)[:, 2] > 0)
        y, s = seeded_sources(src_seed, 8, radii, scale=scale)
        sol = PointSources(y, s)
        out.append(Problem(m, dmask, sol, bc_vector(m, dmask, sol), p))
    return out


def config_wigley_half(nx: int = 400, nz: int = 50, p: int = 8, n_rhs: int = 4, seed: int = 5) -> Problem:
    """The submerged half of the Wigley hull (z in [-T, 0]) as a double body
    mirrored about z = 0, with horizontal inflows U_k (z component dropped, so
    that the exact solution -U_k . x is even in z): the same problem as
    config_wigley with 2 nz cells over the full depth."""
    m = wigley_hull(nx, nz, z_top=0.0)
    dmask = np.zeros(m.n_dofs, dtype=np.uint8)
    dirs = unit_directions(seed, n_rhs) * np.array([1.0, 1.0, 0.0])
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    exact_set = [PointSources(np.zeros((0, 3)), [], linear=-u) for u in dirs]
    rhs_set = np.stack([bc_vector(m, dmask, e) for e in exact_set])
    return Problem(m, dmask, exact_set[0], rhs_set[0], p, rhs_set=rhs_set, exact_set=exact_set, mirror_z0=0.0)


def constraints(prob: Problem):
    """Double-node constraints of BEMProblem::compute_constraints
    (bem_problem.cc:1376-1653) for the exact boundary data.

    Rows (CSR): for each double-node group, the first Dirichlet DoF (else the
    smallest id) is the master.  Neumann partners of a Neumann master get
    phi_it = phi_i (entry 1).  Neumann partners of a Dirichlet master take the
    master's Dirichlet value as inhomogeneity.  Dirichlet-Dirichlet pairs get
    the continuous-gradient normal derivatives as inhomogeneities, here from
    the exact gradient (the reference projects surface gradients).  The
    reference's dead normal test (normal_distance /= normal_distance) is not
    reproduced: coplanar Dirichlet doubles cannot occur on these bodies.
    Returns (ptr, idx, val, inhom, is_constrained)."""
    m = prob.mesh
    n = m.n_dofs
    dsets = m.double_sets()
    dmask = prob.dirichlet
    g = prob.tmp_rhs
    grad = prob.exact.grad(m.support_xyz)
    entries = [[] for _ in range(n)]
    inhom = np.zeros(n)
    isc = np.zeros(n, dtype=np.uint8)
    for i in range(n):
        doubles = sorted(dsets[i])
        first = doubles[0]
        for d in doubles:
            if dmask[d] == 1:
                first = d
                break
        if i != first:
            continue
        others = [d for d in doubles if d != i]
        if dmask[i] == 1:
            for it in others:
                if dmask[it] == 1:
                    isc[i] = 1
                    inhom[i] = float(grad[i] @ m.normals[i])
                    isc[it] = 1
                    inhom[it] = float(grad[it] @ m.normals[it])
                else:
                    isc[it] = 1
                    inhom[it] = g[i]
        else:
            for it in others:
                isc[it] = 1
                entries[it].append((i, 1.0))
    ptr = np.zeros(n + 1, dtype=np.int64)
    for i in range(n):
        ptr[i + 1] = ptr[i] + len(entries[i])
    idx = np.array([e[0] for row in entries for e in row], dtype=np.int64)
    val = np.array([e[1] for row in entries for e in row], dtype=np.float64)
    return ptr, idx, val, inhom, isc
