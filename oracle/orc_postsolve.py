"""TEST INFRASTRUCTURE ONLY: numpy restatement of pi-BEM's post-solve steps.

Only tests/ import this module, and only to check the GPU path against it.
It restates, cell by cell with FEValues-style quadrature,

* BEMProblem::compute_normals            source/bem_problem.cc:1997-2079
* BEMProblem::compute_gradients          source/bem_problem.cc:1746-1876
* BEMProblem::compute_surface_gradients  source/bem_problem.cc:1878-1995
* BEMProblem::hydrostatic_pressure       source/bem_problem.cc:2081-2117
* BEMProblem::hydrodynamic_pressure      source/bem_problem.cc:2119-2167
* BEMProblem::center_of_pressure         source/bem_problem.cc:2169-2219
* BEMProblem::pressure_force_and_moment  source/bem_problem.cc:2221-2272
* BEMProblem::volume_integral            source/bem_problem.cc:2274-2309
* BEMProblem::free_surface_elevation     source/bem_problem.cc:2311-2370

on a Q1/Q2 surface mesh whose geometry is the element interpolant of the
support points (MappingFEField over the position vector,
bem_problem.cc:118-121).  The vector element is three copies of the scalar
one; the reference's vector constraints hold hanging nodes only (none on
uniformly refined meshes), so each projection is one scalar mass matrix with
three right-hand sides, solved here by a sparse direct factorisation (the
reference iterates GMRES + AMG to its solver tolerance).

The Gauss rule comes from numpy.polynomial.legendre.leggauss and the shape
functions from the 1-D Lagrange bases below, independently of the product's
set-up code.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# FE_Q local numbering: (i_x, i_y) of each local DoF on the 1-D support
# points (0, 1, 1/2): vertices, then lines x=0, x=1, y=0, y=1, then centre
_Q_IJ = {4: [(0, 0), (1, 0), (0, 1), (1, 1)],
         9: [(0, 0), (1, 0), (0, 1), (1, 1), (0, 2), (1, 2), (2, 0), (2, 1), (2, 2)]}
_NODES1 = np.array([0.0, 1.0, 0.5])


def _lagrange1(deg, t):
    """Values and derivatives of the 1-D Lagrange basis on _NODES1[:deg+1]."""
    nodes = _NODES1[:deg + 1]
    t = np.asarray(t, dtype=np.float64)
    L = np.ones((deg + 1,) + t.shape)
    dL = np.zeros((deg + 1,) + t.shape)
    for a in range(deg + 1):
        others = [b for b in range(deg + 1) if b != a]
        for b in others:
            L[a] *= (t - nodes[b]) / (nodes[a] - nodes[b])
        for c in others:
            term = np.ones_like(t) / (nodes[a] - nodes[c])
            for b in others:
                if b != c:
                    term = term * (t - nodes[b]) / (nodes[a] - nodes[b])
            dL[a] += term
    return L, dL


def shape(dpc, xi, eta):
    """N, dN/dxi, dN/deta of the FE_Q element at points (xi, eta): (dpc, m)."""
    deg = 1 if dpc == 4 else 2
    Lx, dLx = _lagrange1(deg, xi)
    Ly, dLy = _lagrange1(deg, eta)
    ij = _Q_IJ[dpc]
    N = np.array([Lx[i] * Ly[j] for i, j in ij])
    Nx = np.array([dLx[i] * Ly[j] for i, j in ij])
    Ny = np.array([Lx[i] * dLy[j] for i, j in ij])
    return N, Nx, Ny


def gauss01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


class PostSolve:
    """Post-solve of one mesh (pibem_amd.mesh.Mesh) with QGauss(quad_order)."""

    def __init__(self, mesh, quad_order=4):
        self.mesh = mesh
        self.nq1 = quad_order
        g, w = gauss01(quad_order)
        xi, eta = np.meshgrid(g, g, indexing="xy")  # x fastest
        self.xi, self.eta = xi.ravel(), eta.ravel()
        self.w = np.outer(w, w).ravel()
        self.g1 = g
        dpc = mesh.dpc
        N, Nx, Ny = shape(dpc, self.xi, self.eta)  # (dpc, nq)
        G = mesh.support_xyz[mesh.cell_dofs]        # (nc, dpc, 3)
        self.N = N
        self.X = np.einsum("jq,cjk->cqk", N, G)
        self.T1 = np.einsum("jq,cjk->cqk", Nx, G)
        self.T2 = np.einsum("jq,cjk->cqk", Ny, G)
        cr = np.cross(self.T1, self.T2)
        J = np.linalg.norm(cr, axis=2)
        self.n = cr / J[..., None]                  # t1 x t2 / |t1 x t2|
        self.JxW = J * self.w[None, :]
        # surface gradients of the shape functions: J (J^T J)^{-1} grad_ref
        g11 = np.einsum("cqk,cqk->cq", self.T1, self.T1)
        g12 = np.einsum("cqk,cqk->cq", self.T1, self.T2)
        g22 = np.einsum("cqk,cqk->cq", self.T2, self.T2)
        det = g11 * g22 - g12 * g12
        a = (g22[:, None, :] * Nx[None] - g12[:, None, :] * Ny[None]) / det[:, None, :]
        b = (g11[:, None, :] * Ny[None] - g12[:, None, :] * Nx[None]) / det[:, None, :]
        self.gradN = a[..., None] * self.T1[:, None] + b[..., None] * self.T2[:, None]  # (nc, dpc, nq, 3)
        # consistent scalar mass matrix
        cd = mesh.cell_dofs
        Ml = np.einsum("iq,jq,cq->cij", N, N, self.JxW)
        rows = np.repeat(cd, dpc, axis=1).ravel()
        cols = np.tile(cd, (1, dpc)).ravel()
        self.M = sp.csc_matrix((Ml.ravel(), (rows, cols)), shape=(mesh.n_dofs, mesh.n_dofs))
        self._lu = spla.splu(self.M)

    # -- projections ------------------------------------------------------
    def _project(self, f):
        """Solve M u_d = int N_i f_d for a field f given at the quadrature
        points, (nc, nq, 3) -> (3 n,) component-major."""
        n = self.mesh.n_dofs
        loc = np.einsum("iq,cqd,cq->cdi", self.N, f, self.JxW)  # (nc, 3, dpc)
        b = np.zeros((3, n))
        for d in range(3):
            np.add.at(b[d], self.mesh.cell_dofs.ravel(), loc[:, d, :].ravel())
        return np.concatenate([self._lu.solve(b[d]) for d in range(3)])

    def compute_normals(self):
        return self._project(self.n)

    def _surf_grad(self, u):
        uc = u[self.mesh.cell_dofs]  # (nc, dpc)
        return np.einsum("cj,cjqk->cqk", uc, self.gradN)

    def compute_gradients(self, phi, dphi_dn):
        dn = np.einsum("jq,cj->cq", self.N, dphi_dn[self.mesh.cell_dofs])
        return self._project(self.n * dn[..., None] + self._surf_grad(phi))

    def compute_surface_gradients(self, g):
        return self._project(self._surf_grad(g))

    # -- double-node constraints -------------------------------------------
    def compute_constraints(self, tmp_rhs, dirichlet, continuous_gradient=True):
        """BEMProblem::compute_constraints (bem_problem.cc:1376-1653) with
        this module's projected normals and surface gradients of tmp_rhs
        (bem_problem.cc:1385-1388).  For every double-node set, in DoF
        order, the master is the set's first Dirichlet DoF, else its smallest
        (bem_problem.cc:1405-1416).  Partners of a Neumann master get the
        entry (master, 1) (bem_problem.cc:1567-1575); a Neumann partner of a
        Dirichlet master takes tmp_rhs[master] as inhomogeneity
        (bem_problem.cc:1558-1562); a Dirichlet pair with continuous
        gradients sets both normal derivatives from
        (s_it.n_i + (s_i.n_it)(n_i.n_it)) / (1 - (n_i.n_it)^2) and its
        mirror (bem_problem.cc:1499-1556), the last partner's value staying
        on the master.  The reference's parallel-normal test divides the dot
        product by itself (bem_problem.cc:1490-1491) and never fires; here,
        as in the product, normals within 1e-4 of parallel (|cos| > 1 - 1e-4,
        where that formula divides by nearly zero) give the entry
        (master, sign cos).  Returns (ptr, idx, val, inhom, is_constrained)
        with component-major projections as in the product."""
        m = self.mesh
        n = m.n_dofs
        nrm = self.compute_normals().reshape(3, n).T
        sg = self.compute_surface_gradients(np.asarray(tmp_rhs, dtype=np.float64)).reshape(3, n).T
        dsets = m.double_sets()
        entries = [[] for _ in range(n)]
        inhom = np.zeros(n)
        isc = np.zeros(n, dtype=np.uint8)
        for i in range(n):
            doubles = sorted(dsets[i])
            first = next((d for d in doubles if dirichlet[d] == 1), doubles[0])
            if i != first:
                continue
            for it in doubles:
                if it == i:
                    continue
                if dirichlet[i] == 0:
                    isc[it] = 1
                    entries[it].append((i, 1.0))
                elif dirichlet[it] == 0:
                    isc[it] = 1
                    inhom[it] = tmp_rhs[i]
                else:
                    nn = float(nrm[i] @ nrm[it])
                    c = nn / (np.linalg.norm(nrm[i]) * np.linalg.norm(nrm[it]))
                    if 1.0 - abs(c) < 1e-4:
                        isc[it] = 1
                        entries[it].append((i, 1.0 if c > 0 else -1.0))
                    elif continuous_gradient:
                        a = float(sg[it] @ nrm[i])   # s_it . n_i
                        b = float(sg[i] @ nrm[it])   # s_i . n_it
                        isc[i] = 1
                        inhom[i] = (a + b * nn) / (1.0 - nn * nn)
                        isc[it] = 1
                        inhom[it] = (b + a * nn) / (1.0 - nn * nn)
        ptr = np.zeros(n + 1, dtype=np.int64)
        ptr[1:] = np.cumsum([len(r) for r in entries])
        idx = np.array([e[0] for r in entries for e in r], dtype=np.int64)
        val = np.array([e[1] for r in entries for e in r], dtype=np.float64)
        return ptr, idx, val, inhom, isc

    # -- pressures --------------------------------------------------------
    def hydrostatic_pressure(self, gravity, density, z0):
        return -density * gravity * (self.mesh.support_xyz[:, 2] - z0)

    @staticmethod
    def hydrodynamic_pressure(density, grad, wind):
        """grad, wind: (3 n) component-major."""
        u = grad.reshape(3, -1)
        W = wind.reshape(3, -1)
        return -density * np.sum(u * (0.5 * u - W), axis=0)

    # -- body integrals ---------------------------------------------------
    def _pq(self, p):
        return np.einsum("jq,cj->cq", self.N, p[self.mesh.cell_dofs])

    def center_of_pressure(self, mask, p, center=None):
        c = np.zeros(3) if center is None else np.array(center, dtype=np.float64)
        m = np.asarray(mask, dtype=bool)
        w = (self._pq(p) * self.JxW)[m]
        i2 = w.sum()
        if abs(i2) > 0.0:
            c = np.einsum("cqk,cq->k", self.X[m], w) / i2
        return c

    def pressure_force_and_moment(self, mask, p, cog):
        m = np.asarray(mask, dtype=bool)
        w = (self._pq(p) * self.JxW)[m]
        n = self.n[m]
        r = np.asarray(cog)[None, None, :] - self.X[m]
        force = np.einsum("cq,cqk->k", w, n)
        moment = np.einsum("cq,cqk->k", w, np.cross(r, n))
        return force, moment

    def volume_integral(self, mask):
        m = np.asarray(mask, dtype=bool)
        return -np.einsum("cqk,cqk,cq->k", self.X[m], self.n[m], self.JxW[m])

    def free_surface_elevation(self, gravity, density, cell_lines, p):
        """(x, y, p/(rho g)) at the 1-D Gauss points of each (cell, line)."""
        out = []
        for c, f in cell_lines:
            t = self.g1
            xi = np.zeros_like(t) if f == 0 else np.ones_like(t) if f == 1 else t
            eta = np.zeros_like(t) if f == 2 else np.ones_like(t) if f == 3 else t
            N, _, _ = shape(self.mesh.dpc, xi, eta)
            G = self.mesh.support_xyz[self.mesh.cell_dofs[c]]
            x = N.T @ G
            pv = N.T @ p[self.mesh.cell_dofs[c]]
            out.append(np.stack([x[:, 0], x[:, 1], pv / (density * gravity)], axis=1))
        return np.concatenate(out) if out else np.zeros((0, 3))
