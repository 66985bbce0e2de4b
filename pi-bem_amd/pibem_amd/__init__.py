"""Host-side mirror of pi-BEM's BEM-FMM interface over libpbfmm (C ABI).

``BEMFMA`` carries the method set of ``BEMFMA<3>`` (include/bem_fma.h:74-202):
``init_fma``, ``generate_octree_blocking``, ``compute_m2l_flags``,
``direct_integrals``, ``multipole_integrals``,
``generate_multipole_expansions``, ``multipole_matr_vect_products``.
``BEMProblem`` carries the FMA branch of ``BEMProblem<3>``
(include/bem_problem.h:131-242): ``compute_alpha``, ``vmult``,
``compute_rhs``, ``solve_system``, and the post-solve steps
(``compute_normals``, ``compute_gradients``, ``compute_surface_gradients``,
pressures, body integrals and waterline elevation; bem_problem.cc:1746-2370)
over a ``Body`` (include/Body.h).  Errors raise ``PbfmmError`` (the reference
throws ``ExcMessage`` through ``AssertThrow``).  All computation runs in the
HIP library; nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from ._lib import PBFMM_FP64, PBFMM_MIXED, PBFMM_SING_DUFFY, PBFMM_SING_TELLES, Params, PbfmmError, load

__all__ = ["BEMFMA", "BatchFMA", "BEMProblem", "Body", "setup_batch", "tune_leaf_size", "leaf_walk", "PBFMM_FP64", "PBFMM_MIXED",
           "PBFMM_SING_TELLES", "PBFMM_SING_DUFFY", "PbfmmError", "load"]

_d = C.POINTER(C.c_double)
_i64 = C.POINTER(C.c_int64)
_i32 = C.POINTER(C.c_int32)
_u8 = C.POINTER(C.c_uint8)


def _p(a, t=_d):
    return None if a is None else a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int64)


class BEMFMA:
    """BEMFMA<3> over the C ABI.  Parameters as in BEMFMA::declare_parameters
    (bem_fma.cc:64-85): num_octree_levels=10, max_nodes_per_leaf=20,
    trunc_order=6; quadrature orders as handed to init_fma (4, 5).
    sing_rule selects the singular rule: PBFMM_SING_TELLES (BEMFMA's QTelles,
    bem_fma.cc:127-131) or PBFMM_SING_DUFFY (the dense assembly's
    QSplit(QDuffy), bem_problem.cc:213-229)."""

    def __init__(self, trunc_order=6, num_octree_levels=10, max_nodes_per_leaf=20, quad_order=4,
                 sing_quad_order=5, precision=PBFMM_FP64, device=0, stream=None, preconditioner=1,
                 sing_rule=PBFMM_SING_TELLES):
        self._L = load()
        prm = Params()
        self._L.pbfmm_default_params(C.byref(prm))
        prm.num_octree_levels = num_octree_levels
        prm.max_nodes_per_leaf = max_nodes_per_leaf
        prm.trunc_order = trunc_order
        prm.quad_order = quad_order
        prm.sing_quad_order = sing_quad_order
        prm.precision = precision
        prm.device = device
        prm.stream = stream
        prm.preconditioner = preconditioner
        prm.sing_rule = sing_rule
        self.params = prm
        h = C.c_void_p()
        rc = self._L.pbfmm_create(C.byref(prm), C.byref(h))
        if rc != 0:
            raise PbfmmError(rc, "pbfmm_create failed (invalid parameters or no HIP device)")
        self._h = h
        self.n = 0
        self.p = trunc_order
        self.levels = num_octree_levels

    def close(self):
        if getattr(self, "_h", None):
            self._L.pbfmm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc != 0:
            msg = self._L.pbfmm_last_error(self._h)
            raise PbfmmError(rc, msg.decode() if msg else "")

    # ---- BEMFMA<3> --------------------------------------------------------
    def init_fma(self, mesh, dirichlet_mask):
        """bem_fma.h:103-108: mesh = pibem_amd.mesh.Mesh (support points,
        FE_Q cell DoFs, double-node CSR), Dirichlet mask per DoF."""
        self._keep = [_f64(mesh.support_xyz), _i(mesh.cell_dofs),
                      None if mesh.dbl_ptr is None else _i(mesh.dbl_ptr),
                      None if mesh.dbl_idx is None else _i(mesh.dbl_idx),
                      np.ascontiguousarray(dirichlet_mask, dtype=np.uint8)]
        xyz, cd, dp, di, dm = self._keep
        self._chk(self._L.pbfmm_init_fma(self._h, mesh.n_dofs, _p(xyz), mesh.n_cells, mesh.dpc, _p(cd, _i64), None,
                                         _p(dp, _i64), _p(di, _i64), _p(dm, _u8)))
        self.n = mesh.n_dofs
        self.mesh = mesh

    def set_mixed_scheme(self, scheme=0):
        """Mixed-precision solver scheme: 0 chosen by the timed FP64/FP32
        operator cost ratio, 1 relaxed (FP64 then FP32 operator in one Krylov
        space), 2 iterative refinement (its second inner solve recycling the
        first's Krylov space), 3 refinement without recycling, 4 scheme 2
        with FP32 basis vectors in the inner solves
        (pbfmm_set_mixed_scheme)."""
        self._chk(self._L.pbfmm_set_mixed_scheme(self._h, int(scheme)))

    def set_mixed_tolerances(self, inner_reduction=0.0, relax_switch=0.0):
        """Refinement's inner residual reduction and the relaxed scheme's
        FP32 switch (pbfmm_set_mixed_tolerances; 0 = the defaults)."""
        self._chk(self._L.pbfmm_set_mixed_tolerances(self._h, float(inner_reduction), float(relax_switch)))

    def set_gmres_reorth(self, eta=-1.0):
        """GMRES Gram-Schmidt: the second projection of an Arnoldi step runs
        when the first left less than eta of the vector's norm (default
        0.05; >= 1 always, 0 never; inner solves on FP32 basis vectors
        always project twice; pbfmm_set_gmres_reorth)."""
        self._chk(self._L.pbfmm_set_gmres_reorth(self._h, float(eta)))

    def set_spmv_kernel(self, precision, kernel):
        """Overrides the autotuned near-field SpMV kernel for one precision: 0
        workgroup per leaf, 1 / 2 wavefront per leaf with 2 / 4 rows in
        flight; FP32 also 3..5, the same on the operator's values in row
        pairs, and 6, workgroup per leaf on them in column vectors (FP64: 3)
        (pbfmm_set_spmv_kernel)."""
        self._chk(self._L.pbfmm_set_spmv_kernel(self._h, int(precision), int(kernel)))

    def set_side_fusion(self, precision, mode):
        """Overrides the timed launch plan of the operator: 0 none, 1 SpMV
        beside an M2M level and M2P beside an L2L level, 2 M2P beside M2L
        instead, 3 both beside M2L; 4, 5 plans 0, 1 with the slot M2L; 8, 12
        plans 0, 4 with the SpMV on a second stream; 24, 28 forked after the
        M2M levels; 56, 60 with M2P after the SpMV there, 88, 92 before it;
        -1 times them again (pbfmm_set_side_fusion)."""
        self._chk(self._L.pbfmm_set_side_fusion(self._h, int(precision), int(mode)))

    def tuning_state(self):
        """Choices the handle's timings made so far (-1 = not timed yet):
        SpMV kernels, side fusion modes, mixed scheme, FP64/FP32 operator
        cost ratio (pbfmm_tuning_state)."""
        st = np.zeros(8, np.int32)
        self._chk(self._L.pbfmm_tuning_state(self._h, _p(st, _i32)))
        return dict(spmv_kernel_fp64=int(st[0]), spmv_kernel_fp32=int(st[1]), side_fusion_fp64=int(st[2]),
                    side_fusion_fp32=int(st[3]), mixed_scheme=int(st[4]),
                    op_cost_ratio=(st[5] / 1000.0 if st[5] >= 0 else None), graph_fp64=int(st[6]),
                    graph_fp32=int(st[7]))

    def set_preconditioner_rows(self, rows=128):
        """Row cap of the block-Jacobi blocks, 8..128
        (pbfmm_set_preconditioner_rows)."""
        self._chk(self._L.pbfmm_set_preconditioner_rows(self._h, int(rows)))

    def set_graph_mode(self, mode=-1):
        """GMRES step operator as a HIP graph replay: -1 timed choice, 0 off,
        1 on (pbfmm_set_graph_mode)."""
        self._chk(self._L.pbfmm_set_graph_mode(self._h, int(mode)))

    def graph_timings(self):
        """ms per step behind the graph choice: FP64 direct, graph, FP32
        direct, graph (pbfmm_graph_timings)."""
        ms = np.zeros(4, np.float32)
        self._chk(self._L.pbfmm_graph_timings(self._h, ms.ctypes.data_as(C.POINTER(C.c_float))))
        return ms.tolist()

    def set_double_body(self, enable=True, z0=0.0):
        """Double-body operator: every cell also acts through its mirror image
        about z = z0 (BEMProblem::_assemble_system_double_body,
        bem_problem.cc:957-1135).  Between init_fma and generate_octree_blocking."""
        self._chk(self._L.pbfmm_set_double_body(self._h, 1 if enable else 0, float(z0)))

    def generate_octree_blocking(self):
        self._chk(self._L.pbfmm_generate_octree_blocking(self._h))

    def compute_m2l_flags(self):
        out = np.zeros(self.levels + 1, dtype=np.int64)
        self._chk(self._L.pbfmm_compute_m2l_flags(self._h, _p(out, _i64)))
        return out

    def direct_integrals(self):
        self._chk(self._L.pbfmm_direct_integrals(self._h))

    def multipole_integrals(self):
        self._chk(self._L.pbfmm_multipole_integrals(self._h))

    def generate_multipole_expansions(self, phi, dphi_dn):
        self._chk(self._L.pbfmm_generate_multipole_expansions(self._h, _p(_f64(phi)), _p(_f64(dphi_dn))))

    def multipole_matr_vect_products(self, phi, dphi_dn):
        N, D = np.zeros(self.n), np.zeros(self.n)
        self._chk(self._L.pbfmm_multipole_matr_vect_products(self._h, _p(_f64(phi)), _p(_f64(dphi_dn)), _p(N),
                                                             _p(D)))
        return N, D

    # ---- device-resident entry points -----------------------------------
    def matvec_dev(self, phi_ptr, dphi_ptr, outN_ptr, outD_ptr, precision=PBFMM_FP64):
        self._chk(self._L.pbfmm_matvec_dev(self._h, phi_ptr, dphi_ptr, outN_ptr, outD_ptr, precision))

    def gmres_dev(self, rhs_ptr, x_ptr, rel_tol=1e-8, restart=100, max_it=1000, precision=PBFMM_FP64,
                  check=True):
        it, rr = C.c_int32(0), C.c_double(0)
        rc = self._L.pbfmm_gmres_dev(self._h, rhs_ptr, x_ptr, rel_tol, restart, max_it, precision, C.byref(it),
                                     C.byref(rr))
        if check:
            self._chk(rc)
        return it.value, rr.value

    # ---- introspection ------------------------------------------------------
    def gmres_dev_multi(self, rhs_ptr, x_ptr, nrhs, ld, rel_tol=1e-8, restart=100, max_it=1000,
                        precision=PBFMM_FP64):
        """nrhs device-resident right-hand sides (column r at rhs_ptr + 8 r ld),
        solved into x_ptr; returns (iterations, relative residuals) per column."""
        it = np.zeros(nrhs, np.int32)
        rr = np.zeros(nrhs)
        self._chk(self._L.pbfmm_gmres_dev_multi(self._h, rhs_ptr, x_ptr, nrhs, ld, rel_tol, restart, max_it,
                                                 precision, _p(it, _i32), _p(rr)))
        return it, rr

    def tree(self):
        cnt = np.zeros(8, np.int64)
        self._chk(self._L.pbfmm_tree_counts(self._h, _p(cnt, _i64)))
        nb, L, nleaf, ns, nil, nnil, nn, _ = (int(x) for x in cnt)
        t = dict(level=np.zeros(nb, np.int32), parent=np.zeros(nb, np.int64), pmin=np.zeros(3 * nb),
                 delta=np.zeros(nb), startL=np.zeros(L + 1, np.int64), endL=np.zeros(L + 1, np.int64),
                 childless=np.zeros(nleaf, np.int64), sub_ptr=np.zeros(nb + 1, np.int64),
                 il_ptr=np.zeros(ns + 1, np.int64), il_idx=np.zeros(nil, np.int64),
                 nil_ptr=np.zeros(ns + 1, np.int64), nil_idx=np.zeros(nnil, np.int64),
                 node_ptr=np.zeros(nb + 1, np.int64), node_idx=np.zeros(nn, np.int64))
        args = [_p(t["level"], _i32)]
        for k in ("parent", "pmin", "delta", "startL", "endL", "childless", "sub_ptr", "il_ptr", "il_idx", "nil_ptr",
                  "nil_idx", "node_ptr", "node_idx"):
            args.append(_p(t[k], _i64) if t[k].dtype == np.int64 else _p(t[k]))
        self._chk(self._L.pbfmm_tree_export(self._h, *args))
        t["pmin"] = t["pmin"].reshape(nb, 3)
        t["num_blocks"] = nb
        return t

    def nearfield(self, precision=PBFMM_FP64):
        cnt = np.zeros(8, np.int64)
        self._chk(self._L.pbfmm_tree_counts(self._h, _p(cnt, _i64)))
        nnz = int(cnt[7])
        rp, col = np.zeros(self.n + 1, np.int64), np.zeros(nnz, np.int64)
        vN, vD = np.zeros(nnz), np.zeros(nnz)
        self._chk(self._L.pbfmm_nearfield_export(self._h, precision, _p(rp, _i64), _p(col, _i64), _p(vN), _p(vD)))
        return rp, col, vN, vD

    def expansions(self, which="me"):
        nb = self.tree()["num_blocks"]
        kt = (self.p + 1) * (self.p + 2) // 2
        a, b = np.zeros(2 * kt * nb), np.zeros(2 * kt * nb)
        self._chk(self._L.pbfmm_get_expansions(self._h, 0 if which == "me" else 1, _p(a), _p(b)))
        return (a[0::2] + 1j * a[1::2]).reshape(nb, kt), (b[0::2] + 1j * b[1::2]).reshape(nb, kt)

    def set_phase_timing(self, on=True):
        self._chk(self._L.pbfmm_set_phase_timing(self._h, 1 if on else 0))

    def last_timings(self):
        """Phase times (ms) of the last matvec: spmv, p2m, m2m, l2l, m2l, m2p, l2p, total, and
        m2l_sum, the per-target sum of the slot M2L (0 when M2L adds into its targets)."""
        ms = np.zeros(9)
        self._chk(self._L.pbfmm_last_timings(self._h, _p(ms)))
        return dict(zip(("spmv", "p2m", "m2m", "l2l", "m2l", "m2p", "l2p", "total", "m2l_sum"), ms.tolist()))

    def work_counts(self):
        c = np.zeros(16, np.int64)
        self._chk(self._L.pbfmm_work_counts(self._h, _p(c, _i64)))
        keys = ("n_dofs", "near_values", "near_cols", "near_rows", "leaves", "quad_points", "m2l_pairs",
                "m2p_evals", "m2m_translations", "l2l_translations", "num_blocks", "trunc_order", "kp",
                "p2m_cols", "spmv_kernel_fp64", "spmv_kernel_fp32")
        return dict(zip(keys, (int(x) for x in c)))


class BEMProblem:
    """FMA branch of BEMProblem<3> (bem_problem.cc:1138-1347) over a BEMFMA."""

    def __init__(self, fma: BEMFMA):
        self.fma = fma

    @property
    def _h(self):
        return self.fma._h

    def compute_alpha(self):
        a = np.zeros(self.fma.n)
        self.fma._chk(self.fma._L.pbfmm_compute_alpha(self._h, _p(a)))
        return a

    def set_constraints(self, cons):
        """cons = (ptr, idx, val, inhom, is_constrained) as produced by
        pibem_amd.mesh.constraints (compute_constraints, bem_problem.cc:1376-1653)."""
        if cons is None:
            self.fma._chk(self.fma._L.pbfmm_set_constraints(self._h, None, None, None, None, None))
            return
        ptr, idx, val, inh, isc = cons
        self._ckeep = [_i(ptr), _i(idx), _f64(val), _f64(inh), np.ascontiguousarray(isc, np.uint8)]
        a, b, c, d, e = self._ckeep
        self.fma._chk(self.fma._L.pbfmm_set_constraints(self._h, _p(a, _i64), _p(b, _i64), _p(c), _p(d),
                                                        _p(e, _u8)))

    def compute_constraints(self, tmp_rhs, continuous_gradient=True):
        """BEMProblem::compute_constraints (bem_problem.cc:1376-1653): the
        double-node rows from the projected normals and surface gradients of
        tmp_rhs, installed on the handle; returns them as
        (ptr, idx, val, inhom, is_constrained)."""
        self.fma._chk(self.fma._L.pbfmm_compute_constraints(self._h, _p(_f64(tmp_rhs)), 1 if continuous_gradient else 0))
        return self.constraints()

    def constraints(self):
        """The installed constraint rows (pbfmm_get_constraints)."""
        n = self.fma.n
        nnz = C.c_int64(0)
        self.fma._chk(self.fma._L.pbfmm_get_constraints(self._h, C.byref(nnz), None, None, None, None, None))
        ptr, idx, val = np.zeros(n + 1, np.int64), np.zeros(nnz.value, np.int64), np.zeros(nnz.value)
        inh, isc = np.zeros(n), np.zeros(n, np.uint8)
        self.fma._chk(self.fma._L.pbfmm_get_constraints(self._h, C.byref(nnz), _p(ptr, _i64), _p(idx, _i64), _p(val),
                                                        _p(inh), _p(isc, _u8)))
        return ptr, idx, val, inh, isc

    def fma_preconditioner(self, alpha=None):
        """BEMFMA::FMA_preconditioner (bem_fma.cc:3078-3411) as block Jacobi over
        the rows of the leaves that share a parent block."""
        a = None if alpha is None else _f64(alpha)
        self.fma._chk(self.fma._L.pbfmm_fma_preconditioner(self._h, _p(a)))

    def preconditioner_blocks(self):
        """The block-Jacobi preconditioner before inversion
        (pbfmm_preconditioner_blocks): a list of (rows, R x R matrix)."""
        sz = np.zeros(3, np.int64)
        self.fma._chk(self.fma._L.pbfmm_preconditioner_blocks(self._h, _p(sz, _i64), None, None, None))
        nb, nr, nv = (int(x) for x in sz)
        rb, rows, vals = np.zeros(nb + 1, np.int32), np.zeros(nr, np.int32), np.zeros(nv)
        self.fma._chk(self.fma._L.pbfmm_preconditioner_blocks(self._h, _p(sz, _i64), _p(rb, _i32), _p(rows, _i32),
                                                              _p(vals)))
        out, off = [], 0
        for b in range(nb):
            r = rows[rb[b]:rb[b + 1]]
            out.append((r, vals[off:off + len(r) ** 2].reshape(len(r), len(r))))
            off += len(r) ** 2
        return out

    def vmult(self, src):
        dst = np.zeros(self.fma.n)
        self.fma._chk(self.fma._L.pbfmm_vmult(self._h, _p(_f64(src)), _p(dst)))
        return dst

    def vmult_dev(self, src_ptr, dst_ptr, precision=PBFMM_FP64):
        """vmult on device vectors (pbfmm_vmult_dev); precision PBFMM_MIXED runs
        the FP32 operator."""
        self.fma._chk(self.fma._L.pbfmm_vmult_dev(self._h, src_ptr, dst_ptr, precision))

    def vmult_dev_multi(self, src_ptr, dst_ptr, nrhs, ld=None, precision=PBFMM_FP64):
        """vmult_dev on nrhs device columns (column r at ptr + 8*r*ld) in one
        operator application (pbfmm_vmult_dev_multi)."""
        ld = self.fma.n if ld is None else int(ld)
        self.fma._chk(self.fma._L.pbfmm_vmult_dev_multi(self._h, src_ptr, dst_ptr, int(nrhs), ld, precision))

    def compute_rhs(self, tmp_rhs):
        dst = np.zeros(self.fma.n)
        self.fma._chk(self.fma._L.pbfmm_compute_rhs(self._h, _p(_f64(tmp_rhs)), _p(dst)))
        return dst

    def solve_system(self, tmp_rhs, tol=1e-8, restart=100, max_it=1000, phi=None, dphi_dn=None):
        phi = np.zeros(self.fma.n) if phi is None else _f64(phi).copy()
        dphi = np.zeros(self.fma.n) if dphi_dn is None else _f64(dphi_dn).copy()
        it, rr = C.c_int32(0), C.c_double(0)
        self.fma._chk(self.fma._L.pbfmm_solve_system(self._h, _p(_f64(tmp_rhs)), _p(phi), _p(dphi), tol, restart,
                                                     max_it, C.byref(it), C.byref(rr)))
        return phi, dphi, it.value, rr.value

    # ---- post-solve (bem_problem.cc:1746-2370) ----
    def _vec3(self, fn, *args):
        out = np.zeros(3 * self.fma.n)
        self.fma._chk(fn(self._h, *args, _p(out)))
        return out

    def compute_normals(self):
        """BEMProblem::compute_normals: L2-projected, unnormalised nodal
        normals, component-major (3 n)."""
        self.vector_normals_solution = self._vec3(self.fma._L.pbfmm_compute_normals)
        return self.vector_normals_solution

    def compute_gradients(self, phi, dphi_dn):
        """BEMProblem::compute_gradients: L2 projection of n dphi_dn + the
        surface gradient of phi (3 n, component-major)."""
        self.vector_gradients_solution = self._vec3(self.fma._L.pbfmm_compute_gradients, _p(_f64(phi)),
                                                    _p(_f64(dphi_dn)))
        return self.vector_gradients_solution

    def compute_surface_gradients(self, tmp_rhs):
        """BEMProblem::compute_surface_gradients (3 n, component-major)."""
        self.vector_surface_gradients_solution = self._vec3(self.fma._L.pbfmm_compute_surface_gradients,
                                                            _p(_f64(tmp_rhs)))
        return self.vector_surface_gradients_solution

    def projection_iterations(self):
        it = C.c_int32(0)
        self.fma._chk(self.fma._L.pbfmm_projection_iterations(self._h, C.byref(it)))
        return it.value

    def hydrostatic_pressure(self, gravity, density, z0):
        p = np.zeros(self.fma.n)
        self.fma._chk(self.fma._L.pbfmm_hydrostatic_pressure(self._h, float(gravity), float(density), float(z0),
                                                             _p(p)))
        return p

    def hydrodynamic_pressure(self, density, wind):
        """wind: a function of the support points ((n, 3) -> (n, 3)) or its
        values as a (3 n) component-major array."""
        w = wind(self.fma.mesh.support_xyz).T.ravel() if callable(wind) else _f64(wind)
        p = np.zeros(self.fma.n)
        self.fma._chk(self.fma._L.pbfmm_hydrodynamic_pressure(self._h, float(density), _p(_f64(w)), _p(p)))
        return p

    def center_of_pressure(self, body, pressure, pressure_center=None):
        c = np.zeros(3) if pressure_center is None else _f64(pressure_center).copy()
        m = body.cell_mask(self.fma.mesh)
        self.fma._chk(self.fma._L.pbfmm_center_of_pressure(self._h, _p(m, _u8), _p(_f64(pressure)), _p(c)))
        return c

    def pressure_force_and_moment(self, body, pressure):
        f, mo = np.zeros(3), np.zeros(3)
        m = body.cell_mask(self.fma.mesh)
        cog = _f64(body.center_of_gravity)
        self.fma._chk(self.fma._L.pbfmm_pressure_force_and_moment(self._h, _p(m, _u8), _p(_f64(pressure)),
                                                                  _p(cog), _p(f), _p(mo)))
        return f, mo

    def pressure_force_and_moment_noalloc(self, mask_ptr, pressure_ptr, cog, out_ptr):
        """pbfmm_pressure_force_and_moment_noalloc: device pointers (mask may be
        None for every real cell; out receives force then moment, 6 doubles),
        partial sums in the library's fixed device array, no allocation;
        asynchronous on the handle's stream."""
        c = _f64(cog)
        self.fma._chk(self.fma._L.pbfmm_pressure_force_and_moment_noalloc(self._h, mask_ptr, pressure_ptr, _p(c),
                                                                       out_ptr))

    def volume_integral(self, body):
        v = np.zeros(3)
        m = body.cell_mask(self.fma.mesh)
        self.fma._chk(self.fma._L.pbfmm_volume_integral(self._h, _p(m, _u8), _p(v)))
        return v

    def free_surface_elevation(self, gravity, density, body, pressure):
        """(x, y, p / (density gravity)) at the Gauss points of the body's
        waterline boundary lines, in the reference's order; (m, 3)."""
        cl = np.ascontiguousarray(body.waterline_lines(self.fma.mesh), dtype=np.int32)
        nq = self.fma.params.quad_order
        out = np.zeros(3 * nq * cl.shape[0])
        self.fma._chk(self.fma._L.pbfmm_free_surface_elevation(
            self._h, float(gravity), float(density), cl.ctypes.data_as(_i32), cl.shape[0],
            _p(_f64(pressure)), _p(out)))
        return out.reshape(-1, 3)


class Body:
    """include/Body.h: the cells of a body (material ids), its waterline
    lines (manifold ids), draft and centre of gravity."""

    def __init__(self, material_indices, waterline_indices=(), draft=0.0, center_of_gravity=(0.0, 0.0, 0.0)):
        self.material_indices = [int(i) for i in material_indices]
        self.waterline_indices = [int(i) for i in waterline_indices]
        self.draft = float(draft)
        self.center_of_gravity = np.asarray(center_of_gravity, dtype=np.float64)

    def cell_mask(self, mesh):
        """Body::hasMaterial of each cell's material id (Mesh.cell_face)."""
        return np.ascontiguousarray(np.isin(mesh.cell_face, self.material_indices), dtype=np.uint8)

    def waterline_lines(self, mesh):
        """(cell, line) pairs of body cells whose line is at the boundary of
        the surface and carries a waterline manifold id, ascending
        (bem_problem.cc:2337-2348)."""
        if mesh.line_ids is None or not self.waterline_indices:
            return np.zeros((0, 2), dtype=np.int32)
        from .mesh import boundary_lines
        sel = (boundary_lines(mesh) & np.isin(mesh.line_ids, self.waterline_indices)
               & (self.cell_mask(mesh)[:, None] == 1))
        c, j = np.nonzero(sel)
        return np.stack([c, j], axis=1).astype(np.int32)


class BatchFMA(BEMFMA):
    """Independent problems as one block-diagonal operator
    (pbfmm_batch_create): the parts' octrees and plans are built in parallel
    on the host, the near field and multipoles of all parts are set up
    together, and every call of the one-problem path (matvecs, compute_alpha,
    vmult, compute_rhs, GMRES) acts on the concatenated vectors, part s at
    DoFs offsets[s]:offsets[s + 1], with one kernel launch per phase for all
    parts.  gmres_batch_dev reaches each part's own relative residual."""

    def __init__(self, parts):
        self._L = load()
        arr = (C.c_void_p * len(parts))(*[q._h.value for q in parts])
        h = C.c_void_p()
        rc = self._L.pbfmm_batch_create(arr, len(parts), C.byref(h))
        if rc != 0:
            msg = self._L.pbfmm_last_error(parts[0]._h)
            raise PbfmmError(rc, msg.decode() if msg else "pbfmm_batch_create failed")
        self._h = h
        self.parts = list(parts)
        self.params = parts[0].params
        self.p = parts[0].p
        self.levels = parts[0].levels
        off = np.zeros(len(parts) + 1, dtype=np.int64)
        self._chk(self._L.pbfmm_batch_offsets(self._h, _p(off, _i64)))
        self.offsets = off
        self.n = int(off[-1])

    def split(self, v):
        """A concatenated vector as a list of per-part vectors."""
        return [v[self.offsets[s]:self.offsets[s + 1]] for s in range(len(self.parts))]

    def gmres_batch_dev(self, rhs_ptr, x_ptr, rel_tol=1e-8, restart=100, max_it=1000, precision=PBFMM_FP64,
                        check=True):
        """(iterations, per-part true relative residuals)."""
        it = C.c_int32(0)
        rr = np.zeros(len(self.parts))
        rc = self._L.pbfmm_batch_gmres_dev(self._h, rhs_ptr, x_ptr, rel_tol, restart, max_it, precision,
                                           C.byref(it), _p(rr))
        if check:
            self._chk(rc)
        return it.value, rr


# Leaf sizes the tuner walks: a geometric ladder around the reference's 20.
LEAF_LADDER = (8, 12, 16, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160)


def tune_leaf_size(problem, precision=PBFMM_MIXED, solves=2, start=20, ladder=LEAF_LADDER, tol=1e-8,
                   device=0, log=None):
    """Leaf-size autotuner: sets the problem up with a leaf size
    (max_nodes_per_leaf) and times whole GMRES solves to relative residual
    tol in the given precision (one warm-up solve, then `solves` timed ones,
    wall clock between device synchronisations, right-hand side on the
    device), so the operator's cost, the preconditioner's blocks (cut from
    the leaves) and the step count all count.  From `start` it walks the
    ladder towards faster solves until a leaf size is slower than both its
    evaluated neighbours or the ladder ends, so the best one is bracketed
    where the ladder allows.  Returns (best leaf size, {leaf size: ms per
    solve}, {leaf size: (steps, relres, set-up s)}, bracketed).  The leaf
    size moves the near/far split, so solutions differ within the FMM
    truncation error; the reference's default is 20."""
    import dataclasses
    from . import hip
    times, info = {}, {}

    def measure(leaf):
        if leaf in times:
            return times[leaf]
        pr = dataclasses.replace(problem, max_nodes_per_leaf=int(leaf))
        t0 = time.perf_counter()
        f = setup(pr, precision, device)
        bp = BEMProblem(f)
        rhs = hip.DeviceArray.from_numpy(bp.compute_rhs(pr.tmp_rhs))
        hip.synchronize()
        t_setup = time.perf_counter() - t0
        x = hip.DeviceArray(f.n)
        zero = hip.DeviceArray(f.n)
        f.gmres_dev(rhs.value, x.value, tol, 100, 2000, precision)
        hip.synchronize()
        t0 = time.perf_counter()
        for _ in range(solves):
            x.copy_from(zero)
            its, rr = f.gmres_dev(rhs.value, x.value, tol, 100, 2000, precision)
        hip.synchronize()
        times[leaf] = (time.perf_counter() - t0) * 1e3 / solves
        info[leaf] = (its, rr, t_setup)
        f.close()
        if log:
            log(f"leaf {leaf}: {times[leaf]:.3f} ms per solve, {its} steps, set-up {t_setup:.2f} s")
        return times[leaf]

    best, bracketed = leaf_walk(measure, start, ladder)
    return best, times, info, bracketed


def leaf_walk(measure, start=20, ladder=LEAF_LADDER, gain=0.01):
    """The tuner's walk over the ladder: measure(leaf) -> cost (each leaf is
    measured once here).  From start it measures both neighbours, then walks
    towards the cheaper one (also when neither beats the start: a flat start
    does not end the walk) until three leaf sizes in a row do not beat the
    best so far (so that a plateau or a noisy point does not end the walk
    early), and then the same way from start in the other direction (so that
    one noisy neighbour cannot hide a minimum on its side).  A leaf size
    replaces the best only when it is cheaper by more than `gain` (1%), so
    timing noise does not move the choice.  Returns (best leaf, bracketed:
    both ladder neighbours of the best were measured and cost more)."""
    lad = list(ladder)
    cost = {}

    def m(leaf):
        if leaf not in cost:
            cost[leaf] = measure(leaf)
        return cost[leaf]

    def beats(leaf, best):
        return m(leaf) < cost[best] * (1.0 - gain)

    k = lad.index(start) if start in lad else 0
    m(lad[k])
    nbr = [j for j in (k - 1, k + 1) if 0 <= j < len(lad)]
    for j in nbr:
        m(lad[j])
    jc = min(nbr, key=lambda j: cost[lad[j]]) if nbr else k
    first = 1 if jc > k else -1
    best, kb = lad[k], k
    for step in ((first, -first) if nbr else ()):
        j, misses = k + step, 0
        while 0 <= j < len(lad) and misses < 3:
            if beats(lad[j], best):
                best, kb, misses = lad[j], j, 0
            else:
                misses += 1
            j += step
    bracketed = 0 < kb < len(lad) - 1 and lad[kb - 1] in cost and lad[kb + 1] in cost
    return best, bracketed


def setup_batch(problems, precision=PBFMM_FP64, device=0, preconditioner=1):
    """A BatchFMA over problems (init_fma per part; the rest of the set-up
    runs once for the batch)."""
    parts = []
    for pr in problems:
        f = BEMFMA(trunc_order=pr.p, num_octree_levels=pr.levels, max_nodes_per_leaf=pr.max_nodes_per_leaf,
                   quad_order=pr.quad_order, sing_quad_order=pr.sing_quad_order, precision=precision,
                   device=device, preconditioner=preconditioner, sing_rule=pr.sing_rule)
        f.init_fma(pr.mesh, pr.dirichlet)
        parts.append(f)
    return BatchFMA(parts)


def setup(problem, precision=PBFMM_FP64, device=0, preconditioner=1):
    """The set-up sequence of BEMProblem::solve (bem_problem.cc:1363-1371,
    upstream wiring): init_fma, octree, direct and multipole integrals."""
    f = BEMFMA(trunc_order=problem.p, num_octree_levels=problem.levels,
               max_nodes_per_leaf=problem.max_nodes_per_leaf, quad_order=problem.quad_order,
               sing_quad_order=problem.sing_quad_order, precision=precision, device=device,
               preconditioner=preconditioner, sing_rule=problem.sing_rule)
    f.init_fma(problem.mesh, problem.dirichlet)
    if getattr(problem, "mirror_z0", None) is not None:
        f.set_double_body(True, problem.mirror_z0)
    f.generate_octree_blocking()
    f.direct_integrals()
    f.multipole_integrals()
    return f
