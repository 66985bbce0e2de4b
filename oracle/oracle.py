"""TEST INFRASTRUCTURE ONLY: ctypes front end of the CPU oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only to check or time against it.  The oracle is a plain FP64
C++ restatement of pi-BEM's BEM-FMM path (source/bem_fma.cc,
source/multipole_expansion.cc, source/local_expansion.cc, source/bem_problem.cc)
that follows the reference's loop order and data structures.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = C.POINTER(C.c_double)
_i64 = C.POINTER(C.c_int64)
_i32 = C.POINTER(C.c_int32)
_u8 = C.POINTER(C.c_uint8)


def build(force: bool = False) -> str:
    """Compile oracle/liboracle.so with the Makefile beside this file."""
    cmd = ["make", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int] * 5
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_init.argtypes = [C.c_void_p, C.c_int64, _d, C.c_int64, C.c_int, _i64, _d, _i64, _i64, _u8]
        L.orc_set_double_body.argtypes = [C.c_void_p, C.c_double]
        L.orc_set_sing_rule.argtypes = [C.c_void_p, C.c_int]
        for f in ("orc_generate_octree_blocking", "orc_direct_integrals", "orc_multipole_integrals"):
            getattr(L, f).argtypes = [C.c_void_p]
        L.orc_generate_multipole_expansions.argtypes = [C.c_void_p, _d, _d]
        L.orc_multipole_matr_vect_products.argtypes = [C.c_void_p, _d, _d, _d, _d]
        L.orc_tree_counts.argtypes = [C.c_void_p, _i64]
        L.orc_tree_export.argtypes = [C.c_void_p, _i32, _i64, _d, _d, _i64, _i64, _i64, _i64,
                                      _i64, _i64, _i64, _i64, _i64, _i64]
        L.orc_csr_nnz.restype = C.c_int64
        L.orc_csr_nnz.argtypes = [C.c_void_p]
        L.orc_csr_export.argtypes = [C.c_void_p, _i64, _i64, _d, _d, _d]
        L.orc_get_expansions.argtypes = [C.c_void_p, C.c_int, _d, _d]
        L.orc_quad_export.argtypes = [C.c_void_p, _d, _d, _d, _d]
        L.orc_compute_alpha.argtypes = [C.c_void_p, _d]
        L.orc_vmult.argtypes = [C.c_void_p, _d, _d]
        L.orc_compute_rhs.argtypes = [C.c_void_p, _d, _d]
        L.orc_solve_system.argtypes = [C.c_void_p, _d, _i64, _i64, _d, _d, _u8, _d, _d, C.c_double,
                                       C.c_int, C.c_int, C.POINTER(C.c_int), _d]
        L.orc_dense_assemble.argtypes = [C.c_void_p, C.c_int, _d, _d]
        L.orc_legendre.argtypes = [C.c_int, C.c_double, _d, _d]
        L.orc_me_p2m.argtypes = [C.c_int, _d, C.c_int, _d, _d, _d]
        L.orc_me_p2m_dipole.argtypes = [C.c_int, _d, C.c_int, _d, _d, _d, _d]
        L.orc_me_m2m.argtypes = [C.c_int, _d, _d, _d, _d]
        L.orc_me_eval.restype = C.c_double
        L.orc_me_eval.argtypes = [C.c_int, _d, _d, _d]
        L.orc_le_m2l.argtypes = [C.c_int, _d, _d, _d, _d]
        L.orc_le_l2l.argtypes = [C.c_int, _d, _d, _d, _d]
        L.orc_le_eval.restype = C.c_double
        L.orc_le_eval.argtypes = [C.c_int, _d, _d, _d]
        L.orc_gauss01.argtypes = [C.c_int, _d, _d]
        L.orc_telles01.argtypes = [C.c_int, C.c_double, _d, _d]
        L.orc_duffy.argtypes = [C.c_int, C.c_double, C.c_double, _d, _d]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t=_d):
    return a.ctypes.data_as(t) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def kt(p):
    return (p + 1) * (p + 2) // 2


# ---- expansion algebra -----------------------------------------------------
def legendre(nmax, x, deriv=False):
    P = np.zeros((nmax + 1) * (nmax + 1))
    dP = np.zeros_like(P) if deriv else None
    lib().orc_legendre(nmax, float(x), _p(P), _p(dP))
    P = P.reshape(nmax + 1, nmax + 1)
    return (P, dP.reshape(nmax + 1, nmax + 1)) if deriv else P


def _cplx(a):
    return a[0::2] + 1j * a[1::2]


def _real(c):
    out = np.empty(2 * len(c))
    out[0::2] = np.real(c)
    out[1::2] = np.imag(c)
    return out


def p2m(p, center, q, y):
    out = np.zeros(2 * kt(p))
    q, y = _f64(q), _f64(y)
    lib().orc_me_p2m(p, _p(_f64(center)), len(q), _p(q), _p(y), _p(out))
    return _cplx(out)


def p2m_dipole(p, center, q, y, nrm):
    out = np.zeros(2 * kt(p))
    q, y, nrm = _f64(q), _f64(y), _f64(nrm)
    lib().orc_me_p2m_dipole(p, _p(_f64(center)), len(q), _p(q), _p(y), _p(nrm), _p(out))
    return _cplx(out)


def m2m(p, child_center, child, parent_center):
    out = np.zeros(2 * kt(p))
    lib().orc_me_m2m(p, _p(_f64(child_center)), _p(_real(child)), _p(_f64(parent_center)), _p(out))
    return _cplx(out)


def me_eval(p, center, coef, x):
    return lib().orc_me_eval(p, _p(_f64(center)), _p(_real(coef)), _p(_f64(x)))


def m2l(p, me_center, me, le_center):
    out = np.zeros(2 * kt(p))
    lib().orc_le_m2l(p, _p(_f64(me_center)), _p(_real(me)), _p(_f64(le_center)), _p(out))
    return _cplx(out)


def l2l(p, parent_center, parent, child_center):
    out = np.zeros(2 * kt(p))
    lib().orc_le_l2l(p, _p(_f64(parent_center)), _p(_real(parent)), _p(_f64(child_center)), _p(out))
    return _cplx(out)


def le_eval(p, center, coef, x):
    return lib().orc_le_eval(p, _p(_f64(center)), _p(_real(coef)), _p(_f64(x)))


def gauss01(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().orc_gauss01(n, _p(x), _p(w))
    return x, w


def telles01(n, x0):
    x, w = np.zeros(n), np.zeros(n)
    k = lib().orc_telles01(n, float(x0), _p(x), _p(w))
    return x[:k], w[:k]


def duffy(n, sx, sy):
    pts, w = np.zeros(2 * 8 * n * n), np.zeros(8 * n * n)
    k = lib().orc_duffy(n, float(sx), float(sy), _p(pts), _p(w))
    return pts[:2 * k].reshape(k, 2), w[:k]


# ---- the BEMFMA / BEMProblem restatement -------------------------------------
class OracleFMM:
    """CPU restatement of BEMFMA<3> + the FMA branch of BEMProblem<3>."""

    def __init__(self, problem=None, *, p=6, levels=10, max_nodes=20, quad_order=4, sing_order=5, mesh=None,
                 dirichlet=None, sing_rule=0):
        if problem is not None:
            mesh, dirichlet = problem.mesh, problem.dirichlet
            p, levels, max_nodes = problem.p, problem.levels, problem.max_nodes_per_leaf
            quad_order, sing_order = problem.quad_order, problem.sing_quad_order
            sing_rule = problem.sing_rule
        self.p = p
        self.levels = levels
        self.mesh = mesh
        self.h = lib().orc_create(p, levels, max_nodes, quad_order, sing_order)
        self.sing_rule = sing_rule
        if lib().orc_set_sing_rule(self.h, int(sing_rule)) != 0:
            raise ValueError(f"singular rule {sing_rule}")
        m = mesh
        self._keep = [_f64(m.support_xyz), _i(m.cell_dofs), _i(m.dbl_ptr), _i(m.dbl_idx),
                      np.ascontiguousarray(dirichlet, dtype=np.uint8)]
        xyz, cd, dp, di, dm = self._keep
        rc = lib().orc_init(self.h, m.n_dofs, _p(xyz), m.n_cells, m.dpc, _p(cd, _i64), None, _p(dp, _i64),
                            _p(di, _i64), _p(dm, _u8))
        if rc != 0:
            raise RuntimeError("oracle init failed")
        self.n = m.n_dofs
        z0 = getattr(problem, "mirror_z0", None) if problem is not None else None
        if z0 is not None:
            lib().orc_set_double_body(self.h, C.c_double(z0))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def setup(self):
        L = lib()
        L.orc_generate_octree_blocking(self.h)
        L.orc_direct_integrals(self.h)
        L.orc_multipole_integrals(self.h)
        return self

    def generate_octree_blocking(self):
        lib().orc_generate_octree_blocking(self.h)

    def direct_integrals(self):
        lib().orc_direct_integrals(self.h)

    def multipole_integrals(self):
        lib().orc_multipole_integrals(self.h)

    def generate_multipole_expansions(self, phi, dphi):
        lib().orc_generate_multipole_expansions(self.h, _p(_f64(phi)), _p(_f64(dphi)))

    def multipole_matr_vect_products(self, phi, dphi):
        N, D = np.zeros(self.n), np.zeros(self.n)
        lib().orc_multipole_matr_vect_products(self.h, _p(_f64(phi)), _p(_f64(dphi)), _p(N), _p(D))
        return N, D

    def matvec(self, phi, dphi):
        self.generate_multipole_expansions(phi, dphi)
        return self.multipole_matr_vect_products(phi, dphi)

    def tree(self):
        cnt = np.zeros(8, dtype=np.int64)
        lib().orc_tree_counts(self.h, _p(cnt, _i64))
        nb, L, nleaf, ns, nil, nnil, nn, _ = (int(x) for x in cnt)
        t = dict(level=np.zeros(nb, np.int32), parent=np.zeros(nb, np.int64), pmin=np.zeros(3 * nb),
                 delta=np.zeros(nb), startL=np.zeros(L + 1, np.int64), endL=np.zeros(L + 1, np.int64),
                 childless=np.zeros(nleaf, np.int64), sub_ptr=np.zeros(nb + 1, np.int64),
                 il_ptr=np.zeros(ns + 1, np.int64), il_idx=np.zeros(nil, np.int64),
                 nil_ptr=np.zeros(ns + 1, np.int64), nil_idx=np.zeros(nnil, np.int64),
                 node_ptr=np.zeros(nb + 1, np.int64), node_idx=np.zeros(nn, np.int64))
        lib().orc_tree_export(self.h, _p(t["level"], _i32), *[_p(t[k], _i64) if t[k].dtype == np.int64 else _p(t[k])
                                                              for k in ("parent", "pmin", "delta", "startL", "endL",
                                                                        "childless", "sub_ptr", "il_ptr", "il_idx",
                                                                        "nil_ptr", "nil_idx", "node_ptr",
                                                                        "node_idx")])
        t["pmin"] = t["pmin"].reshape(nb, 3)
        t["num_blocks"] = nb
        return t

    def csr(self):
        nnz = lib().orc_csr_nnz(self.h)
        rp, col = np.zeros(self.n + 1, np.int64), np.zeros(nnz, np.int64)
        vN, vD, vP = np.zeros(nnz), np.zeros(nnz), np.zeros(nnz)
        lib().orc_csr_export(self.h, _p(rp, _i64), _p(col, _i64), _p(vN), _p(vD), _p(vP))
        return rp, col, vN, vD, vP

    def expansions(self, which="me"):
        nb = self.tree()["num_blocks"]
        k = kt(self.p)
        a, b = np.zeros(2 * k * nb), np.zeros(2 * k * nb)
        lib().orc_get_expansions(self.h, 0 if which == "me" else 1, _p(a), _p(b))
        return _cplx(a).reshape(nb, k), _cplx(b).reshape(nb, k)

    def compute_alpha(self):
        a = np.zeros(self.n)
        lib().orc_compute_alpha(self.h, _p(a))
        return a

    def vmult(self, src):
        dst = np.zeros(self.n)
        lib().orc_vmult(self.h, _p(_f64(src)), _p(dst))
        return dst

    def compute_rhs(self, g):
        dst = np.zeros(self.n)
        lib().orc_compute_rhs(self.h, _p(_f64(g)), _p(dst))
        return dst

    def solve_system(self, tmp_rhs, cons=None, tol=1e-8, restart=100, maxit=1000):
        phi, dphi = np.zeros(self.n), np.zeros(self.n)
        it, rr = C.c_int(0), C.c_double(0)
        if cons is None:
            args = [None, None, None, None, None]
        else:
            ptr, idx, val, inh, isc = cons
            self._ckeep = [_i(ptr), _i(idx), _f64(val), _f64(inh), np.ascontiguousarray(isc, np.uint8)]
            a, b, c, d, e = self._ckeep
            args = [_p(a, _i64), _p(b, _i64), _p(c), _p(d), _p(e, _u8)]
        lib().orc_solve_system(self.h, _p(_f64(tmp_rhs)), *args, _p(phi), _p(dphi), tol, restart, maxit,
                               C.byref(it), C.byref(rr))
        return phi, dphi, it.value, rr.value

    def dense(self, rule=None):
        """Dense collocation matrices (bem_problem.cc:477-950); rule None
        takes this oracle's near-field rule (0 Telles, 1 Duffy split)."""
        rule = self.sing_rule if rule is None else rule
        Nm, Dm = np.zeros((self.n, self.n)), np.zeros((self.n, self.n))
        lib().orc_dense_assemble(self.h, rule, _p(Nm), _p(Dm))
        return Nm, Dm
