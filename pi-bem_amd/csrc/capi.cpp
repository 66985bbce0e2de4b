// C ABI of libpbfmm (include/pbfmm.h): argument checks, error reporting, host
// copies, and the call order of BEMFMA / BEMProblem.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <atomic>
#include <exception>
#include <string>
#include <thread>
#include <vector>

#include "engine.h"

using namespace pbfmm;

struct pbfmm_handle {
  pbfmm_params prm;
  Engine eng;
  std::string err;
  int stage = 0;  // 1 init_fma, 2 octree, 3 direct, 4 multipole
  bool alpha_ready = false;
  bool device_ready = false;
  bool batch = false;  // a forest of parts (pbfmm_batch_create)
};

namespace {
int fail(pbfmm_handle *h, int code, const std::string &msg) {
  if (h) h->err = msg;
  return code;
}
// Host-only steps (init_fma, generate_octree_blocking, tree export) run
// without touching the GPU; every other call selects the handle's device and
// initialises it on first use.
template <typename F>
int guarded(pbfmm_handle *h, F &&f, bool device = true) {
  if (!h) return PBFMM_ERR_ARG;
  try {
    if (device) {
      hipError_t e = hipSetDevice(h->prm.device);
      if (e != hipSuccess) return fail(h, PBFMM_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
      if (!h->device_ready) {
        h->eng.init_device(h->prm.device, h->prm.stream);
        h->device_ready = true;
      }
    }
    int rc = f();
    if (rc == PBFMM_OK) h->err.clear();
    return rc;
  } catch (const HipError &ex) {
    return fail(h, PBFMM_ERR_HIP, ex.what());
  } catch (const std::bad_alloc &) {
    return fail(h, PBFMM_ERR_ALLOC, "host allocation failed");
  } catch (const std::exception &ex) {
    return fail(h, PBFMM_ERR_ARG, ex.what());
  }
}
void h2d(Engine &e, double *dst, const double *src, int64_t n) {
  if (hipMemcpyAsync(dst, src, n * 8, hipMemcpyHostToDevice, e.stream) != hipSuccess)
    throw HipError("host to device copy failed");
}
void d2h(Engine &e, double *dst, const double *src, int64_t n) {
  if (hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToHost, e.stream) != hipSuccess)
    throw HipError("device to host copy failed");
  e.sync();
}
int need(pbfmm_handle *h, int stage, const char *what) {
  if (h->stage < stage) return fail(h, PBFMM_ERR_STATE, std::string(what) + " called before its set-up steps");
  return PBFMM_OK;
}
// per-mesh set-up steps and tree exports do not apply to a forest
int not_batch(pbfmm_handle *h, const char *what) {
  if (h->batch) return fail(h, PBFMM_ERR_UNSUPPORTED, std::string(what) + " is not available on a batch handle");
  return PBFMM_OK;
}
void ensure_alpha(pbfmm_handle *h) {
  if (!h->alpha_ready) {
    h->eng.compute_alpha();
    h->alpha_ready = true;
    h->eng.have_prec = false;
  }
}
// device staging buffer of at least `bytes`, grown on demand
void *stage(DevBuf &b, size_t bytes) {
  if (!b.p || b.bytes < bytes) {
    b.release();
    bytes = std::max<size_t>(bytes, 16);
    if (hipMalloc(&b.p, bytes) != hipSuccess) {
      b.p = nullptr;
      throw HipError("device allocation of a post-solve buffer failed");
    }
    b.bytes = bytes;
  }
  return b.p;
}
void h2d_bytes(Engine &e, void *dst, const void *src, size_t bytes) {
  if (bytes && hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, e.stream) != hipSuccess)
    throw HipError("host to device copy failed");
}
// the post-solve data (mass matrix and quadrature tables) on first use
int ps_ready(pbfmm_handle *h, const char *what) {
  if (int rc = need(h, 1, what)) return rc;
  h->eng.ps_setup(h->prm.quad_order);
  return PBFMM_OK;
}
// masked body integrals; mask NULL selects every real cell
void ps_sums(pbfmm_handle *h, const uint8_t *mask, const double *pressure, const double *cog, double out[13]) {
  Engine &e = h->eng;
  const int64_t nc = e.ps.ncells;
  uint8_t *dm = (uint8_t *)stage(e.ps.mask, (size_t)nc);
  if (mask) {
    h2d_bytes(e, dm, mask, (size_t)nc);
  } else if (hipMemsetAsync(dm, 1, (size_t)nc, e.stream) != hipSuccess) {
    throw HipError("mask fill failed");
  }
  double *dp = (double *)stage(e.ps.in1, (size_t)e.n * 8);
  if (pressure)
    h2d(e, dp, pressure, e.n);
  else if (hipMemsetAsync(dp, 0, (size_t)e.n * 8, e.stream) != hipSuccess)
    throw HipError("pressure fill failed");
  const double zero[3] = {0, 0, 0};
  e.ps_integrals(dm, dp, cog ? cog : zero, out);
}
// FMA_preconditioner as called from solve_system (bem_problem.cc:1302-1304)
void ensure_prec(pbfmm_handle *h) {
  ensure_alpha(h);
  if (h->prm.preconditioner == 1 && !h->eng.have_prec) h->eng.build_preconditioner();
}
}  // namespace

extern "C" {

void pbfmm_default_params(pbfmm_params *p) {
  if (!p) return;
  p->num_octree_levels = 10;
  p->max_nodes_per_leaf = 20;
  p->trunc_order = 6;
  p->quad_order = 4;
  p->sing_quad_order = 5;
  p->precision = PBFMM_FP64;
  p->device = 0;
  p->stream = nullptr;
  p->preconditioner = 1;
  p->sing_rule = PBFMM_SING_TELLES;
}

int pbfmm_create(const pbfmm_params *p, pbfmm_handle **out) {
  if (!p || !out) return PBFMM_ERR_ARG;
  *out = nullptr;
  if (p->trunc_order < 1 || p->trunc_order > 10 || p->num_octree_levels < 1 || p->num_octree_levels > 20 ||
      p->max_nodes_per_leaf < 1 || p->quad_order < 1 || p->quad_order > 8 || p->sing_quad_order < 1 ||
      p->sing_quad_order > 16 || (p->precision != PBFMM_FP64 && p->precision != PBFMM_MIXED) ||
      p->preconditioner < 0 || p->preconditioner > 1 ||
      (p->sing_rule != PBFMM_SING_TELLES && p->sing_rule != PBFMM_SING_DUFFY))
    return PBFMM_ERR_ARG;
  pbfmm_handle *h = nullptr;
  try {
    h = new pbfmm_handle();
    h->prm = *p;
    h->eng.p = p->trunc_order;
    h->eng.kt = kt_of(p->trunc_order);
    h->eng.kp = kp_of(p->trunc_order);
    h->eng.L = p->num_octree_levels;
    h->eng.sing_order = p->sing_quad_order;
    h->eng.sing_rule = p->sing_rule;
  } catch (...) {
    delete h;
    return PBFMM_ERR_HIP;
  }
  *out = h;
  return PBFMM_OK;
}

int pbfmm_destroy(pbfmm_handle *h) {
  if (!h) return PBFMM_ERR_ARG;
  if (h->device_ready) (void)hipSetDevice(h->prm.device);
  delete h;
  return PBFMM_OK;
}

const char *pbfmm_last_error(const pbfmm_handle *h) { return h ? h->err.c_str() : "null handle"; }

int pbfmm_init_fma(pbfmm_handle *h, int64_t n_dofs, const double *xyz, int64_t n_cells, int32_t dpc,
                   const int64_t *cell_dofs, const double *cell_geom, const int64_t *dbl_ptr,
                   const int64_t *dbl_idx, const uint8_t *dirichlet_mask) {
  return guarded(h, [&]() -> int {
    if (int rc = not_batch(h, "init_fma")) return rc;
    if (n_dofs <= 0 || n_cells <= 0 || !xyz || !cell_dofs || !dirichlet_mask || (dpc != 4 && dpc != 9))
      return fail(h, PBFMM_ERR_ARG, "init_fma: need n_dofs > 0, n_cells > 0, dpc 4 or 9 and non-null arrays");
    if (n_dofs >= (int64_t(1) << 30) || n_cells * dpc >= (int64_t(1) << 31))
      return fail(h, PBFMM_ERR_ARG, "init_fma: sizes exceed the 32-bit index range of this build");
    if ((dbl_ptr == nullptr) != (dbl_idx == nullptr))
      return fail(h, PBFMM_ERR_ARG, "init_fma: dbl_ptr and dbl_idx must both be given or both be null");
    Engine &e = h->eng;
    HostMesh &m = e.mesh;
    m = HostMesh();
    m.n = n_dofs;
    m.ncells = n_cells;
    m.ncells_real = n_cells;
    m.dpc = dpc;
    m.xyz.assign(xyz, xyz + 3 * n_dofs);
    m.cdofs.resize(n_cells * dpc);
    for (int64_t i = 0; i < n_cells * dpc; ++i) {
      if (cell_dofs[i] < 0 || cell_dofs[i] >= n_dofs)
        return fail(h, PBFMM_ERR_ARG, "init_fma: cell_dofs entry out of range");
      m.cdofs[i] = (int32_t)cell_dofs[i];
    }
    m.geom.resize(n_cells * dpc * 3);
    for (int64_t c = 0; c < n_cells; ++c)
      for (int j = 0; j < dpc; ++j)
        for (int k = 0; k < 3; ++k)
          m.geom[(c * dpc + j) * 3 + k] =
              cell_geom ? cell_geom[(c * dpc + j) * 3 + k] : xyz[3 * (int64_t)m.cdofs[c * dpc + j] + k];
    // double_nodes_set: sorted, unique, holding the dof itself (bem_problem.cc:413-454)
    m.dbl_ptr.assign(n_dofs + 1, 0);
    std::vector<int32_t> tmp;
    for (int64_t i = 0; i < n_dofs; ++i) {
      tmp.clear();
      tmp.push_back((int32_t)i);
      if (dbl_ptr) {
        if (dbl_ptr[i + 1] < dbl_ptr[i]) return fail(h, PBFMM_ERR_ARG, "init_fma: dbl_ptr not monotone");
        for (int64_t k = dbl_ptr[i]; k < dbl_ptr[i + 1]; ++k) {
          if (dbl_idx[k] < 0 || dbl_idx[k] >= n_dofs) return fail(h, PBFMM_ERR_ARG, "init_fma: dbl_idx out of range");
          tmp.push_back((int32_t)dbl_idx[k]);
        }
      }
      std::sort(tmp.begin(), tmp.end());
      tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
      m.dbl_idx.insert(m.dbl_idx.end(), tmp.begin(), tmp.end());
      m.dbl_ptr[i + 1] = (int32_t)m.dbl_idx.size();
    }
    m.dir.assign(dirichlet_mask, dirichlet_mask + n_dofs);
    e.have_dirichlet = false;
    for (auto v : m.dir)
      if (v) e.have_dirichlet = true;
    // dof -> cells
    m.d2c_ptr.assign(n_dofs + 1, 0);
    for (int64_t i = 0; i < n_cells * dpc; ++i) m.d2c_ptr[m.cdofs[i] + 1]++;
    for (int64_t i = 0; i < n_dofs; ++i) m.d2c_ptr[i + 1] += m.d2c_ptr[i];
    m.d2c_idx.resize(n_cells * dpc);
    std::vector<int32_t> fill(m.d2c_ptr.begin(), m.d2c_ptr.end() - 1);
    for (int64_t c = 0; c < n_cells; ++c)
      for (int j = 0; j < dpc; ++j) m.d2c_idx[fill[m.cdofs[c * dpc + j]]++] = (int32_t)c;
    e.n = n_dofs;
    e.dpc = dpc;
    h->stage = 1;
    h->alpha_ready = false;
    e.have_constraints = false;
    return PBFMM_OK;
  }, false);
}

int pbfmm_set_double_body(pbfmm_handle *h, int32_t enable, double z0) {
  return guarded(h, [&]() -> int {
    if (int rc = not_batch(h, "set_double_body")) return rc;
    if (int rc = need(h, 1, "set_double_body")) return rc;
    if (h->stage > 1) return fail(h, PBFMM_ERR_STATE, "set_double_body: call it before generate_octree_blocking");
    HostMesh &m = h->eng.mesh;
    m.cdofs.resize(m.ncells_real * m.dpc);
    m.geom.resize(m.ncells_real * m.dpc * 3);
    m.ncells = m.ncells_real;
    m.mirror_z0 = 0.0;
    if (enable) {
      if (!std::isfinite(z0)) return fail(h, PBFMM_ERR_ARG, "set_double_body: z0 must be finite");
      add_mirror_cells(m, z0);
    }
    return PBFMM_OK;
  }, false);
}

int pbfmm_generate_octree_blocking(pbfmm_handle *h) {
  return guarded(h, [&]() -> int {
    if (int rc = not_batch(h, "generate_octree_blocking")) return rc;
    if (int rc = need(h, 1, "generate_octree_blocking")) return rc;
    Engine &e = h->eng;
    std::vector<double> gx, gw;
    gauss_legendre01(h->prm.quad_order, gx, gw);
    Quadrature2D rule = tensor_rule(gx, gw, gx, gw);
    compute_quad_data(e.mesh, rule, e.qd);
    e.nq = e.qd.nq;
    build_tree(e.mesh, e.qd, h->prm.num_octree_levels, h->prm.max_nodes_per_leaf, e.tree);
    h->stage = 2;
    h->alpha_ready = false;
    return PBFMM_OK;
  }, false);
}

// The census of compute_m2l_flags (bem_fma.cc:1846-1956): per level, the
// nonIntList entries of every sublevel summed over the blocks that hold
// nodes (dofs_filled_blocks), which the reference splits across MPI ranks;
// one process here, so no split.
int pbfmm_compute_m2l_flags(pbfmm_handle *h, int64_t *counts) {
  return guarded(h, [&]() -> int {
    if (int rc = not_batch(h, "compute_m2l_flags")) return rc;
    if (int rc = need(h, 2, "compute_m2l_flags")) return rc;
    if (!counts) return fail(h, PBFMM_ERR_ARG, "compute_m2l_flags: null output");
    const Tree &t = h->eng.tree;
    for (int l = 0; l <= t.L; ++l) {
      int64_t c = 0;
      if (l >= 1)
        for (int32_t b : t.dofs_filled[l])
          for (int s = 0; s < t.nsub[b]; ++s) c += (int64_t)t.nil[t.sub_beg[b] + s].size();
      counts[l] = c;
    }
    return PBFMM_OK;
  }, false);
}

int pbfmm_direct_integrals(pbfmm_handle *h) {
  return guarded(h, [&]() -> int {
    if (int rc = not_batch(h, "direct_integrals")) return rc;
    if (int rc = need(h, 2, "direct_integrals")) return rc;
    Engine &e = h->eng;
    build_near_plan(e.mesh, e.qd, e.tree, e.near);
    if (e.near.nnz >= (int64_t(1) << 40)) return fail(h, PBFMM_ERR_ARG, "direct_integrals: near field too large");
    e.upload_mesh_and_tree();
    e.upload_quad(h->prm.precision == PBFMM_MIXED);
    e.upload_near_plan();
    e.assemble_near(h->prm.precision == PBFMM_MIXED);
    h->stage = 3;
    h->alpha_ready = false;
    return PBFMM_OK;
  });
}

int pbfmm_multipole_integrals(pbfmm_handle *h) {
  return guarded(h, [&]() -> int {
    if (int rc = not_batch(h, "multipole_integrals")) return rc;
    if (int rc = need(h, 3, "multipole_integrals")) return rc;
    Engine &e = h->eng;
    build_far_plan(e.tree, e.p, e.far);
    build_p2m_plan(e.mesh, e.qd, e.tree, e.far);
    validate_plans(e);
    e.upload_far(h->prm.precision == PBFMM_MIXED);
    e.build_unit_multipoles(h->prm.precision == PBFMM_MIXED);
    h->stage = 4;
    return PBFMM_OK;
  });
}

int pbfmm_generate_multipole_expansions(pbfmm_handle *h, const double *phi, const double *dphi) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 4, "generate_multipole_expansions")) return rc;
    if (!phi || !dphi) return fail(h, PBFMM_ERR_ARG, "generate_multipole_expansions: null vector");
    Engine &e = h->eng;
    h2d(e, e.d_in.as<double>(), phi, e.n);
    h2d(e, e.d_out.as<double>(), dphi, e.n);
    e.generate<double, 2>(e.d_in.as<double>(), e.d_out.as<double>());
    e.sync();
    return PBFMM_OK;
  });
}

int pbfmm_multipole_matr_vect_products(pbfmm_handle *h, const double *phi, const double *dphi, double *outN,
                                       double *outD) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 4, "multipole_matr_vect_products")) return rc;
    if (!phi || !dphi || !outN || !outD) return fail(h, PBFMM_ERR_ARG, "multipole_matr_vect_products: null vector");
    Engine &e = h->eng;
    h2d(e, e.d_in.as<double>(), phi, e.n);
    h2d(e, e.d_out.as<double>(), dphi, e.n);
    e.products<double, 2>(e.d_in.as<double>(), e.d_out.as<double>(), e.d_outN.as<double>(), e.d_outD.as<double>());
    d2h(e, outN, e.d_outN.as<double>(), e.n);
    d2h(e, outD, e.d_outD.as<double>(), e.n);
    return PBFMM_OK;
  });
}

int pbfmm_compute_alpha(pbfmm_handle *h, double *alpha) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 4, "compute_alpha")) return rc;
    h->eng.compute_alpha();
    h->alpha_ready = true;
    if (alpha) d2h(h->eng, alpha, h->eng.d_alpha.as<double>(), h->eng.n);
    return PBFMM_OK;
  });
}

// Installs constraint rows (CSR over the DoFs, see pbfmm_set_constraints) in
// the handle: host copies for the preconditioner, device copies for vmult.
static int install_constraints(pbfmm_handle *h, const int64_t *c_ptr, const int64_t *c_idx, const double *c_val,
                               const double *c_inhom, const uint8_t *is_c) {
  {
    Engine &e = h->eng;
    const int64_t n = e.n;
    ++e.constraint_version;  // a captured operator holds the old rows
    std::vector<int32_t> ptr(n + 1), idx;
    std::vector<double> val, inh(n, 0.0);
    std::vector<uint8_t> isc(is_c, is_c + n);
    for (int64_t i = 0; i < n; ++i) {
      if (c_ptr[i + 1] < c_ptr[i]) return fail(h, PBFMM_ERR_ARG, "set_constraints: c_ptr not monotone");
      ptr[i + 1] = (int32_t)c_ptr[i + 1];
      if (c_inhom) inh[i] = c_inhom[i];
    }
    for (int64_t k = 0; k < c_ptr[n]; ++k) {
      if (c_idx[k] < 0 || c_idx[k] >= n) return fail(h, PBFMM_ERR_ARG, "set_constraints: c_idx out of range");
      idx.push_back((int32_t)c_idx[k]);
      val.push_back(c_val ? c_val[k] : 1.0);
    }
    auto up = [&](DevBuf &b, const void *src, size_t bytes) {
      b.release();
      if (hipMalloc(&b.p, std::max<size_t>(bytes, 16)) != hipSuccess) throw HipError("hipMalloc failed");
      b.bytes = bytes;
      if (bytes && hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice) != hipSuccess) throw HipError("hipMemcpy failed");
    };
    up(e.d_c_ptr, ptr.data(), ptr.size() * 4);
    up(e.d_c_idx, idx.data(), idx.size() * 4);
    up(e.d_c_val, val.data(), val.size() * 8);
    up(e.d_c_inh, inh.data(), inh.size() * 8);
    up(e.d_c_is, isc.data(), isc.size());
    e.hc_ptr = ptr;
    e.hc_idx = idx;
    e.hc_val = val;
    e.hc_is = isc;
    e.hc_inh = inh;
    e.have_constraints = true;
    e.have_prec = false;
    return PBFMM_OK;
  }
}

int pbfmm_set_constraints(pbfmm_handle *h, const int64_t *c_ptr, const int64_t *c_idx, const double *c_val,
                          const double *c_inhom, const uint8_t *is_c) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 1, "set_constraints")) return rc;
    Engine &e = h->eng;
    if (!c_ptr || !is_c) {
      e.have_constraints = false;
      e.have_prec = false;
      return PBFMM_OK;
    }
    return install_constraints(h, c_ptr, c_idx, c_val, c_inhom, is_c);
  });
}

// BEMProblem::compute_constraints (bem_problem.cc:1376-1653) for one process
// and meshes without hanging nodes: the double-node rows from the projected
// normals and the projected surface gradients of tmp_rhs.
int pbfmm_compute_constraints(pbfmm_handle *h, const double *tmp_rhs, int32_t continuous_gradient) {
  return guarded(h, [&]() -> int {
    if (int rc = not_batch(h, "compute_constraints")) return rc;
    if (int rc = ps_ready(h, "compute_constraints")) return rc;
    if (!tmp_rhs) return fail(h, PBFMM_ERR_ARG, "compute_constraints: null tmp_rhs");
    Engine &e = h->eng;
    const int64_t n = e.n;
    // vector_normals_solution and vector_surface_gradients_solution
    // (compute_normals, compute_surface_gradients; bem_problem.cc:1385),
    // component-major
    std::vector<double> nrm((size_t)3 * n), sg((size_t)3 * n);
    double *dout = (double *)stage(e.ps.out, (size_t)3 * n * 8);
    e.ps_project(0, nullptr, nullptr, dout);
    d2h(e, nrm.data(), dout, 3 * n);
    double *d1 = (double *)stage(e.ps.in1, (size_t)n * 8);
    h2d(e, d1, tmp_rhs, n);
    e.ps_project(2, d1, nullptr, dout);
    d2h(e, sg.data(), dout, 3 * n);
    auto dot = [&](const std::vector<double> &a, int64_t i, const std::vector<double> &b, int64_t j) {
      return a[i] * b[j] + a[n + i] * b[n + j] + a[2 * n + i] * b[2 * n + j];
    };
    const HostMesh &m = e.mesh;
    std::vector<std::vector<std::pair<int64_t, double>>> ent(n);
    std::vector<double> inh(n, 0.0);
    std::vector<uint8_t> isc(n, 0);
    for (int64_t i = 0; i < n; ++i) {
      const int32_t *d0 = m.dbl_idx.data() + m.dbl_ptr[i], *d1e = m.dbl_idx.data() + m.dbl_ptr[i + 1];
      // the first Dirichlet DoF of the set, else its first DoF (ascending)
      int64_t first = *d0;
      for (const int32_t *d = d0; d < d1e; ++d)
        if (m.dir[*d]) {
          first = *d;
          break;
        }
      if (i != first) continue;
      for (const int32_t *d = d0; d < d1e; ++d) {
        const int64_t it = *d;
        if (it == i) continue;
        if (!m.dir[i]) {  // Neumann set: phi(it) = phi(i)
          isc[it] = 1;
          ent[it].push_back({i, 1.0});
        } else if (!m.dir[it]) {  // the Neumann double takes the Dirichlet value
          isc[it] = 1;
          inh[it] = tmp_rhs[i];
        } else {
          // Dirichlet-Dirichlet.  The reference's parallel-normal test
          // divides the dot product by itself (bem_problem.cc:1490), so it
          // never fires and parallel normals divide by 1 - (n.n)^2 = 0; the
          // intended test is made here: normals within 1e-4 of parallel
          // give normal derivatives equal up to the sign of n_i.n_it.
          const double nn = dot(nrm, i, nrm, it);
          const double ni = std::sqrt(dot(nrm, i, nrm, i)), nt = std::sqrt(dot(nrm, it, nrm, it));
          const double cosang = (ni > 0 && nt > 0) ? nn / (ni * nt) : 1.0;
          if (1.0 - std::fabs(cosang) < 1e-4) {
            isc[it] = 1;
            ent[it].push_back({i, cosang > 0 ? 1.0 : -1.0});
          } else if (continuous_gradient) {
            // both normal derivatives from the surface gradients
            // (bem_problem.cc:1499-1557)
            const double s_it_n_i = dot(sg, it, nrm, i), s_i_n_it = dot(sg, i, nrm, it);
            const double den = 1.0 - nn * nn;
            isc[i] = 1;
            inh[i] = (s_it_n_i + s_i_n_it * nn) / den;
            isc[it] = 1;
            inh[it] = (s_i_n_it + s_it_n_i * nn) / den;
          }
        }
      }
    }
    std::vector<int64_t> ptr(n + 1, 0), idx;
    std::vector<double> val;
    for (int64_t i = 0; i < n; ++i) {
      for (auto &pv : ent[i]) {
        idx.push_back(pv.first);
        val.push_back(pv.second);
      }
      ptr[i + 1] = (int64_t)idx.size();
    }
    return install_constraints(h, ptr.data(), idx.data(), val.data(), inh.data(), isc.data());
  });
}

int pbfmm_get_constraints(pbfmm_handle *h, int64_t *nnz, int64_t *c_ptr, int64_t *c_idx, double *c_val,
                          double *c_inhom, uint8_t *is_c) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 1, "get_constraints")) return rc;
    if (!nnz) return fail(h, PBFMM_ERR_ARG, "get_constraints: null nnz");
    Engine &e = h->eng;
    const int64_t n = e.n;
    *nnz = e.have_constraints ? (int64_t)e.hc_idx.size() : 0;
    if (!c_ptr && !c_idx && !c_val && !c_inhom && !is_c) return PBFMM_OK;
    if (!c_ptr || !c_inhom || !is_c || (*nnz > 0 && (!c_idx || !c_val)))
      return fail(h, PBFMM_ERR_ARG, "get_constraints: null output");
    for (int64_t i = 0; i <= n; ++i) c_ptr[i] = e.have_constraints ? e.hc_ptr[i] : 0;
    for (int64_t k = 0; k < *nnz; ++k) {
      c_idx[k] = e.hc_idx[k];
      c_val[k] = e.hc_val[k];
    }
    for (int64_t i = 0; i < n; ++i) {
      c_inhom[i] = e.have_constraints ? e.hc_inh[i] : 0.0;
      is_c[i] = e.have_constraints ? e.hc_is[i] : 0;
    }
    return PBFMM_OK;
  });
}

int pbfmm_fma_preconditioner(pbfmm_handle *h, const double *alpha) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 4, "fma_preconditioner")) return rc;
    Engine &e = h->eng;
    if (alpha) {
      h2d(e, e.d_alpha.as<double>(), alpha, e.n);
      h->alpha_ready = true;
    }
    ensure_alpha(h);
    e.build_preconditioner();
    return PBFMM_OK;
  });
}

int pbfmm_set_preconditioner_rows(pbfmm_handle *h, int32_t rows) {
  return guarded(h, [&]() -> int {
    if (rows < 8 || rows > 128) return fail(h, PBFMM_ERR_ARG, "set_preconditioner_rows: rows must be 8..128");
    Engine &e = h->eng;
    e.prec_rows = rows;
    if (e.have_prec) e.build_preconditioner();
    return PBFMM_OK;
  });
}

int pbfmm_vmult(pbfmm_handle *h, const double *src, double *dst) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 4, "vmult")) return rc;
    if (!src || !dst) return fail(h, PBFMM_ERR_ARG, "vmult: null vector");
    ensure_alpha(h);
    Engine &e = h->eng;
    h2d(e, e.d_in.as<double>(), src, e.n);
    e.vmult(e.d_in.as<double>(), e.d_out.as<double>(), PBFMM_FP64);
    d2h(e, dst, e.d_out.as<double>(), e.n);
    return PBFMM_OK;
  });
}

int pbfmm_compute_rhs(pbfmm_handle *h, const double *tmp_rhs, double *rhs) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 4, "compute_rhs")) return rc;
    if (!tmp_rhs || !rhs) return fail(h, PBFMM_ERR_ARG, "compute_rhs: null vector");
    ensure_alpha(h);
    Engine &e = h->eng;
    h2d(e, e.d_in.as<double>(), tmp_rhs, e.n);
    e.compute_rhs(e.d_in.as<double>(), e.d_out.as<double>());
    d2h(e, rhs, e.d_out.as<double>(), e.n);
    return PBFMM_OK;
  });
}

int pbfmm_solve_system(pbfmm_handle *h, const double *tmp_rhs, double *phi, double *dphi, double rel_tol,
                       int32_t restart, int32_t max_it, int32_t *iters, double *relres) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 4, "solve_system")) return rc;
    if (!tmp_rhs || !phi || !dphi) return fail(h, PBFMM_ERR_ARG, "solve_system: null vector");
    if (restart < 1 || restart > 127 || max_it < 1 || !(rel_tol > 0))
      return fail(h, PBFMM_ERR_ARG, "solve_system: need 1 <= restart <= 127, max_it >= 1, rel_tol > 0");
    ensure_prec(h);
    Engine &e = h->eng;
    const int64_t n = e.n;
    // rhs (compute_rhs) and distribute_rhs of the constrained rows
    h2d(e, e.d_in.as<double>(), tmp_rhs, n);
    DevBuf rhs, sol;
    if (hipMalloc(&rhs.p, n * 8) != hipSuccess || hipMalloc(&sol.p, n * 8) != hipSuccess)
      throw HipError("hipMalloc failed");
    e.compute_rhs(e.d_in.as<double>(), rhs.as<double>());
    std::vector<double> hr(n);
    d2h(e, hr.data(), rhs.as<double>(), n);
    if (e.have_constraints) {
      std::vector<uint8_t> isc(n);
      std::vector<double> inh(n);
      if (hipMemcpy(isc.data(), e.d_c_is.p, n, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(inh.data(), e.d_c_inh.p, n * 8, hipMemcpyDeviceToHost) != hipSuccess)
        throw HipError("constraint copy failed");
      for (int64_t i = 0; i < n; ++i)
        if (isc[i]) hr[i] = inh[i];
      h2d(e, rhs.as<double>(), hr.data(), n);
    }
    // initial guess sol = rhs (bem_problem.cc:1295)
    h2d(e, sol.as<double>(), hr.data(), n);
    int it = 0;
    double rr = 0;
    e.gmres(rhs.as<double>(), sol.as<double>(), rel_tol, restart, max_it, h->prm.precision, &it, &rr);
    std::vector<double> hs(n);
    d2h(e, hs.data(), sol.as<double>(), n);
    for (int64_t i = 0; i < n; ++i) {
      if (e.mesh.dir[i]) dphi[i] = hs[i];
      else phi[i] = hs[i];
    }
    if (iters) *iters = it;
    if (relres) *relres = rr;
    return rr <= rel_tol ? PBFMM_OK : fail(h, PBFMM_ERR_STATE, "solve_system: GMRES did not reach rel_tol");
  });
}

int pbfmm_matvec_dev(pbfmm_handle *h, const double *phi, const double *dphi, double *outN, double *outD,
                     int32_t precision) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 4, "matvec_dev")) return rc;
    if (!phi || !dphi || !outN || !outD) return fail(h, PBFMM_ERR_ARG, "matvec_dev: null vector");
    Engine &e = h->eng;
    if (precision == PBFMM_MIXED) {
      if (!e.have_f32) return fail(h, PBFMM_ERR_STATE, "matvec_dev: handle was not created with PBFMM_MIXED");
      e.matvec<float>(phi, dphi, outN, outD);
    } else {
      e.matvec<double>(phi, dphi, outN, outD);
    }
    return PBFMM_OK;
  });
}

int pbfmm_vmult_dev(pbfmm_handle *h, const double *src, double *dst, int32_t precision) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 4, "vmult_dev")) return rc;
    if (!src || !dst) return fail(h, PBFMM_ERR_ARG, "vmult_dev: null vector");
    if (src == dst) return fail(h, PBFMM_ERR_ARG, "vmult_dev: src and dst must differ");
    if (precision != PBFMM_FP64 && precision != PBFMM_MIXED)
      return fail(h, PBFMM_ERR_ARG, "vmult_dev: unknown precision");
    if (precision == PBFMM_MIXED && !h->eng.have_f32)
      return fail(h, PBFMM_ERR_STATE, "vmult_dev: handle was not created with PBFMM_MIXED");
    ensure_alpha(h);
    h->eng.vmult(src, dst, precision);
    return PBFMM_OK;
  });
}

int pbfmm_vmult_dev_multi(pbfmm_handle *h, const double *src, double *dst, int32_t nrhs, int64_t ld,
                          int32_t precision) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 4, "vmult_dev_multi")) return rc;
    Engine &e = h->eng;
    const int64_t n = e.n;
    if (!src || !dst || nrhs < 1 || ld < n)
      return fail(h, PBFMM_ERR_ARG, "vmult_dev_multi: need src, dst, nrhs >= 1 and ld >= n_dofs");
    const char *s0 = reinterpret_cast<const char *>(src), *d0 = reinterpret_cast<const char *>(dst);
    const size_t span = (size_t)((nrhs - 1) * ld + n) * 8;
    if (s0 < d0 + span && d0 < s0 + span) return fail(h, PBFMM_ERR_ARG, "vmult_dev_multi: src and dst overlap");
    if (precision != PBFMM_FP64 && precision != PBFMM_MIXED)
      return fail(h, PBFMM_ERR_ARG, "vmult_dev_multi: unknown precision");
    if (precision == PBFMM_MIXED && !e.have_f32)
      return fail(h, PBFMM_ERR_STATE, "vmult_dev_multi: handle was not created with PBFMM_MIXED");
    ensure_alpha(h);
    if (!e.have_dirichlet) {
      // vmult's norm shifts (bem_problem.cc:1169-1172, 1214-1215) act on each column
      for (int32_t r = 0; r < nrhs; ++r) e.vmult(src + r * ld, dst + r * ld, precision);
      return PBFMM_OK;
    }
    e.multi_setup(nrhs);
    DevBuf pb, px;
    const double *S = src;
    double *D = dst;
    if (ld != n) {
      if (hipMalloc(&pb.p, (size_t)nrhs * n * 8) != hipSuccess || hipMalloc(&px.p, (size_t)nrhs * n * 8) != hipSuccess)
        throw HipError("vmult_dev_multi: packing buffers");
      if (hipMemcpy2DAsync(pb.p, n * 8, src, ld * 8, n * 8, nrhs, hipMemcpyDeviceToDevice, e.stream) != hipSuccess)
        throw HipError("vmult_dev_multi: packing copy");
      S = pb.as<double>();
      D = px.as<double>();
    }
    e.vmult_multi(S, D, nrhs, precision);
    if (ld != n) {
      if (hipMemcpy2DAsync(dst, ld * 8, D, n * 8, n * 8, nrhs, hipMemcpyDeviceToDevice, e.stream) != hipSuccess)
        throw HipError("vmult_dev_multi: unpacking copy");
      e.sync();  // the packing buffers are freed on return
    }
    return PBFMM_OK;
  });
}

// Byte ranges [a, a + span) and [b, b + span) share memory.
static bool ranges_overlap(const void *a, const void *b, size_t span) {
  const char *p = static_cast<const char *>(a), *q = static_cast<const char *>(b);
  return p < q + span && q < p + span;
}

int pbfmm_gmres_dev(pbfmm_handle *h, const double *rhs, double *x, double rel_tol, int32_t restart, int32_t max_it,
                    int32_t precision, int32_t *iters, double *relres) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 4, "gmres_dev")) return rc;
    if (!rhs || !x) return fail(h, PBFMM_ERR_ARG, "gmres_dev: null vector");
    // the solve reads rhs again after updating x (restart residuals, the
    // outer FP64 residual), so the two may not share memory
    if (ranges_overlap(rhs, x, (size_t)h->eng.n * 8))
      return fail(h, PBFMM_ERR_ARG, "gmres_dev: rhs and x overlap");
    if (restart < 1 || restart > 127 || max_it < 1 || !(rel_tol > 0))
      return fail(h, PBFMM_ERR_ARG, "gmres_dev: need 1 <= restart <= 127, max_it >= 1, rel_tol > 0");
    if (precision == PBFMM_MIXED && !h->eng.have_f32)
      return fail(h, PBFMM_ERR_STATE, "gmres_dev: handle was not created with PBFMM_MIXED");
    ensure_prec(h);
    int it = 0;
    double rr = 0;
    h->eng.gmres(rhs, x, rel_tol, restart, max_it, precision, &it, &rr);
    if (iters) *iters = it;
    if (relres) *relres = rr;
    return rr <= rel_tol ? PBFMM_OK : fail(h, PBFMM_ERR_STATE, "gmres_dev: GMRES did not reach rel_tol");
  });
}

int pbfmm_gmres_dev_multi(pbfmm_handle *h, const double *rhs, double *x, int32_t nrhs, int64_t ld, double rel_tol,
                          int32_t restart, int32_t max_it, int32_t precision, int32_t *iters, double *relres) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 4, "gmres_dev_multi")) return rc;
    if (!rhs || !x || nrhs < 1 || ld < h->eng.n)
      return fail(h, PBFMM_ERR_ARG, "gmres_dev_multi: need rhs, x, nrhs >= 1 and ld >= n_dofs");
    if (ranges_overlap(rhs, x, (size_t)((nrhs - 1) * ld + h->eng.n) * 8))
      return fail(h, PBFMM_ERR_ARG, "gmres_dev_multi: rhs and x overlap");
    if (restart < 1 || restart > 127 || max_it < 1 || !(rel_tol > 0))
      return fail(h, PBFMM_ERR_ARG, "gmres_dev_multi: need 1 <= restart <= 127, max_it >= 1, rel_tol > 0");
    if (precision == PBFMM_MIXED && !h->eng.have_f32)
      return fail(h, PBFMM_ERR_STATE, "gmres_dev_multi: handle was not created with PBFMM_MIXED");
    ensure_prec(h);
    Engine &e = h->eng;
    const int64_t n = e.n;
    // lockstep solve on contiguous columns (packed when ld > n_dofs)
    DevBuf pb, px;
    const double *B = rhs;
    double *X = x;
    if (ld != n) {
      if (hipMalloc(&pb.p, (size_t)nrhs * n * 8) != hipSuccess || hipMalloc(&px.p, (size_t)nrhs * n * 8) != hipSuccess)
        throw HipError("gmres_dev_multi: packing buffers");
      if (hipMemcpy2DAsync(pb.p, n * 8, rhs, ld * 8, n * 8, nrhs, hipMemcpyDeviceToDevice, e.stream) != hipSuccess ||
          hipMemcpy2DAsync(px.p, n * 8, x, ld * 8, n * 8, nrhs, hipMemcpyDeviceToDevice, e.stream) != hipSuccess)
        throw HipError("gmres_dev_multi: packing copy");
      B = pb.as<double>();
      X = px.as<double>();
    }
    std::vector<int> its(nrhs, 0);
    std::vector<double> rrs(nrhs, 0.0);
    const int bad = e.gmres_multi(B, X, nrhs, rel_tol, restart, max_it, precision, its.data(), rrs.data());
    if (ld != n) {
      if (hipMemcpy2DAsync(x, ld * 8, X, n * 8, n * 8, nrhs, hipMemcpyDeviceToDevice, e.stream) != hipSuccess)
        throw HipError("gmres_dev_multi: unpacking copy");
      e.sync();
    }
    int failed = 0;
    for (int32_t r = 0; r < nrhs; ++r) {
      if (iters) iters[r] = its[r];
      if (relres) relres[r] = rrs[r];
      failed += rrs[r] > rel_tol;
    }
    if (bad || failed)
      return fail(h, PBFMM_ERR_STATE, "gmres_dev_multi: " + std::to_string(failed) + " right-hand side(s) did not "
                                      "reach rel_tol");
    return PBFMM_OK;
  });
}

int pbfmm_tree_counts(pbfmm_handle *h, int64_t *cnt) {
  return guarded(h, [&]() -> int {
    if (int rc = not_batch(h, "tree_counts")) return rc;
    if (int rc = need(h, 2, "tree_counts")) return rc;
    const Tree &t = h->eng.tree;
    cnt[0] = t.nb;
    cnt[1] = t.L;
    cnt[2] = (int64_t)t.leaves.size();
    cnt[3] = t.sub_beg[t.nb];
    int64_t a = 0, b = 0;
    for (auto &v : t.il) a += (int64_t)v.size();
    for (auto &v : t.nil) b += (int64_t)v.size();
    cnt[4] = a;
    cnt[5] = b;
    int64_t nn = 0;
    for (int i = 0; i < t.nb; ++i) nn += t.node_cnt[i];
    cnt[6] = nn;
    cnt[7] = h->stage >= 3 ? h->eng.near.nnz : 0;
    return PBFMM_OK;
  }, false);
}

int pbfmm_tree_export(pbfmm_handle *h, int32_t *level, int64_t *parent, double *pmin, double *delta, int64_t *startL,
                      int64_t *endL, int64_t *childless, int64_t *sub_ptr, int64_t *il_ptr, int64_t *il_idx,
                      int64_t *nil_ptr, int64_t *nil_idx, int64_t *node_ptr, int64_t *node_idx) {
  return guarded(h, [&]() -> int {
    if (int rc = not_batch(h, "tree_export")) return rc;
    if (int rc = need(h, 2, "tree_export")) return rc;
    const Tree &t = h->eng.tree;
    for (int l = 0; l <= t.L; ++l) {
      startL[l] = t.startL[l];
      endL[l] = t.endL[l];
    }
    for (size_t k = 0; k < t.leaves.size(); ++k) childless[k] = t.leaves[k];
    int64_t a = 0, b = 0, nd = 0;
    il_ptr[0] = nil_ptr[0] = node_ptr[0] = sub_ptr[0] = 0;
    std::vector<int32_t> nodes;
    for (int i = 0; i < t.nb; ++i) {
      level[i] = t.level[i];
      parent[i] = t.parent[i];
      for (int k = 0; k < 3; ++k) pmin[3 * i + k] = t.pmin[3 * i + k];
      delta[i] = t.delta[i];
      for (int s = 0; s < t.nsub[i]; ++s) {
        const int q = t.sub_beg[i] + s;
        for (int x : t.il[q]) il_idx[a++] = x;
        for (int x : t.nil[q]) nil_idx[b++] = x;
        il_ptr[q + 1] = a;
        nil_ptr[q + 1] = b;
      }
      sub_ptr[i + 1] = t.sub_beg[i + 1];
      nodes.assign(t.nperm.begin() + t.node_beg[i], t.nperm.begin() + t.node_beg[i] + t.node_cnt[i]);
      std::sort(nodes.begin(), nodes.end());
      for (int x : nodes) node_idx[nd++] = x;
      node_ptr[i + 1] = nd;
    }
    return PBFMM_OK;
  }, false);
}

int pbfmm_nearfield_export(pbfmm_handle *h, int32_t precision, int64_t *rowptr, int64_t *col, double *valN,
                           double *valD) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 3, "nearfield_export")) return rc;
    Engine &e = h->eng;
    const NearPlan &np = e.near;
    std::vector<double> v(2 * np.nnz);
    if (precision == PBFMM_MIXED) {
      if (!e.have_f32) return fail(h, PBFMM_ERR_STATE, "nearfield_export: no FP32 near field in this handle");
      std::vector<float> f(2 * np.nnz);
      if (hipMemcpy(f.data(), e.f32.vals.p, f.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
        throw HipError("copy failed");
      for (size_t i = 0; i < f.size(); ++i) v[i] = f[i];
    } else if (hipMemcpy(v.data(), e.f64.vals.p, v.size() * 8, hipMemcpyDeviceToHost) != hipSuccess) {
      throw HipError("copy failed");
    }
    // rows in DoF order
    std::vector<int32_t> row_of(e.n, -1);
    for (size_t r = 0; r < np.rows.size(); ++r) row_of[np.rows[r]] = (int32_t)r;
    rowptr[0] = 0;
    int64_t k = 0;
    for (int64_t i = 0; i < e.n; ++i) {
      const int r = row_of[i];
      if (r >= 0) {
        const int leaf = np.row_leaf[r], ri = r - np.row_beg[leaf];
        const int cb = np.col_beg[leaf], C = np.col_beg[leaf + 1] - cb;
        const int64_t base = np.val_beg[leaf] + (int64_t)ri * C;
        for (int c = 0; c < C; ++c) {
          col[k] = np.cols[cb + c];
          valN[k] = v[2 * (base + c)];
          valD[k] = v[2 * (base + c) + 1];
          ++k;
        }
      }
      rowptr[i + 1] = k;
    }
    return PBFMM_OK;
  });
}

int pbfmm_get_expansions(pbfmm_handle *h, int32_t which, double *k1, double *k2) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 4, "get_expansions")) return rc;
    Engine &e = h->eng;
    const int nb = e.tree.nb, kp = e.kp, kt = e.kt;
    std::vector<double> v((size_t)nb * 2 * kp);
    const DevBuf &src = which == 0 ? e.f64.me : e.f64.le;
    if (hipMemcpy(v.data(), src.p, v.size() * 8, hipMemcpyDeviceToHost) != hipSuccess) throw HipError("copy failed");
    // the device keeps level-scaled expansions (ME_n / delta^n, LE_n delta^(n+1))
    for (int b = 0; b < nb; ++b)
      for (int nn = 0; nn <= e.p; ++nn) {
        const double sc = std::pow(e.tree.delta[b], which == 0 ? nn : -(nn + 1));
        for (int m = 0; m <= nn; ++m) {
          const int i = nn * (nn + 1) / 2 + m;
          k1[(size_t)b * 2 * kt + 2 * i] = v[(size_t)b * 2 * kp + i] * sc;
          k1[(size_t)b * 2 * kt + 2 * i + 1] = v[(size_t)b * 2 * kp + kt + i] * sc;
          k2[(size_t)b * 2 * kt + 2 * i] = v[(size_t)b * 2 * kp + kp + i] * sc;
          k2[(size_t)b * 2 * kt + 2 * i + 1] = v[(size_t)b * 2 * kp + kp + kt + i] * sc;
        }
      }
    return PBFMM_OK;
  });
}

int pbfmm_set_phase_timing(pbfmm_handle *h, int32_t enable) {
  return guarded(h, [&]() -> int {
    h->eng.time_phases = enable != 0;
    return PBFMM_OK;
  });
}

int pbfmm_work_counts(pbfmm_handle *h, int64_t *c) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 4, "work_counts")) return rc;
    if (!c) return fail(h, PBFMM_ERR_ARG, "work_counts: null output");
    const Engine &e = h->eng;
    c[0] = e.n;
    c[1] = e.near.nnz;
    c[2] = (int64_t)e.near.cols.size();
    c[3] = (int64_t)e.near.rows.size();
    c[4] = (int64_t)e.far.p2m_leaf.size();
    c[5] = e.mesh.ncells * e.nq;
    c[6] = (int64_t)e.far.m2l.tgt.size();
    c[7] = e.far.m2p_work_beg.empty() ? 0 : e.far.m2p_work_beg.back();
    c[8] = (int64_t)e.far.m2m.tgt.size();
    c[9] = (int64_t)e.far.l2l.tgt.size();
    c[10] = e.tree.nb;
    c[11] = e.p;
    c[12] = e.kp;
    c[13] = (int64_t)e.far.p2m_col_dof.size();
    c[14] = e.spmv_variant[0][0] + 8 * e.spmv_variant[0][1];
    c[15] = e.spmv_variant[1][0] + 8 * e.spmv_variant[1][1];
    return PBFMM_OK;
  }, false);
}

int pbfmm_last_timings(pbfmm_handle *h, double *ms) {
  return guarded(h, [&]() -> int {
    if (!ms) return fail(h, PBFMM_ERR_ARG, "last_timings: null output");
    for (int i = 0; i < 9; ++i) ms[i] = h->eng.last_ms[i];
    return PBFMM_OK;
  }, false);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Post-solve (BEMProblem, bem_problem.cc:1746-2370)
// ---------------------------------------------------------------------------
static int ps_project_host(pbfmm_handle *h, int mode, const double *in1, const double *in2, double *out,
                           const char *what) {
  return guarded(h, [&]() -> int {
    if (int rc = ps_ready(h, what)) return rc;
    if (!out || (mode >= 1 && !in1) || (mode == 1 && !in2))
      return fail(h, PBFMM_ERR_ARG, std::string(what) + ": null vector");
    Engine &e = h->eng;
    double *d1 = nullptr, *d2 = nullptr;
    if (mode >= 1) {
      d1 = (double *)stage(e.ps.in1, (size_t)e.n * 8);
      h2d(e, d1, in1, e.n);
    }
    if (mode == 1) {
      d2 = (double *)stage(e.ps.in2, (size_t)e.n * 8);
      h2d(e, d2, in2, e.n);
    }
    double *dout = (double *)stage(e.ps.out, (size_t)3 * e.n * 8);
    e.ps_project(mode, d1, d2, dout);
    if (mode == 1) {
      double *g = (double *)stage(e.ps.grad, (size_t)3 * e.n * 8);
      if (hipMemcpyAsync(g, dout, (size_t)3 * e.n * 8, hipMemcpyDeviceToDevice, e.stream) != hipSuccess)
        throw HipError("gradient copy failed");
      e.ps.have_grad = true;
    }
    d2h(e, out, dout, 3 * e.n);
    return PBFMM_OK;
  });
}

extern "C" {

int pbfmm_compute_normals(pbfmm_handle *h, double *normals) {
  return ps_project_host(h, 0, nullptr, nullptr, normals, "compute_normals");
}

int pbfmm_compute_gradients(pbfmm_handle *h, const double *phi, const double *dphi_dn, double *gradients) {
  return ps_project_host(h, 1, phi, dphi_dn, gradients, "compute_gradients");
}

int pbfmm_compute_surface_gradients(pbfmm_handle *h, const double *tmp_rhs, double *gradients) {
  return ps_project_host(h, 2, tmp_rhs, nullptr, gradients, "compute_surface_gradients");
}

int pbfmm_preconditioner_blocks(pbfmm_handle *h, int64_t *sizes, int32_t *row_beg, int32_t *rows, double *values) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 4, "preconditioner_blocks")) return rc;
    if (!sizes) return fail(h, PBFMM_ERR_ARG, "preconditioner_blocks: null sizes");
    if (h->prm.preconditioner != 1) return fail(h, PBFMM_ERR_STATE, "preconditioner_blocks: no preconditioner");
    ensure_prec(h);
    Engine &e = h->eng;
    sizes[0] = e.nprec;
    sizes[1] = (int64_t)e.prows.size();
    sizes[2] = e.pinv_beg.empty() ? 0 : e.pinv_beg[e.nprec];
    if (!row_beg && !rows && !values) return PBFMM_OK;
    if (!row_beg || !rows || !values) return fail(h, PBFMM_ERR_ARG, "preconditioner_blocks: null output");
    std::copy(e.prow_beg.begin(), e.prow_beg.end(), row_beg);
    std::copy(e.prows.begin(), e.prows.end(), rows);
    std::vector<double> blk;
    e.prec_blocks(blk);
    std::copy(blk.begin(), blk.end(), values);
    return PBFMM_OK;
  });
}

int pbfmm_set_mixed_scheme(pbfmm_handle *h, int32_t scheme) {
  if (!h) return PBFMM_ERR_ARG;
  if (scheme < 0 || scheme > 4) return fail(h, PBFMM_ERR_ARG, "set_mixed_scheme: scheme must be 0..4");
  h->eng.mixed_scheme = scheme;
  return PBFMM_OK;
}

int pbfmm_set_mixed_tolerances(pbfmm_handle *h, double inner_reduction, double relax_switch) {
  if (!h) return PBFMM_ERR_ARG;
  if (!(inner_reduction < 1.0) || std::isnan(relax_switch))
    return fail(h, PBFMM_ERR_ARG, "set_mixed_tolerances: inner_reduction must be below 1");
  h->eng.inner_reduction = inner_reduction > 0 ? inner_reduction : 0.0;
  h->eng.relax_switch = relax_switch > 0 ? relax_switch : 0.0;
  return PBFMM_OK;
}

int pbfmm_set_gmres_reorth(pbfmm_handle *h, double eta) {
  if (!h) return PBFMM_ERR_ARG;
  if (std::isnan(eta)) return fail(h, PBFMM_ERR_ARG, "set_gmres_reorth: eta is NaN");
  h->eng.reorth_eta = eta < 0 ? Engine::kReorthEta : eta;
  return PBFMM_OK;
}

int pbfmm_set_spmv_kernel(pbfmm_handle *h, int32_t precision, int32_t kernel) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 3, "set_spmv_kernel")) return rc;
    if (precision != PBFMM_FP64 && precision != PBFMM_MIXED)
      return fail(h, PBFMM_ERR_ARG, "set_spmv_kernel: unknown precision");
    if (kernel < 0 || kernel > (precision == PBFMM_MIXED ? 6 : 3))
      return fail(h, PBFMM_ERR_ARG, "set_spmv_kernel: kernel must be 0..3 (FP64) or 0..6 (FP32)");
    if (precision == PBFMM_MIXED && !h->eng.have_f32)
      return fail(h, PBFMM_ERR_STATE, "set_spmv_kernel: handle was not created with PBFMM_MIXED");
    const int pi = precision == PBFMM_MIXED ? 1 : 0;
    h->eng.spmv_variant[pi][0] = kernel;
    h->eng.spmv_variant[pi][1] = kernel % 3;
    return PBFMM_OK;
  });
}

int pbfmm_set_side_fusion(pbfmm_handle *h, int32_t precision, int32_t mode) {
  if (!h) return PBFMM_ERR_ARG;
  if (mode != -1 && !Engine::valid_plan(mode))
    return fail(h, PBFMM_ERR_ARG, "set_side_fusion: plan must be -1, 0, 1, 2, 3, 4, 5, 8, 12, 24, 28, 56, 60, 88 or 92");
  if (precision != PBFMM_FP64 && precision != PBFMM_MIXED)
    return fail(h, PBFMM_ERR_ARG, "set_side_fusion: unknown precision");
  h->eng.side_fusion[precision == PBFMM_MIXED ? 1 : 0] = mode;
  return PBFMM_OK;
}

int pbfmm_tuning_state(pbfmm_handle *h, int32_t *state) {
  if (!h || !state) return PBFMM_ERR_ARG;
  const Engine &e = h->eng;
  const bool timed = h->stage >= 3;
  state[0] = timed ? e.spmv_variant[0][0] : -1;
  state[1] = timed && e.have_f32 ? e.spmv_variant[1][0] : -1;
  state[2] = e.side_fusion[0];
  state[3] = e.side_fusion[1];
  state[4] = e.mixed_scheme != 0 ? e.mixed_scheme
                                 : (e.op_cost_ratio > 0 ? (e.op_cost_ratio < kRelaxCostRatio ? 1 : 2) : -1);
  state[5] = e.op_cost_ratio > 0 ? (int32_t)(e.op_cost_ratio * 1000.0 + 0.5) : -1;
  state[6] = e.graph_enabled ? e.graph_use[0] : 0;
  state[7] = e.graph_enabled ? e.graph_use[1] : 0;
  return PBFMM_OK;
}

int pbfmm_set_graph_mode(pbfmm_handle *h, int32_t mode) {
  if (!h) return PBFMM_ERR_ARG;
  if (mode < -1 || mode > 1) return fail(h, PBFMM_ERR_ARG, "set_graph_mode: mode must be -1, 0 or 1");
  Engine &e = h->eng;
  e.graph_enabled = mode != 0;
  e.graph_use[0] = e.graph_use[1] = mode;
  return PBFMM_OK;
}

int pbfmm_graph_timings(pbfmm_handle *h, float *ms) {
  if (!h || !ms) return PBFMM_ERR_ARG;
  for (int f = 0; f < 2; ++f)
    for (int k = 0; k < 2; ++k) ms[2 * f + k] = h->eng.graph_ms[f][k];
  return PBFMM_OK;
}

int pbfmm_projection_iterations(pbfmm_handle *h, int32_t *iters) {
  if (!h || !iters) return PBFMM_ERR_ARG;
  *iters = h->eng.ps.last_iters;
  return PBFMM_OK;
}

int pbfmm_hydrostatic_pressure(pbfmm_handle *h, double gravity, double density, double z0, double *pressure) {
  return guarded(h, [&]() -> int {
    if (int rc = ps_ready(h, "hydrostatic_pressure")) return rc;
    if (!pressure) return fail(h, PBFMM_ERR_ARG, "hydrostatic_pressure: null vector");
    Engine &e = h->eng;
    double *dp = (double *)stage(e.ps.out, (size_t)3 * e.n * 8);
    e.ps_hydrostatic(gravity, density, z0, dp);
    d2h(e, pressure, dp, e.n);
    return PBFMM_OK;
  });
}

int pbfmm_hydrodynamic_pressure(pbfmm_handle *h, double density, const double *wind, double *pressure) {
  return guarded(h, [&]() -> int {
    if (int rc = ps_ready(h, "hydrodynamic_pressure")) return rc;
    if (!wind || !pressure) return fail(h, PBFMM_ERR_ARG, "hydrodynamic_pressure: null vector");
    Engine &e = h->eng;
    if (!e.ps.have_grad)
      return fail(h, PBFMM_ERR_STATE, "hydrodynamic_pressure called before compute_gradients");
    double *dw = (double *)stage(e.ps.aux, (size_t)3 * e.n * 8);
    h2d(e, dw, wind, 3 * e.n);
    double *dp = (double *)stage(e.ps.out, (size_t)3 * e.n * 8);
    e.ps_hydrodynamic(density, dw, e.ps.grad.as<double>(), dp);
    d2h(e, pressure, dp, e.n);
    return PBFMM_OK;
  });
}

int pbfmm_center_of_pressure(pbfmm_handle *h, const uint8_t *cell_mask, const double *pressure, double *center) {
  return guarded(h, [&]() -> int {
    if (int rc = ps_ready(h, "center_of_pressure")) return rc;
    if (!pressure || !center) return fail(h, PBFMM_ERR_ARG, "center_of_pressure: null argument");
    double s[13];
    ps_sums(h, cell_mask, pressure, nullptr, s);
    if (std::fabs(s[9]) > 0.0)
      for (int d = 0; d < 3; ++d) center[d] = s[6 + d] / s[9];
    return PBFMM_OK;
  });
}

int pbfmm_pressure_force_and_moment(pbfmm_handle *h, const uint8_t *cell_mask, const double *pressure,
                                    const double *cog, double *force, double *moment) {
  return guarded(h, [&]() -> int {
    if (int rc = ps_ready(h, "pressure_force_and_moment")) return rc;
    if (!pressure || !cog || !force || !moment)
      return fail(h, PBFMM_ERR_ARG, "pressure_force_and_moment: null argument");
    double s[13];
    ps_sums(h, cell_mask, pressure, cog, s);
    for (int d = 0; d < 3; ++d) {
      force[d] = s[d];
      moment[d] = s[3 + d];
    }
    return PBFMM_OK;
  });
}

int pbfmm_pressure_force_and_moment_noalloc(pbfmm_handle *h, const uint8_t *d_cell_mask, const double *d_pressure,
                                            const double *cog, double *d_out) {
  return guarded(h, [&]() -> int {
    if (int rc = not_batch(h, "pressure_force_and_moment_noalloc")) return rc;
    if (int rc = need(h, 1, "pressure_force_and_moment_noalloc")) return rc;
    Engine &e = h->eng;
    if (!e.ps.ready)
      return fail(h, PBFMM_ERR_STATE,
                  "pressure_force_and_moment_noalloc: post-solve data not set up (any post-solve call sets it up)");
    if (!d_pressure || !cog || !d_out) return fail(h, PBFMM_ERR_ARG, "pressure_force_and_moment_noalloc: null argument");
    if (e.ps.ncells > ps_scratch_max_cells())
      return fail(h, PBFMM_ERR_ARG, "pressure_force_and_moment_noalloc: more than " +
                                        std::to_string(ps_scratch_max_cells()) + " cells");
    e.ps_force_moment_noalloc(d_cell_mask, d_pressure, cog, d_out);
    return PBFMM_OK;
  });
}

int pbfmm_volume_integral(pbfmm_handle *h, const uint8_t *cell_mask, double *volume) {
  return guarded(h, [&]() -> int {
    if (int rc = ps_ready(h, "volume_integral")) return rc;
    if (!volume) return fail(h, PBFMM_ERR_ARG, "volume_integral: null argument");
    double s[13];
    ps_sums(h, cell_mask, nullptr, nullptr, s);
    for (int d = 0; d < 3; ++d) volume[d] = s[10 + d];
    return PBFMM_OK;
  });
}

int pbfmm_free_surface_elevation(pbfmm_handle *h, double gravity, double density, const int32_t *cell_lines,
                                 int64_t n_lines, const double *pressure, double *elevation) {
  return guarded(h, [&]() -> int {
    if (int rc = ps_ready(h, "free_surface_elevation")) return rc;
    if (n_lines < 0 || (n_lines > 0 && (!cell_lines || !pressure || !elevation)))
      return fail(h, PBFMM_ERR_ARG, "free_surface_elevation: bad arguments");
    if (!(density * gravity != 0.0)) return fail(h, PBFMM_ERR_ARG, "free_surface_elevation: density * gravity is 0");
    Engine &e = h->eng;
    for (int64_t k = 0; k < n_lines; ++k)
      if (cell_lines[2 * k] < 0 || cell_lines[2 * k] >= e.ps.ncells || cell_lines[2 * k + 1] < 0 ||
          cell_lines[2 * k + 1] > 3)
        return fail(h, PBFMM_ERR_ARG, "free_surface_elevation: (cell, line) pair out of range");
    if (n_lines == 0) return PBFMM_OK;
    int32_t *df = (int32_t *)stage(e.ps.mask, (size_t)n_lines * 8);
    h2d_bytes(e, df, cell_lines, (size_t)n_lines * 8);
    double *dp = (double *)stage(e.ps.in1, (size_t)e.n * 8);
    h2d(e, dp, pressure, e.n);
    const int64_t nout = n_lines * e.ps.nq1 * 3;
    double *dout = (double *)stage(e.ps.out, (size_t)std::max<int64_t>(nout, 3 * e.n) * 8);
    e.ps_elevation(df, n_lines, dp, density * gravity, dout);
    d2h(e, elevation, dout, nout);
    return PBFMM_OK;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches of independent problems (a forest of octrees as one operator)
// ---------------------------------------------------------------------------
extern "C" {

int pbfmm_batch_create(pbfmm_handle *const *parts, int32_t n_parts, pbfmm_handle **out) {
  if (!out) return PBFMM_ERR_ARG;
  *out = nullptr;
  if (!parts || n_parts < 1) return PBFMM_ERR_ARG;
  for (int32_t s = 0; s < n_parts; ++s)
    if (!parts[s] || parts[s]->stage < 1 || parts[s]->batch) return PBFMM_ERR_STATE;
  // host set-up of every part (octree, near-field and far-field plans) on
  // the host's cores, parts in parallel; nothing here touches the GPU
  std::atomic<int32_t> next{0};
  std::vector<std::string> errs(n_parts);
  auto work = [&]() {
    for (int32_t s; (s = next++) < n_parts;) {
      pbfmm_handle *q = parts[s];
      Engine &e = q->eng;
      try {
        if (q->stage < 2) {
          std::vector<double> gx, gw;
          gauss_legendre01(q->prm.quad_order, gx, gw);
          compute_quad_data(e.mesh, tensor_rule(gx, gw, gx, gw), e.qd);
          e.nq = e.qd.nq;
          build_tree(e.mesh, e.qd, q->prm.num_octree_levels, q->prm.max_nodes_per_leaf, e.tree);
          q->stage = 2;
        }
        if (q->stage < 4) {
          build_near_plan(e.mesh, e.qd, e.tree, e.near);
          build_far_plan(e.tree, e.p, e.far);
          build_p2m_plan(e.mesh, e.qd, e.tree, e.far);
        }
      } catch (const std::exception &ex) {
        errs[s] = ex.what();
      }
    }
  };
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const unsigned nt = std::min<unsigned>(hw, (unsigned)n_parts);
  std::vector<std::thread> pool;
  for (unsigned k = 1; k < nt; ++k) pool.emplace_back(work);
  work();
  for (auto &th : pool) th.join();
  for (int32_t s = 0; s < n_parts; ++s)
    if (!errs[s].empty()) {
      parts[s]->err = errs[s];
      return PBFMM_ERR_ARG;
    }
  pbfmm_handle *h = nullptr;
  try {
    h = new pbfmm_handle();
  } catch (...) {
    return PBFMM_ERR_ALLOC;
  }
  h->prm = parts[0]->prm;
  h->batch = true;
  std::vector<const Engine *> engs;
  for (int32_t s = 0; s < n_parts; ++s) engs.push_back(&parts[s]->eng);
  // host merge and index checks first: nothing reaches the device unless
  // every index of the merged plans is in range
  int rc = guarded(h, [&]() -> int {
    Engine &e = h->eng;
    for (int32_t s = 0; s < n_parts; ++s)
      if (parts[s]->prm.precision != h->prm.precision || parts[s]->prm.device != h->prm.device)
        return fail(h, PBFMM_ERR_ARG, "batch_create: parts differ in precision or device");
    merge_forest(engs, e, e.seg_off);
    validate_plans(e);
    if (e.near.nnz >= (int64_t(1) << 40)) return fail(h, PBFMM_ERR_ARG, "batch_create: near field too large");
    return PBFMM_OK;
  }, false);
  if (rc == PBFMM_OK) rc = guarded(h, [&]() -> int {
    Engine &e = h->eng;
    const bool mixed = h->prm.precision == PBFMM_MIXED;
    e.upload_mesh_and_tree();
    e.upload_quad(mixed);
    e.upload_near_plan();
    e.assemble_near(mixed);
    e.upload_far(mixed);
    e.build_unit_multipoles(mixed);
    e.forest_segments();
    h->stage = 4;
    return PBFMM_OK;
  });
  if (rc != PBFMM_OK) {
    if (parts[0]) parts[0]->err = "batch_create: " + h->err;
    delete h;
    return rc;
  }
  *out = h;
  return PBFMM_OK;
}

int pbfmm_batch_offsets(pbfmm_handle *h, int64_t *dof_off) {
  if (!h || !dof_off) return PBFMM_ERR_ARG;
  if (!h->batch) return fail(h, PBFMM_ERR_STATE, "batch_offsets: not a batch handle");
  for (size_t s = 0; s < h->eng.seg_off.size(); ++s) dof_off[s] = h->eng.seg_off[s];
  return PBFMM_OK;
}

int pbfmm_batch_gmres_dev(pbfmm_handle *h, const double *d_rhs, double *d_x, double rel_tol, int32_t restart,
                          int32_t max_it, int32_t precision, int32_t *iters, double *relres) {
  return guarded(h, [&]() -> int {
    if (int rc = need(h, 4, "batch_gmres_dev")) return rc;
    if (!h->batch) return fail(h, PBFMM_ERR_STATE, "batch_gmres_dev: not a batch handle");
    if (!d_rhs || !d_x || restart < 1 || max_it < 1 || !(rel_tol > 0))
      return fail(h, PBFMM_ERR_ARG, "batch_gmres_dev: bad arguments");
    if (ranges_overlap(d_rhs, d_x, (size_t)h->eng.n * 8))
      return fail(h, PBFMM_ERR_ARG, "batch_gmres_dev: rhs and x overlap");
    if (precision != PBFMM_FP64 && precision != PBFMM_MIXED)
      return fail(h, PBFMM_ERR_ARG, "batch_gmres_dev: unknown precision");
    ensure_prec(h);
    int it = 0;
    const int bad = h->eng.gmres_forest(d_rhs, d_x, rel_tol, restart, max_it, precision, &it, relres);
    if (iters) *iters = it;
    if (bad) return fail(h, PBFMM_ERR_STATE, "batch_gmres_dev: a part did not reach rel_tol");
    return PBFMM_OK;
  });
}

}  // extern "C"
