// Post-solve on the GPU: the L2-projected nodal normals and gradients of
// BEMProblem (bem_problem.cc:1746-2079), the pressures and the body integrals
// of the Driver's post_process (bem_problem.cc:2081-2370, driver.cc:105-160).
//
// The projections solve M u_d = b_d for the three components at once, where M
// is the consistent scalar mass matrix of the real cells (the vector element
// is three copies of the scalar one, and the reference's vector constraints
// hold only hanging nodes, of which uniformly refined meshes have none).  M
// is assembled without atomics: per-cell local matrices first, then one
// thread per row adds its cells' contributions in ascending cell order, so
// the result does not depend on scheduling.  The solve is Jacobi-
// preconditioned CG with its scalars kept on the device; the host reads the
// residuals every few steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "engine.h"

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess)                                                                \
      throw pbfmm::HipError(std::string(#x) + ": " + hipGetErrorString(e_));             \
  } while (0)

namespace pbfmm {

namespace {

constexpr int kPsThreads = 256;
constexpr int kPsRedChunk = kPsThreads * 8;
constexpr double kPsTol = 1e-13;   // relative CG residual per component
constexpr int kPsMaxIt = 4000;
constexpr int kPsCheck = 8;        // host reads the residuals every kPsCheck steps

template <typename T>
void ps_up(DevBuf &b, const std::vector<T> &v) {
  b.release();
  const size_t bytes = std::max<size_t>(v.size() * sizeof(T), 16);
  HIPCHK(hipMalloc(&b.p, bytes));
  b.bytes = bytes;
  if (!v.empty()) HIPCHK(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}
void ps_alloc(DevBuf &b, size_t bytes) {
  if (b.p && b.bytes >= bytes) return;
  b.release();
  bytes = std::max<size_t>(bytes, 16);
  HIPCHK(hipMalloc(&b.p, bytes));
  b.bytes = bytes;
}
inline unsigned ps_blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

__device__ __forceinline__ double ps_wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
// Block sum returned on thread 0.
__device__ __forceinline__ double ps_block_sum(double v, double *sh) {
  v = ps_wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
  return r;
}

// Geometry of one quadrature point of a cell: position, tangents, unit
// normal t1 x t2 / |t1 x t2| (the convention of the near-field assembly) and
// JxW.  tab holds per point [w, N[DPC], dNx[DPC], dNy[DPC]].
template <int DPC>
struct QGeo {
  double x[3], t1[3], t2[3], n[3], jxw;
  const double *N, *dNx, *dNy;
  __device__ void eval(const double *__restrict__ G, const double *__restrict__ tq) {
    N = tq + 1;
    dNx = tq + 1 + DPC;
    dNy = tq + 1 + 2 * DPC;
    for (int k = 0; k < 3; ++k) x[k] = t1[k] = t2[k] = 0.0;
    for (int j = 0; j < DPC; ++j)
      for (int k = 0; k < 3; ++k) {
        const double g = G[3 * j + k];
        x[k] += N[j] * g;
        t1[k] += dNx[j] * g;
        t2[k] += dNy[j] * g;
      }
    const double c0 = t1[1] * t2[2] - t1[2] * t2[1];
    const double c1 = t1[2] * t2[0] - t1[0] * t2[2];
    const double c2 = t1[0] * t2[1] - t1[1] * t2[0];
    const double J = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    n[0] = c0 / J;
    n[1] = c1 / J;
    n[2] = c2 / J;
    jxw = J * tq[0];
  }
  // surface gradient of sum_j u_j N_j: J (J^T J)^{-1} grad_ref, J = [t1 t2]
  __device__ void surf_grad(const double *u, double g[3]) const {
    double ux = 0, uy = 0;
    for (int j = 0; j < DPC; ++j) {
      ux += u[j] * dNx[j];
      uy += u[j] * dNy[j];
    }
    const double g11 = t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2];
    const double g12 = t1[0] * t2[0] + t1[1] * t2[1] + t1[2] * t2[2];
    const double g22 = t2[0] * t2[0] + t2[1] * t2[1] + t2[2] * t2[2];
    const double det = g11 * g22 - g12 * g12;
    const double a = (g22 * ux - g12 * uy) / det;
    const double b = (g11 * uy - g12 * ux) / det;
    for (int k = 0; k < 3; ++k) g[k] = a * t1[k] + b * t2[k];
  }
};

// Local mass matrices (MODE 0, loc[c][DPC*DPC]) or projection right-hand
// sides (loc[c][3][DPC]): MODE 1 the cell normal (compute_normals,
// bem_problem.cc:2041-2055), MODE 2 n dphi_dn + surface gradient of phi
// (compute_gradients, 1818-1852), MODE 3 the surface gradient of in1
// (compute_surface_gradients, 1943-1968).
template <int DPC, int MODE>
__global__ void __launch_bounds__(kPsThreads)
    k_ps_local(int64_t ncells, int nq, const double *__restrict__ tab, const double *__restrict__ geom,
               const int32_t *__restrict__ cdofs, const double *__restrict__ in1,
               const double *__restrict__ in2, double *__restrict__ loc) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const double *G = geom + c * DPC * 3;
  constexpr int NL = MODE == 0 ? DPC * DPC : 3 * DPC;
  double acc[NL];
  for (int i = 0; i < NL; ++i) acc[i] = 0.0;
  double u[DPC], v[DPC];
  if constexpr (MODE >= 2) {
    for (int j = 0; j < DPC; ++j) {
      const int32_t d = cdofs[c * DPC + j];
      u[j] = in1[d];
      v[j] = MODE == 2 ? in2[d] : 0.0;
    }
  }
  QGeo<DPC> g;
  for (int q = 0; q < nq; ++q) {
    g.eval(G, tab + q * (1 + 3 * DPC));
    if constexpr (MODE == 0) {
      for (int i = 0; i < DPC; ++i)
        for (int j = 0; j < DPC; ++j) acc[i * DPC + j] += g.N[i] * g.N[j] * g.jxw;
    } else {
      double f[3];
      if constexpr (MODE == 1) {
        f[0] = g.n[0];
        f[1] = g.n[1];
        f[2] = g.n[2];
      } else {
        g.surf_grad(u, f);
        if constexpr (MODE == 2) {
          double dn = 0;
          for (int j = 0; j < DPC; ++j) dn += g.N[j] * v[j];
          for (int k = 0; k < 3; ++k) f[k] = g.n[k] * dn + f[k];
        }
      }
      for (int k = 0; k < 3; ++k)
        for (int i = 0; i < DPC; ++i) acc[k * DPC + i] += g.N[i] * f[k] * g.jxw;
    }
  }
  double *o = loc + c * NL;
  for (int i = 0; i < NL; ++i) o[i] = acc[i];
}

// Mass matrix rows: row r adds, for each of its cells in ascending order, the
// local row of r at the CSR positions pos[c][i][.].
template <int DPC>
__global__ void k_ps_gather_mass(int64_t n, const int32_t *__restrict__ d2c_ptr, const int32_t *__restrict__ d2c_cell,
                                 const int8_t *__restrict__ d2c_loc, const int32_t *__restrict__ pos,
                                 const int32_t *__restrict__ m_ptr, const int32_t *__restrict__ m_col,
                                 const double *__restrict__ loc, double *__restrict__ val,
                                 double *__restrict__ diag) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int32_t k = m_ptr[r]; k < m_ptr[r + 1]; ++k) val[k] = 0.0;
  for (int32_t t = d2c_ptr[r]; t < d2c_ptr[r + 1]; ++t) {
    const int64_t c = d2c_cell[t];
    const int i = d2c_loc[t];
    for (int j = 0; j < DPC; ++j) val[pos[(c * DPC + i) * DPC + j]] += loc[(c * DPC + i) * DPC + j];
  }
  double dg = 0.0;
  for (int32_t k = m_ptr[r]; k < m_ptr[r + 1]; ++k)
    if (m_col[k] == r) dg = val[k];
  diag[r] = dg != 0.0 ? dg : 1.0;  // a DoF without cells keeps a zero solution
}

// b[d*n + r] = sum over the cells of r (ascending) of loc[c][d][i]
template <int DPC>
__global__ void k_ps_gather_rhs(int64_t n, const int32_t *__restrict__ d2c_ptr, const int32_t *__restrict__ d2c_cell,
                                const int8_t *__restrict__ d2c_loc, const double *__restrict__ loc,
                                double *__restrict__ b) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  double a0 = 0, a1 = 0, a2 = 0;
  for (int32_t t = d2c_ptr[r]; t < d2c_ptr[r + 1]; ++t) {
    const double *l = loc + (int64_t)d2c_cell[t] * 3 * DPC + d2c_loc[t];
    a0 += l[0];
    a1 += l[DPC];
    a2 += l[2 * DPC];
  }
  b[r] = a0;
  b[n + r] = a1;
  b[2 * n + r] = a2;
}

// CG state cg[]: 0..2 rz, 3..5 pq, 6..8 rr, 9..11 bb, 12..14 alpha, 15..17 beta
enum { kRZ = 0, kPQ = 3, kRR = 6, kBB = 9, kAL = 12, kBE = 15, kCgLen = 18 };

// x = 0, r = b, z = r / diag, p = z; partial sums of r.z and b.b per component
__global__ void __launch_bounds__(kPsThreads)
    k_ps_cg_init(int64_t n, const double *__restrict__ b, const double *__restrict__ diag, double *__restrict__ x,
                 double *__restrict__ r, double *__restrict__ z, double *__restrict__ p,
                 double *__restrict__ part) {
  __shared__ double sh[kPsThreads / 64];
  const int64_t beg = (int64_t)blockIdx.x * kPsRedChunk, end = min(n, beg + (int64_t)kPsRedChunk);
  for (int d = 0; d < 3; ++d) {
    double rz = 0, bb = 0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kPsThreads) {
      const int64_t k = d * n + i;
      const double bi = b[k], zi = bi / diag[i];
      x[k] = 0.0;
      r[k] = bi;
      z[k] = zi;
      p[k] = zi;
      rz += bi * zi;
      bb += bi * bi;
    }
    const double s0 = ps_block_sum(rz, sh);
    if (threadIdx.x == 0) part[(kRZ + d) * gridDim.x + blockIdx.x] = s0;
    const double s1 = ps_block_sum(bb, sh);
    if (threadIdx.x == 0) part[(kBB + d) * gridDim.x + blockIdx.x] = s1;
  }
}
// q = M p (three components), partial sums of p.q
__global__ void __launch_bounds__(kPsThreads)
    k_ps_cg_spmv(int64_t n, const int32_t *__restrict__ m_ptr, const int32_t *__restrict__ m_col,
                 const double *__restrict__ val, const double *__restrict__ p, double *__restrict__ q,
                 double *__restrict__ part) {
  __shared__ double sh[kPsThreads / 64];
  const int64_t beg = (int64_t)blockIdx.x * kPsRedChunk, end = min(n, beg + (int64_t)kPsRedChunk);
  double pq[3] = {0, 0, 0};
  for (int64_t i = beg + threadIdx.x; i < end; i += kPsThreads) {
    double a0 = 0, a1 = 0, a2 = 0;
    for (int32_t k = m_ptr[i]; k < m_ptr[i + 1]; ++k) {
      const double v = val[k];
      const int64_t j = m_col[k];
      a0 += v * p[j];
      a1 += v * p[n + j];
      a2 += v * p[2 * n + j];
    }
    q[i] = a0;
    q[n + i] = a1;
    q[2 * n + i] = a2;
    pq[0] += p[i] * a0;
    pq[1] += p[n + i] * a1;
    pq[2] += p[2 * n + i] * a2;
  }
  for (int d = 0; d < 3; ++d) {
    const double s = ps_block_sum(pq[d], sh);
    if (threadIdx.x == 0) part[(kPQ + d) * gridDim.x + blockIdx.x] = s;
  }
}
// finish the partial sums of slots [s0, s0 + cnt) of cg[]
__global__ void k_ps_finish(int nchunk, int s0, const double *__restrict__ part, double *__restrict__ cg) {
  __shared__ double sh[kPsThreads / 64];
  const int s = s0 + blockIdx.x;
  double acc = 0;
  for (int i = threadIdx.x; i < nchunk; i += blockDim.x) acc += part[(int64_t)s * nchunk + i];
  const double r = ps_block_sum(acc, sh);
  if (threadIdx.x == 0) cg[s] = r;
}
__global__ void k_fill_zero6(double *out) {
  if (threadIdx.x < 6) out[threadIdx.x] = 0.0;
}
__global__ void k_ps_alpha(double *cg) {
  const int d = threadIdx.x;
  if (d < 3) cg[kAL + d] = cg[kPQ + d] > 0.0 ? cg[kRZ + d] / cg[kPQ + d] : 0.0;
}
// x += alpha p, r -= alpha q, z = r / diag; partial sums of r.r and the new r.z
// (written to the pq slots, which are free until the next product)
__global__ void __launch_bounds__(kPsThreads)
    k_ps_cg_update(int64_t n, const double *__restrict__ cg, const double *__restrict__ diag,
                   const double *__restrict__ p, const double *__restrict__ q, double *__restrict__ x,
                   double *__restrict__ r, double *__restrict__ z, double *__restrict__ part) {
  __shared__ double sh[kPsThreads / 64];
  const int64_t beg = (int64_t)blockIdx.x * kPsRedChunk, end = min(n, beg + (int64_t)kPsRedChunk);
  for (int d = 0; d < 3; ++d) {
    const double al = cg[kAL + d];
    double rr = 0, rz = 0;
    for (int64_t i = beg + threadIdx.x; i < end; i += kPsThreads) {
      const int64_t k = d * n + i;
      x[k] += al * p[k];
      const double ri = r[k] - al * q[k];
      const double zi = ri / diag[i];
      r[k] = ri;
      z[k] = zi;
      rr += ri * ri;
      rz += ri * zi;
    }
    const double s0 = ps_block_sum(rr, sh);
    if (threadIdx.x == 0) part[(kRR + d) * gridDim.x + blockIdx.x] = s0;
    const double s1 = ps_block_sum(rz, sh);
    if (threadIdx.x == 0) part[(kPQ + d) * gridDim.x + blockIdx.x] = s1;
  }
}
// beta = rz_new / rz, rz = rz_new (rz_new was finished into the pq slots)
__global__ void k_ps_beta(double *cg) {
  const int d = threadIdx.x;
  if (d < 3) {
    const double rzn = cg[kPQ + d];
    cg[kBE + d] = cg[kRZ + d] > 0.0 ? rzn / cg[kRZ + d] : 0.0;
    cg[kRZ + d] = rzn;
  }
}
__global__ void k_ps_cg_dir(int64_t n3, int64_t n, const double *__restrict__ cg, const double *__restrict__ z,
                            double *__restrict__ p) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= n3) return;
  p[k] = z[k] + cg[kBE + (int)(k / n)] * p[k];
}

// hydrostatic_pressure (bem_problem.cc:2081-2117): p = -rho g (z - z0)
__global__ void k_ps_hydrostatic(int64_t n, const double *__restrict__ xyz, double rg, double z0,
                                 double *__restrict__ p) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = -rg * (xyz[3 * i + 2] - z0);
}
// hydrodynamic_pressure (bem_problem.cc:2119-2167):
// p = -rho (u (u/2 - W_x) + v (v/2 - W_y) + w (w/2 - W_z))
__global__ void k_ps_hydrodynamic(int64_t n, const double *__restrict__ grad, const double *__restrict__ wind,
                                  double rho, double *__restrict__ p) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0;
  for (int d = 0; d < 3; ++d) {
    const double u = grad[d * n + i];
    s += u * (0.5 * u - wind[d * n + i]);
  }
  p[i] = -rho * s;
}

// Body integrals over the masked real cells (center_of_pressure 2169-2219,
// pressure_force_and_moment 2221-2272, volume_integral 2274-2309):
// per-block partial sums of the 13 quantities, finished by k_ps_finish
// (mask NULL: every cell).
template <int DPC>
__global__ void __launch_bounds__(kPsThreads)
    k_ps_integrals(int64_t ncells, int nq, const double *__restrict__ tab, const double *__restrict__ geom,
                   const int32_t *__restrict__ cdofs, const uint8_t *__restrict__ mask,
                   const double *__restrict__ pres, double cx, double cy, double cz, double *__restrict__ part) {
  __shared__ double sh[kPsThreads / 64];
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double acc[13];
  for (int k = 0; k < 13; ++k) acc[k] = 0.0;
  if (c < ncells && (!mask || mask[c])) {
    const double *G = geom + c * DPC * 3;
    double pj[DPC];
    for (int j = 0; j < DPC; ++j) pj[j] = pres[cdofs[c * DPC + j]];
    QGeo<DPC> g;
    for (int q = 0; q < nq; ++q) {
      g.eval(G, tab + q * (1 + 3 * DPC));
      double pq = 0;
      for (int j = 0; j < DPC; ++j) pq += g.N[j] * pj[j];
      const double r0 = cx - g.x[0], r1 = cy - g.x[1], r2 = cz - g.x[2];
      const double m0 = r1 * g.n[2] - r2 * g.n[1];
      const double m1 = r2 * g.n[0] - r0 * g.n[2];
      const double m2 = r0 * g.n[1] - r1 * g.n[0];
      const double w = pq * g.jxw;
      acc[0] += w * g.n[0];
      acc[1] += w * g.n[1];
      acc[2] += w * g.n[2];
      acc[3] += w * m0;
      acc[4] += w * m1;
      acc[5] += w * m2;
      acc[6] += g.x[0] * w;
      acc[7] += g.x[1] * w;
      acc[8] += g.x[2] * w;
      acc[9] += w;
      acc[10] += -g.x[0] * g.n[0] * g.jxw;
      acc[11] += -g.x[1] * g.n[1] * g.jxw;
      acc[12] += -g.x[2] * g.n[2] * g.jxw;
    }
  }
  for (int k = 0; k < 13; ++k) {
    const double s = ps_block_sum(acc[k], sh);
    if (threadIdx.x == 0) part[(int64_t)k * gridDim.x + blockIdx.x] = s;
  }
}

// free_surface_elevation (bem_problem.cc:2311-2370): at the 1-D Gauss points
// of each listed (cell, line), the point's x and y and p / (rho g).
// tab1 holds per line and point the shape values of the cell's DoFs.
template <int DPC>
__global__ void k_ps_elevation(int64_t nitems, int nq1, const int32_t *__restrict__ faces,
                               const double *__restrict__ tab1, const double *__restrict__ geom,
                               const int32_t *__restrict__ cdofs, const double *__restrict__ pres, double rho_g,
                               double *__restrict__ out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nitems * nq1) return;
  const int64_t it = t / nq1;
  const int q = (int)(t % nq1);
  const int64_t c = faces[2 * it];
  const int f = faces[2 * it + 1];
  const double *N = tab1 + ((int64_t)f * nq1 + q) * DPC;
  const double *G = geom + c * DPC * 3;
  double x = 0, y = 0, p = 0;
  for (int j = 0; j < DPC; ++j) {
    x += N[j] * G[3 * j];
    y += N[j] * G[3 * j + 1];
    p += pres[cdofs[c * DPC + j]] * N[j];
  }
  out[3 * t] = x;
  out[3 * t + 1] = y;
  out[3 * t + 2] = p / rho_g;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
void Engine::ps_setup(int quad_order) {
  if (ps.ready) return;
  HIPCHK(hipSetDevice(device));
  const int D = mesh.dpc;
  const int64_t nc = mesh.ncells_real > 0 ? mesh.ncells_real : mesh.ncells;
  ps.ncells = nc;
  // 2-D Gauss rule of the reference's QGauss(quadrature order), x fastest
  std::vector<double> gx, gw;
  gauss_legendre01(quad_order, gx, gw);
  const int nq = quad_order * quad_order;
  std::vector<double> tab((size_t)nq * (1 + 3 * D));
  for (int b = 0; b < quad_order; ++b)
    for (int a = 0; a < quad_order; ++a) {
      const int q = b * quad_order + a;
      double *t = &tab[(size_t)q * (1 + 3 * D)];
      t[0] = gw[a] * gw[b];
      lagrange_q(D, gx[a], gx[b], t + 1, t + 1 + D, t + 1 + 2 * D);
    }
  ps.nq = nq;
  ps_up(ps.tab, tab);
  // 1-D rule on the four lines of the unit cell (deal.II numbering: line 0
  // x = 0, 1 x = 1, 2 y = 0, 3 y = 1, parameter increasing with the free
  // coordinate): shape values of the cell's DoFs at each point
  std::vector<double> tab1((size_t)4 * quad_order * D);
  for (int f = 0; f < 4; ++f)
    for (int q = 0; q < quad_order; ++q) {
      const double s = gx[q];
      const double xi = f == 0 ? 0.0 : f == 1 ? 1.0 : s;
      const double eta = f == 2 ? 0.0 : f == 3 ? 1.0 : s;
      lagrange_q(D, xi, eta, &tab1[((size_t)f * quad_order + q) * D], nullptr, nullptr);
    }
  ps.nq1 = quad_order;
  ps_up(ps.tab1, tab1);
  // real cells' geometry and DoFs
  ps_up(ps.geom, std::vector<double>(mesh.geom.begin(), mesh.geom.begin() + nc * D * 3));
  ps_up(ps.cdofs, std::vector<int32_t>(mesh.cdofs.begin(), mesh.cdofs.begin() + nc * D));
  ps_up(ps.xyz, mesh.xyz);
  // dof -> (cell, local index), cells ascending
  std::vector<int32_t> dptr(n + 1, 0), dcell((size_t)nc * D);
  std::vector<int8_t> dloc((size_t)nc * D);
  for (int64_t k = 0; k < nc * D; ++k) dptr[mesh.cdofs[k] + 1]++;
  for (int64_t i = 0; i < n; ++i) dptr[i + 1] += dptr[i];
  {
    std::vector<int32_t> fill(dptr.begin(), dptr.end() - 1);
    for (int64_t c = 0; c < nc; ++c)
      for (int j = 0; j < D; ++j) {
        const int32_t d = mesh.cdofs[c * D + j];
        dcell[fill[d]] = (int32_t)c;
        dloc[fill[d]++] = (int8_t)j;
      }
  }
  // scalar mass-matrix pattern: the DoFs sharing a cell
  std::vector<int32_t> mptr(n + 1, 0), mcol;
  {
    std::vector<int32_t> row;
    for (int64_t r = 0; r < n; ++r) {
      row.clear();
      for (int32_t t = dptr[r]; t < dptr[r + 1]; ++t)
        for (int j = 0; j < D; ++j) row.push_back(mesh.cdofs[(int64_t)dcell[t] * D + j]);
      std::sort(row.begin(), row.end());
      row.erase(std::unique(row.begin(), row.end()), row.end());
      mcol.insert(mcol.end(), row.begin(), row.end());
      mptr[r + 1] = (int32_t)mcol.size();
    }
  }
  std::vector<int32_t> pos((size_t)nc * D * D);
  for (int64_t c = 0; c < nc; ++c)
    for (int i = 0; i < D; ++i) {
      const int32_t r = mesh.cdofs[c * D + i];
      for (int j = 0; j < D; ++j) {
        const int32_t col = mesh.cdofs[c * D + j];
        const auto b = mcol.begin() + mptr[r], e = mcol.begin() + mptr[r + 1];
        pos[((size_t)c * D + i) * D + j] = (int32_t)(std::lower_bound(b, e, col) - mcol.begin());
      }
    }
  ps.nnz = (int64_t)mcol.size();
  ps_up(ps.d2c_ptr, dptr);
  ps_up(ps.d2c_cell, dcell);
  ps_up(ps.d2c_loc, dloc);
  ps_up(ps.m_ptr, mptr);
  ps_up(ps.m_col, mcol);
  ps_up(ps.m_pos, pos);
  ps_alloc(ps.m_val, (size_t)ps.nnz * 8);
  ps_alloc(ps.m_diag, (size_t)n * 8);
  ps_alloc(ps.loc, (size_t)nc * std::max(D * D, 3 * D) * 8);
  for (DevBuf *b : {&ps.b, &ps.x, &ps.r, &ps.z, &ps.pv, &ps.qv}) ps_alloc(*b, (size_t)3 * n * 8);
  ps_alloc(ps.cg, kCgLen * 8);
  const int64_t nblk_red = ps_blocks(n, kPsRedChunk);
  const int64_t nblk_cell = ps_blocks(nc, kPsThreads);
  ps_alloc(ps.part, (size_t)std::max<int64_t>(kCgLen * nblk_red, 13 * nblk_cell) * 8);
  // assemble the mass matrix
  const unsigned gc = ps_blocks(nc, kPsThreads), gr = ps_blocks(n, kPsThreads);
  if (D == 4) {
    hipLaunchKernelGGL((k_ps_local<4, 0>), dim3(gc), dim3(kPsThreads), 0, stream, nc, nq, ps.tab.as<double>(),
                       ps.geom.as<double>(), ps.cdofs.as<int32_t>(), nullptr, nullptr, ps.loc.as<double>());
    hipLaunchKernelGGL((k_ps_gather_mass<4>), dim3(gr), dim3(kPsThreads), 0, stream, n, ps.d2c_ptr.as<int32_t>(),
                       ps.d2c_cell.as<int32_t>(), ps.d2c_loc.as<int8_t>(), ps.m_pos.as<int32_t>(),
                       ps.m_ptr.as<int32_t>(), ps.m_col.as<int32_t>(), ps.loc.as<double>(), ps.m_val.as<double>(),
                       ps.m_diag.as<double>());
  } else {
    hipLaunchKernelGGL((k_ps_local<9, 0>), dim3(gc), dim3(kPsThreads), 0, stream, nc, nq, ps.tab.as<double>(),
                       ps.geom.as<double>(), ps.cdofs.as<int32_t>(), nullptr, nullptr, ps.loc.as<double>());
    hipLaunchKernelGGL((k_ps_gather_mass<9>), dim3(gr), dim3(kPsThreads), 0, stream, n, ps.d2c_ptr.as<int32_t>(),
                       ps.d2c_cell.as<int32_t>(), ps.d2c_loc.as<int8_t>(), ps.m_pos.as<int32_t>(),
                       ps.m_ptr.as<int32_t>(), ps.m_col.as<int32_t>(), ps.loc.as<double>(), ps.m_val.as<double>(),
                       ps.m_diag.as<double>());
  }
  HIPCHK(hipGetLastError());
  sync();
  ps.ready = true;
}

template <int D>
static void ps_rhs(Engine &e, int mode, const double *in1, const double *in2) {
  auto &s = e.ps;
  const unsigned gc = ps_blocks(s.ncells, kPsThreads);
  const double *tab = s.tab.as<double>(), *geom = s.geom.as<double>();
  const int32_t *cd = s.cdofs.as<int32_t>();
  double *loc = s.loc.as<double>();
  if (mode == 0)
    hipLaunchKernelGGL((k_ps_local<D, 1>), dim3(gc), dim3(kPsThreads), 0, e.stream, s.ncells, s.nq, tab, geom, cd,
                       in1, in2, loc);
  else if (mode == 1)
    hipLaunchKernelGGL((k_ps_local<D, 2>), dim3(gc), dim3(kPsThreads), 0, e.stream, s.ncells, s.nq, tab, geom, cd,
                       in1, in2, loc);
  else
    hipLaunchKernelGGL((k_ps_local<D, 3>), dim3(gc), dim3(kPsThreads), 0, e.stream, s.ncells, s.nq, tab, geom, cd,
                       in1, in2, loc);
  hipLaunchKernelGGL((k_ps_gather_rhs<D>), dim3(ps_blocks(e.n, kPsThreads)), dim3(kPsThreads), 0, e.stream, e.n,
                     s.d2c_ptr.as<int32_t>(), s.d2c_cell.as<int32_t>(), s.d2c_loc.as<int8_t>(), loc,
                     s.b.as<double>());
}

void Engine::ps_project(int mode, const double *in1, const double *in2, double *out) {
  if (!ps.ready) throw std::runtime_error("post-solve data not set up");
  if (mesh.dpc == 4)
    ps_rhs<4>(*this, mode, in1, in2);
  else
    ps_rhs<9>(*this, mode, in1, in2);
  const int nchunk = (int)ps_blocks(n, kPsRedChunk);
  double *cg = ps.cg.as<double>(), *part = ps.part.as<double>();
  double *x = ps.x.as<double>(), *r = ps.r.as<double>(), *z = ps.z.as<double>(), *p = ps.pv.as<double>(),
         *q = ps.qv.as<double>();
  const double *diag = ps.m_diag.as<double>();
  hipLaunchKernelGGL(k_ps_cg_init, dim3(nchunk), dim3(kPsThreads), 0, stream, n, ps.b.as<double>(), diag, x, r, z,
                     p, part);
  hipLaunchKernelGGL(k_ps_finish, dim3(3), dim3(kPsThreads), 0, stream, nchunk, (int)kRZ, part, cg);
  hipLaunchKernelGGL(k_ps_finish, dim3(3), dim3(kPsThreads), 0, stream, nchunk, (int)kBB, part, cg);
  double h[kCgLen];
  HIPCHK(hipMemcpyAsync(h, cg, sizeof(h), hipMemcpyDeviceToHost, stream));
  sync();
  double bb[3] = {h[kBB], h[kBB + 1], h[kBB + 2]};
  int it = 0;
  bool done = bb[0] == 0.0 && bb[1] == 0.0 && bb[2] == 0.0;
  while (!done) {
    if (it >= kPsMaxIt) throw std::runtime_error("mass-matrix CG did not converge");
    for (int k = 0; k < kPsCheck; ++k, ++it) {
      hipLaunchKernelGGL(k_ps_cg_spmv, dim3(nchunk), dim3(kPsThreads), 0, stream, n, ps.m_ptr.as<int32_t>(),
                         ps.m_col.as<int32_t>(), ps.m_val.as<double>(), p, q, part);
      hipLaunchKernelGGL(k_ps_finish, dim3(3), dim3(kPsThreads), 0, stream, nchunk, (int)kPQ, part, cg);
      hipLaunchKernelGGL(k_ps_alpha, dim3(1), dim3(64), 0, stream, cg);
      hipLaunchKernelGGL(k_ps_cg_update, dim3(nchunk), dim3(kPsThreads), 0, stream, n, cg, diag, p, q, x, r, z,
                         part);
      hipLaunchKernelGGL(k_ps_finish, dim3(3), dim3(kPsThreads), 0, stream, nchunk, (int)kRR, part, cg);
      hipLaunchKernelGGL(k_ps_finish, dim3(3), dim3(kPsThreads), 0, stream, nchunk, (int)kPQ, part, cg);
      hipLaunchKernelGGL(k_ps_beta, dim3(1), dim3(64), 0, stream, cg);
      hipLaunchKernelGGL(k_ps_cg_dir, dim3(ps_blocks(3 * n, kPsThreads)), dim3(kPsThreads), 0, stream, 3 * n, n, cg,
                         z, p);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h, cg, sizeof(h), hipMemcpyDeviceToHost, stream));
    sync();
    done = true;
    for (int d = 0; d < 3; ++d)
      if (h[kRR + d] > kPsTol * kPsTol * bb[d]) done = false;
  }
  ps.last_iters = it;
  HIPCHK(hipMemcpyAsync(out, x, (size_t)3 * n * 8, hipMemcpyDeviceToDevice, stream));
}

void Engine::ps_hydrostatic(double gravity, double density, double z0, double *p) {
  hipLaunchKernelGGL(k_ps_hydrostatic, dim3(ps_blocks(n, kPsThreads)), dim3(kPsThreads), 0, stream, n,
                     ps.xyz.as<double>(), density * gravity, z0, p);
  HIPCHK(hipGetLastError());
}

void Engine::ps_hydrodynamic(double density, const double *wind, const double *grad, double *p) {
  hipLaunchKernelGGL(k_ps_hydrodynamic, dim3(ps_blocks(n, kPsThreads)), dim3(kPsThreads), 0, stream, n, grad, wind,
                     density, p);
  HIPCHK(hipGetLastError());
}

void Engine::ps_integrals(const uint8_t *mask, const double *p, const double cog[3], double out[13]) {
  if (!ps.ready) throw std::runtime_error("post-solve data not set up");
  const unsigned gc = ps_blocks(ps.ncells, kPsThreads);
  double *part = ps.part.as<double>(), *res = ps.cg.as<double>();
  if (mesh.dpc == 4)
    hipLaunchKernelGGL((k_ps_integrals<4>), dim3(gc), dim3(kPsThreads), 0, stream, ps.ncells, ps.nq,
                       ps.tab.as<double>(), ps.geom.as<double>(), ps.cdofs.as<int32_t>(), mask, p, cog[0], cog[1],
                       cog[2], part);
  else
    hipLaunchKernelGGL((k_ps_integrals<9>), dim3(gc), dim3(kPsThreads), 0, stream, ps.ncells, ps.nq,
                       ps.tab.as<double>(), ps.geom.as<double>(), ps.cdofs.as<int32_t>(), mask, p, cog[0], cog[1],
                       cog[2], part);
  // 13 sums into cg[0..12] (the CG scalars are not live between calls)
  hipLaunchKernelGGL(k_ps_finish, dim3(13), dim3(kPsThreads), 0, stream, (int)gc, 0, part, res);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, res, 13 * 8, hipMemcpyDeviceToHost, stream));
  sync();
}

// ---------------------------------------------------------------------------
// The allocation-free body integrals (pbfmm_pressure_force_and_moment_noalloc)
// keep their per-block partial sums in this fixed device array instead of
// the handle's post-solve buffers: 13 sums per workgroup of kPsThreads cells,
// kScratchBlocks workgroups at most (16,777,216 cells, 6.8 MB per GPU, held
// until the process ends).  Each GPU holds its own copy at its own address,
// looked up once per device at set-up (ps_scratch_init); host code passes
// that address, never the array's name, to kernels.
// ---------------------------------------------------------------------------
constexpr int kScratchBlocks = 65536;
constexpr int kScratchSums = 13;
static __device__ __align__(256) double scratch[(size_t)kScratchSums * kScratchBlocks];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};
}  // namespace

int64_t ps_scratch_max_cells() { return (int64_t)kScratchBlocks * kPsThreads; }

// Called by Engine::init_device with the device current.
void ps_scratch_init(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal out of range for the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[dev]) return;
  void *ptr = nullptr;
  HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) throw HipError("hipGetSymbolAddress returned a null address for the scratch array");
  g_scratch[dev] = static_cast<double *>(ptr);
}

// force (out[0..2]) and moment (out[3..5]) over the masked real cells, all
// device pointers, partial sums in the device's scratch array; no allocation.
void Engine::ps_force_moment_noalloc(const uint8_t *mask, const double *p, const double cog[3], double *out) {
  if (!ps.ready) throw std::runtime_error("post-solve data not set up");
  double *part = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    part = (device >= 0 && device < kMaxDevices) ? g_scratch[device] : nullptr;
  }
  if (!part) throw std::runtime_error("scratch array of this device not looked up");
  const unsigned gc = ps_blocks(ps.ncells, kPsThreads);
  if (gc > (unsigned)kScratchBlocks) throw std::runtime_error("more cells than the scratch array holds");
  if (gc == 0) {
    hipLaunchKernelGGL(k_fill_zero6, dim3(1), dim3(64), 0, stream, out);
  } else {
    if (mesh.dpc == 4)
      hipLaunchKernelGGL((k_ps_integrals<4>), dim3(gc), dim3(kPsThreads), 0, stream, ps.ncells, ps.nq,
                         ps.tab.as<double>(), ps.geom.as<double>(), ps.cdofs.as<int32_t>(), mask, p, cog[0], cog[1],
                         cog[2], part);
    else
      hipLaunchKernelGGL((k_ps_integrals<9>), dim3(gc), dim3(kPsThreads), 0, stream, ps.ncells, ps.nq,
                         ps.tab.as<double>(), ps.geom.as<double>(), ps.cdofs.as<int32_t>(), mask, p, cog[0], cog[1],
                         cog[2], part);
    // sums 0..5 (force, moment) straight into out; every partial read was
    // written by this launch (stride gc), so nothing needs clearing
    hipLaunchKernelGGL(k_ps_finish, dim3(6), dim3(kPsThreads), 0, stream, (int)gc, 0, part, out);
  }
  HIPCHK(hipGetLastError());
}

void Engine::ps_elevation(const int32_t *faces, int64_t nfaces, const double *p, double rho_g, double *out) {
  if (!ps.ready) throw std::runtime_error("post-solve data not set up");
  const int64_t nt = nfaces * ps.nq1;
  if (nt == 0) return;
  if (mesh.dpc == 4)
    hipLaunchKernelGGL((k_ps_elevation<4>), dim3(ps_blocks(nt, kPsThreads)), dim3(kPsThreads), 0, stream, nfaces,
                       ps.nq1, faces, ps.tab1.as<double>(), ps.geom.as<double>(), ps.cdofs.as<int32_t>(), p, rho_g,
                       out);
  else
    hipLaunchKernelGGL((k_ps_elevation<9>), dim3(ps_blocks(nt, kPsThreads)), dim3(kPsThreads), 0, stream, nfaces,
                       ps.nq1, faces, ps.tab1.as<double>(), ps.geom.as<double>(), ps.cdofs.as<int32_t>(), p, rho_g,
                       out);
  HIPCHK(hipGetLastError());
}

}  // namespace pbfmm
