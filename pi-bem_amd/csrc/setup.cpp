// Host-side setup of libpbfmm: quadrature, FE values, the octree and its
// interaction lists, the near-field work plan and the far-field plan with its
// translation operators.  Tree semantics follow
// BEMFMA::generate_octree_blocking (source/bem_fma.cc:3614-4801); the data
// layout is flat (sorted index ranges and CSR lists) rather than per-block
// std::set / std::map objects.
#include <algorithm>
#include <atomic>
#include <exception>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstring>
#include <array>
#include <map>
#include <numeric>
#include <stdexcept>

#include "pbfmm_impl.h"

namespace pbfmm {

// ---------------------------------------------------------------------------
// Quadrature rules
// ---------------------------------------------------------------------------
void gauss_legendre01(int n, std::vector<double> &x, std::vector<double> &w) {
  x.resize(n);
  w.resize(n);
  for (int i = 0; i < n; ++i) {
    long double z = std::cos(kPi * (n - i - 0.25) / (n + 0.5));  // ascending roots
    long double d = 0;
    for (int it = 0; it < 100; ++it) {
      long double pm1 = 1, p0 = z;
      for (int k = 2; k <= n; ++k) {
        long double pk = ((2 * k - 1) * z * p0 - (k - 1) * pm1) / k;
        pm1 = p0;
        p0 = pk;
      }
      if (n == 1) { pm1 = 1; p0 = z; }
      d = n * (z * p0 - pm1) / (z * z - 1);
      long double dz = p0 / d;
      z -= dz;
      if (std::fabs((double)dz) < 1e-19) break;
    }
    x[i] = (double)((z + 1) / 2);
    w[i] = (double)(1 / ((1 - z * z) * d * d));
  }
}

// Telles (1987) cubic transformation on [-1,1] towards eta_bar = 2 x0 - 1.
void telles_rule01(int n, double x0, std::vector<double> &x, std::vector<double> &w) {
  std::vector<double> gx, gw;
  gauss_legendre01(n, gx, gw);
  const double eb = 2.0 * x0 - 1.0;
  const double es = eb * eb - 1.0;
  const double gb = std::cbrt(eb * es + std::fabs(es)) + std::cbrt(eb * es - std::fabs(es)) + eb;
  const double den = 1.0 + 3.0 * gb * gb;
  x.clear();
  w.clear();
  for (int i = 0; i < n; ++i) {
    if (std::fabs(gx[i] - x0) <= 1e-10) continue;  // Jacobian vanishes there
    const double g = 2.0 * gx[i] - 1.0, t = g - gb;
    x.push_back(0.5 * ((t * t * t + gb * (gb * gb + 3.0)) / den + 1.0));
    w.push_back(gw[i] * 3.0 * t * t / den);
  }
}

Quadrature2D tensor_rule(const std::vector<double> &x1, const std::vector<double> &w1,
                         const std::vector<double> &x2, const std::vector<double> &w2) {
  Quadrature2D r;
  for (size_t j = 0; j < x2.size(); ++j)
    for (size_t i = 0; i < x1.size(); ++i) {
      r.pts.push_back(x1[i]);
      r.pts.push_back(x2[j]);
      r.w.push_back(w1[i] * w2[j]);
    }
  return r;
}

// QSplit<2>(QDuffy(n, 1), s), the dense assembly's singular rule
// (bem_problem.cc:213-229; deal.II's QSplit and QDuffy, written here from
// Duffy 1982).  The unit square is cut into the triangles (s, v0, v1) over
// its four sides, in deal.II's face order x=0, x=1, y=0, y=1; a side through
// s spans no area and is skipped.  Each triangle takes the Duffy collapse of
// an n x n Gauss rule on [0,1]^2, radial coordinate q0 (fastest) and angular
// q1: reference point (q0 (1 - q1), q0 q1) with weight q0 w0 w1, whose factor
// q0 cancels the 1/r of the Laplace kernels at s, mapped affinely with the
// triangle's |det|.  Evaluated in deal.II's order of operations, so that the
// points and weights match its rule to the last bit.
Quadrature2D duffy_split_rule(int n, double sx, double sy) {
  std::vector<double> g, gw;
  gauss_legendre01(n, g, gw);
  static const double side[4][4] = {{0, 0, 0, 1}, {1, 0, 1, 1}, {0, 0, 1, 0}, {0, 1, 1, 1}};
  Quadrature2D r;
  for (const auto &sd : side) {
    const double a0 = sd[0] - sx, a1 = sd[1] - sy, b0 = sd[2] - sx, b1 = sd[3] - sy;
    const double det = std::fabs(a0 * b1 - b0 * a1);
    if (det < 1e-12) continue;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const double q0 = g[i], q1 = g[j];
        const double xk = q0 * (1.0 - q1), yk = q0 * q1;
        r.pts.push_back(sx + (a0 * xk + b0 * yk));
        r.pts.push_back(sy + (a1 * xk + b1 * yk));
        r.w.push_back(det * (q0 * (gw[i] * gw[j])));
      }
  }
  return r;
}

// FE_Q numbering: vertices (0,0),(1,0),(0,1),(1,1); lines x=0, x=1, y=0, y=1;
// interior.  1-D factors indexed 0 -> node 0, 1 -> node 1, 2 -> midpoint.
static const int kQmap[9][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}, {0, 2}, {1, 2}, {2, 0}, {2, 1}, {2, 2}};

void unit_support_point(int dpc, int j, double &xi, double &eta) {
  static const double c[3] = {0.0, 1.0, 0.5};
  (void)dpc;
  xi = c[kQmap[j][0]];
  eta = c[kQmap[j][1]];
}

void lagrange_q(int dpc, double xi, double eta, double *N, double *dNx, double *dNy) {
  double a[3], b[3], da[3], db[3];
  auto f = [&](double t, double *L, double *dL) {
    if (dpc == 4) {
      L[0] = 1 - t; L[1] = t; L[2] = 0;
      dL[0] = -1; dL[1] = 1; dL[2] = 0;
    } else {
      L[0] = (1 - t) * (1 - 2 * t); L[1] = t * (2 * t - 1); L[2] = 4 * t * (1 - t);
      dL[0] = 4 * t - 3; dL[1] = 4 * t - 1; dL[2] = 4 - 8 * t;
    }
  };
  f(xi, a, da);
  f(eta, b, db);
  for (int j = 0; j < dpc; ++j) {
    const int u = kQmap[j][0], v = kQmap[j][1];
    if (N) N[j] = a[u] * b[v];
    if (dNx) dNx[j] = da[u] * b[v];
    if (dNy) dNy[j] = a[u] * db[v];
  }
}

void compute_quad_data(const HostMesh &m, const Quadrature2D &rule, QuadData &qd) {
  const int nq = (int)rule.w.size(), dpc = m.dpc;
  qd.nq = nq;
  const int64_t NQ = m.ncells * nq;
  qd.x.resize(3 * NQ);
  qd.n.resize(3 * NQ);
  qd.jxw.resize(NQ);
  qd.sh.resize(NQ * dpc);
  std::vector<double> N(nq * dpc), dx(nq * dpc), dy(nq * dpc);
  for (int q = 0; q < nq; ++q)
    lagrange_q(dpc, rule.pts[2 * q], rule.pts[2 * q + 1], &N[q * dpc], &dx[q * dpc], &dy[q * dpc]);
  for (int64_t c = 0; c < m.ncells; ++c) {
    const double *G = &m.geom[c * dpc * 3];
    for (int q = 0; q < nq; ++q) {
      double x[3] = {0, 0, 0}, a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
      for (int j = 0; j < dpc; ++j)
        for (int k = 0; k < 3; ++k) {
          x[k] += N[q * dpc + j] * G[3 * j + k];
          a[k] += dx[q * dpc + j] * G[3 * j + k];
          b[k] += dy[q * dpc + j] * G[3 * j + k];
        }
      const double cr[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
      const double J = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
      const int64_t g = c * nq + q;
      for (int k = 0; k < 3; ++k) {
        qd.x[3 * g + k] = x[k];
        qd.n[3 * g + k] = cr[k] / J;
      }
      qd.jxw[g] = J * rule.w[q];
      for (int j = 0; j < dpc; ++j) qd.sh[g * dpc + j] = N[q * dpc + j];
    }
  }
}

// ---------------------------------------------------------------------------
// Octree (bem_fma.cc:3614-4801)
// ---------------------------------------------------------------------------
static inline int child_index(const double pm[3], double dl, const double *x) {
  if (x[2] <= pm[2] + dl) {
    if (x[1] <= pm[1] + dl) return (x[0] <= pm[0] + dl) ? 0 : 1;
    return (x[0] <= pm[0] + dl) ? 3 : 2;
  }
  if (x[1] <= pm[1] + dl) return (x[0] <= pm[0] + dl) ? 4 : 5;
  return (x[0] <= pm[0] + dl) ? 7 : 6;
}
static const double kOffs[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                   {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// Face/edge/corner contact with the absolute tolerance (bem_fma.cc:4323-4357).
static bool boxes_touch(const Tree &t, int b1, int b2) {
  const double c1 = t.delta[b1] / 2., c2 = t.delta[b2] / 2.;
  double P1[3], Q1[3], P2[3], Q2[3];
  for (int k = 0; k < 3; ++k) {
    P1[k] = t.pmin[3 * b1 + k];
    Q1[k] = 2. * c1 + P1[k];
    P2[k] = t.pmin[3 * b2 + k];
    Q2[k] = 2. * c2 + P2[k];
  }
  auto ov = [&](int k) { return (P1[k] - kToll <= Q2[k]) && (Q1[k] + kToll >= P2[k]); };
  auto fc = [&](int k) { return (std::fabs(P1[k] - Q2[k]) <= kToll) || (std::fabs(Q1[k] - P2[k]) <= kToll); };
  return (fc(0) && ov(1) && ov(2)) || (fc(1) && ov(0) && ov(2)) || (fc(2) && ov(0) && ov(1));
}

static void sort_unique(std::vector<int32_t> &v) {
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
}

void build_tree(const HostMesh &m, const QuadData &qd, int L, int max_nodes, Tree &t) {
  const int64_t N = m.n, NQ = m.ncells * qd.nq;
  t = Tree();
  t.L = L;
  double maxc = 0;
  for (int64_t i = 0; i < 3 * N; ++i) maxc = std::max(maxc, std::fabs(m.xyz[i]));
  // a double body's mirror images must lie in the root box too
  for (int64_t i = 3 * m.ncells_real * m.dpc; i < 3 * m.ncells * m.dpc; ++i) maxc = std::max(maxc, std::fabs(m.geom[i]));
  t.nperm.resize(N);
  std::iota(t.nperm.begin(), t.nperm.end(), 0);
  t.qperm.resize(NQ);
  std::iota(t.qperm.begin(), t.qperm.end(), 0);
  t.qblock.assign((size_t)(L + 1) * NQ, 0);
  t.startL.assign(L + 1, 0);
  t.endL.assign(L + 1, 0);
  t.dofs_filled.assign(L + 1, {});

  auto add_block = [&](int lev, int par, const double pm[3], double dl, int nb0, int nc, int qb0, int qc) {
    const int id = t.nb++;
    t.level.push_back(lev);
    t.parent.push_back(par);
    t.nchild.push_back(0);
    for (int k = 0; k < 8; ++k) t.child.push_back(-1);
    t.octant.push_back(0);
    for (int k = 0; k < 3; ++k) t.pmin.push_back(pm[k]);
    t.delta.push_back(dl);
    t.node_beg.push_back(nb0);
    t.node_cnt.push_back(nc);
    t.quad_beg.push_back(qb0);
    t.quad_cnt.push_back(qc);
    return id;
  };
  double delta = 2.2 * maxc;
  const double rootp[3] = {-1.1 * maxc, -1.1 * maxc, -1.1 * maxc};
  add_block(0, 0, rootp, delta, 0, (int)N, 0, (int)NQ);

  std::vector<int32_t> parents = {0}, tmp;
  std::vector<uint8_t> cls;
  for (int level = 1; level <= L; ++level) {
    delta /= 2.;
    for (int P : parents) {
      const double pm[3] = {t.pmin[3 * P], t.pmin[3 * P + 1], t.pmin[3 * P + 2]};
      int cn[8] = {0}, cq[8] = {0};
      // stable partition of the parent's node and quad ranges into octants
      auto part = [&](std::vector<int32_t> &perm, int beg, int cnt, const double *X, int *counts) {
        cls.resize(cnt);
        for (int i = 0; i < cnt; ++i) {
          cls[i] = (uint8_t)child_index(pm, delta, X + 3 * (int64_t)perm[beg + i]);
          counts[cls[i]]++;
        }
        int off[8];
        off[0] = 0;
        for (int c = 1; c < 8; ++c) off[c] = off[c - 1] + counts[c - 1];
        tmp.resize(cnt);
        for (int i = 0; i < cnt; ++i) tmp[off[cls[i]]++] = perm[beg + i];
        std::copy(tmp.begin(), tmp.begin() + cnt, perm.begin() + beg);
      };
      part(t.nperm, t.node_beg[P], t.node_cnt[P], m.xyz.data(), cn);
      part(t.qperm, t.quad_beg[P], t.quad_cnt[P], qd.x.data(), cq);
      int nb0 = t.node_beg[P], qb0 = t.quad_beg[P];
      for (int c = 0; c < 8; ++c) {
        if (cn[c] + cq[c] > 0) {
          double cp[3];
          for (int k = 0; k < 3; ++k) cp[k] = pm[k] + (kOffs[c][k] != 0 ? delta : 0.);
          const int id = add_block(level, P, cp, delta, nb0, cn[c], qb0, cq[c]);
          t.child[8 * P + t.nchild[P]++] = id;
          t.octant[id] = c;
          for (int i = 0; i < cq[c]; ++i) t.qblock[(size_t)level * NQ + t.qperm[qb0 + i]] = id;
        }
        nb0 += cn[c];
        qb0 += cq[c];
      }
    }
    t.startL[level] = t.endL[level - 1] + 1;
    t.endL[level] = t.nb - 1;
    std::vector<int32_t> next;
    for (int b = t.startL[level]; b <= t.endL[level]; ++b) {
      double bn = 0.0;  // nodes ascending within the range: same summation order as the reference
      for (int i = 0; i < t.node_cnt[b]; ++i) {
        const int node = t.nperm[t.node_beg[b] + i];
        bn += 1.0 / double(m.dbl_ptr[node + 1] - m.dbl_ptr[node]);
      }
      if (std::round(bn) <= max_nodes) {
        t.leaves.push_back(b);
        for (int i = 0; i < t.quad_cnt[b]; ++i)
          for (int k = level + 1; k <= L; ++k) t.qblock[(size_t)k * NQ + t.qperm[t.quad_beg[b] + i]] = b;
      } else {
        next.push_back(b);
      }
      if (bn > 0) t.dofs_filled[level].push_back(b);
    }
    parents.swap(next);
  }
  for (int b : parents) t.leaves.push_back(b);

  // sublevels: childless blocks below level 1 carry lists down to level L
  t.nsub.assign(t.nb, 1);
  for (int b = 0; b < t.nb; ++b)
    if (t.level[b] >= 2 && t.nchild[b] == 0) t.nsub[b] = L - t.level[b] + 1;
  t.sub_beg.assign(t.nb + 1, 0);
  for (int b = 0; b < t.nb; ++b) t.sub_beg[b + 1] = t.sub_beg[b] + t.nsub[b];
  const int ns = t.sub_beg[t.nb];
  t.nn.assign(ns, {});
  t.il.assign(ns, {});
  t.nil.assign(ns, {});

  // near neighbours (bem_fma.cc:4271-4579)
  for (int b = t.startL[1]; b <= t.endL[1] && L >= 1; ++b)
    for (int c = t.startL[1]; c <= t.endL[1]; ++c) t.nn[t.sub_beg[b]].push_back(c);
  // every block of a level needs only its parent's lists: the blocks of a
  // level run on host threads, level after level
  for (int level = 2; level <= L; ++level)
    parallel_for(t.endL[level] - t.startL[level] + 1, 64, [&](int64_t kb) {
      const int b = t.startL[level] + (int)kb;
      std::vector<int32_t> &n0 = t.nn[t.sub_beg[b]];
      n0.push_back(b);
      for (int pos : t.nn[t.sub_beg[t.parent[b]]]) {
        if (t.nchild[pos] == 0 && boxes_touch(t, b, pos)) n0.push_back(pos);
        for (int k = 0; k < t.nchild[pos]; ++k)
          if (boxes_touch(t, b, t.child[8 * pos + k])) n0.push_back(t.child[8 * pos + k]);
      }
      sort_unique(n0);
      for (int s = 1; s < t.nsub[b]; ++s) {
        std::vector<int32_t> &ns_ = t.nn[t.sub_beg[b] + s];
        for (int pos : t.nn[t.sub_beg[b] + s - 1]) {
          if (t.nchild[pos] == 0) ns_.push_back(pos);
          for (int k = 0; k < t.nchild[pos]; ++k)
            if (boxes_touch(t, b, t.child[8 * pos + k])) ns_.push_back(t.child[8 * pos + k]);
        }
        sort_unique(ns_);
      }
    });

  // interaction lists (bem_fma.cc:4610-4711)
  for (int b = t.startL[1]; b <= t.endL[1] && L >= 1; ++b)
    for (int c = t.startL[1]; c <= t.endL[1]; ++c) t.il[t.sub_beg[b]].push_back(c);
  for (int level = 2; level <= L; ++level)
    parallel_for(t.endL[level] - t.startL[level] + 1, 64, [&](int64_t kb) {
      const int b = t.startL[level] + (int)kb;
      // cells touching any double of any node of b
      std::vector<int32_t> cells;
      for (int i = 0; i < t.node_cnt[b]; ++i) {
        const int node = t.nperm[t.node_beg[b] + i];
        for (int k = m.dbl_ptr[node]; k < m.dbl_ptr[node + 1]; ++k) {
          const int d = m.dbl_idx[k];
          for (int c = m.d2c_ptr[d]; c < m.d2c_ptr[d + 1]; ++c) cells.push_back(m.d2c_idx[c]);
        }
      }
      sort_unique(cells);
      for (int s = 0; s < t.nsub[b]; ++s) {
        std::vector<int32_t> &il = t.il[t.sub_beg[b] + s];
        il = t.nn[t.sub_beg[b] + s];
        const size_t lev = level + s;
        for (int c : cells)
          for (int q = 0; q < qd.nq; ++q) il.push_back(t.qblock[lev * NQ + (int64_t)c * qd.nq + q]);
        sort_unique(il);
      }
      for (int s = 0; s < t.nsub[b]; ++s) {
        const std::vector<int32_t> &il = t.il[t.sub_beg[b] + s];
        const std::vector<int32_t> &pil = (s == 0) ? t.il[t.sub_beg[t.parent[b]]] : t.il[t.sub_beg[b] + s - 1];
        std::vector<int32_t> &nl = t.nil[t.sub_beg[b] + s];
        auto in_il = [&](int x) { return std::binary_search(il.begin(), il.end(), x); };
        for (int pos : pil) {
          if (t.nchild[pos] == 0) {
            if (!in_il(pos)) nl.push_back(pos);
          } else {
            for (int k = 0; k < t.nchild[pos]; ++k)
              if (!in_il(t.child[8 * pos + k])) nl.push_back(t.child[8 * pos + k]);
          }
        }
        sort_unique(nl);
      }
    });
}

// ---------------------------------------------------------------------------
// Near-field plan (direct_integrals, bem_fma.cc:107-1600)
// ---------------------------------------------------------------------------
void parallel_for(int64_t n, int64_t grain, const std::function<void(int64_t)> &body) {
  if (n <= 0) return;
  grain = std::max<int64_t>(grain, 1);
  const int nth = (int)std::min<int64_t>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())),
                                         (n + grain - 1) / grain);
  std::atomic<int64_t> next{0};
  std::exception_ptr err;
  std::mutex mu;
  auto work = [&]() {
    try {
      for (;;) {
        const int64_t b = next.fetch_add(grain);
        if (b >= n) break;
        for (int64_t i = b; i < std::min(n, b + grain); ++i) body(i);
      }
    } catch (...) {
      std::lock_guard<std::mutex> lk(mu);
      if (!err) err = std::current_exception();
      next.store(n);
    }
  };
  std::vector<std::thread> th;
  for (int k = 1; k < nth; ++k) th.emplace_back(work);
  work();
  for (auto &x : th) x.join();
  if (err) std::rethrow_exception(err);
}

// The near field of one leaf: its entries (cell, quad-point mask, near
// flag), the column of each entry's DoFs and its sorted rows and columns.
namespace {
struct LeafNear {
  std::vector<int32_t> ent_cell, ent_colpos, rows, cols;
  std::vector<uint64_t> ent_mask;
  std::vector<uint8_t> ent_near;
};

void leaf_near(const HostMesh &m, const QuadData &qd, const Tree &t, int leaf, LeafNear &o) {
  const int dpc = m.dpc;
  std::vector<int32_t> qs;
  auto add_entries = [&](std::vector<int32_t> &q, uint8_t near) {
    sort_unique(q);
    size_t i = 0;
    while (i < q.size()) {
      const int cell = q[i] / qd.nq;
      uint64_t mask = 0;
      while (i < q.size() && q[i] / qd.nq == cell) mask |= (uint64_t)1 << (q[i++] % qd.nq);
      o.ent_cell.push_back(cell);
      o.ent_mask.push_back(mask);
      o.ent_near.push_back(near && cell < m.ncells_real);  // mirror images are never singular
    }
  };
  auto quads_of = [&](int b, std::vector<int32_t> &q) {
    for (int i = 0; i < t.quad_cnt[b]; ++i) q.push_back(t.qperm[t.quad_beg[b] + i]);
  };
  // near quad points: blocks of the deepest intList
  for (int b : t.il[t.sub_beg[leaf] + t.nsub[leaf] - 1]) quads_of(b, qs);
  add_entries(qs, 1);
  // bigger far blocks in the nonIntLists of the leaf and of every ancestor
  for (int a = leaf; a != 0; a = t.parent[a]) {
    const int lev = t.level[a];
    qs.clear();
    for (int s = 0; s < t.nsub[a]; ++s)
      for (int b : t.nil[t.sub_beg[a] + s])
        if (b < t.startL[lev]) quads_of(b, qs);
    if (!qs.empty()) add_entries(qs, 0);
  }
  const size_t ne = o.ent_cell.size();
  for (size_t e = 0; e < ne; ++e)
    for (int j = 0; j < dpc; ++j) o.cols.push_back(m.cdofs[(int64_t)o.ent_cell[e] * dpc + j]);
  sort_unique(o.cols);
  o.ent_colpos.reserve(ne * dpc);
  for (size_t e = 0; e < ne; ++e)
    for (int j = 0; j < dpc; ++j) {
      const int d = m.cdofs[(int64_t)o.ent_cell[e] * dpc + j];
      o.ent_colpos.push_back((int32_t)(std::lower_bound(o.cols.begin(), o.cols.end(), d) - o.cols.begin()));
    }
  o.rows.assign(t.nperm.begin() + t.node_beg[leaf], t.nperm.begin() + t.node_beg[leaf] + t.node_cnt[leaf]);
  std::sort(o.rows.begin(), o.rows.end());
}
}  // namespace

// Leaves are independent: each is built on a host thread, then they are
// concatenated in leaf order, so the plan does not depend on the threads.
void build_near_plan(const HostMesh &m, const QuadData &qd, const Tree &t, NearPlan &np) {
  np = NearPlan();
  if (qd.nq > 64) throw std::runtime_error("quadrature order above 8 is not supported");
  std::vector<int> leaves;
  for (int leaf : t.leaves)
    if (t.node_cnt[leaf] > 0) leaves.push_back(leaf);
  std::vector<LeafNear> part(leaves.size());
  parallel_for((int64_t)leaves.size(), 64, [&](int64_t k) { leaf_near(m, qd, t, leaves[k], part[k]); });
  size_t ne = 0, nr = 0, nc = 0;
  for (const auto &o : part) {
    ne += o.ent_cell.size();
    nr += o.rows.size();
    nc += o.cols.size();
  }
  np.ent_cell.reserve(ne);
  np.ent_mask.reserve(ne);
  np.ent_near.reserve(ne);
  np.ent_colpos.reserve(ne * m.dpc);
  np.rows.reserve(nr);
  np.row_leaf.reserve(nr);
  np.cols.reserve(nc);
  np.row_beg.push_back(0);
  np.col_beg.push_back(0);
  np.ent_beg.push_back(0);
  np.work_beg.push_back(0);
  for (size_t k = 0; k < leaves.size(); ++k) {
    LeafNear &o = part[k];
    const int64_t e0 = (int64_t)np.ent_cell.size();
    np.ent_cell.insert(np.ent_cell.end(), o.ent_cell.begin(), o.ent_cell.end());
    np.ent_mask.insert(np.ent_mask.end(), o.ent_mask.begin(), o.ent_mask.end());
    np.ent_near.insert(np.ent_near.end(), o.ent_near.begin(), o.ent_near.end());
    np.ent_colpos.insert(np.ent_colpos.end(), o.ent_colpos.begin(), o.ent_colpos.end());
    const int64_t e1 = (int64_t)np.ent_cell.size();
    np.val_beg.push_back(np.nnz);
    np.nnz += (int64_t)o.rows.size() * (int64_t)o.cols.size();
    for (int r : o.rows) {
      np.rows.push_back(r);
      np.row_leaf.push_back(np.nleaf);
    }
    np.cols.insert(np.cols.end(), o.cols.begin(), o.cols.end());
    np.leaf_block.push_back(leaves[k]);
    np.row_beg.push_back((int32_t)np.rows.size());
    np.col_beg.push_back((int32_t)np.cols.size());
    np.ent_beg.push_back((int32_t)e1);
    np.work_beg.push_back(np.work_beg.back() + (int64_t)o.rows.size() * (e1 - e0));
    np.nleaf++;
    o = LeafNear();  // free as we go
  }
}

// ---------------------------------------------------------------------------
// Translation operators as real (2Kt x 2Kt) maps on [Re M, Im M]
// ---------------------------------------------------------------------------
namespace {
struct Leg {
  int nmax;
  std::vector<double> P;
  Leg(int nm, double x) : nmax(nm), P((nm + 1) * (nm + 1), 0.0) {
    const int ld = nm + 1;
    const double s = std::sqrt(std::max(0.0, 1.0 - x * x));
    double pmm = 1.0;
    for (int mm = 0; mm <= nm; ++mm) {
      if (mm > 0) pmm *= -std::sqrt((2.0 * mm - 1.0) / (2.0 * mm)) * s;
      P[mm * ld + mm] = pmm;
      if (mm + 1 <= nm) P[(mm + 1) * ld + mm] = x * std::sqrt(2.0 * mm + 1.0) * pmm;
      for (int n = mm + 2; n <= nm; ++n)
        P[n * ld + mm] = ((2.0 * n - 1.0) * x * P[(n - 1) * ld + mm] -
                          std::sqrt(double((n - 1) * (n - 1) - mm * mm)) * P[(n - 2) * ld + mm]) /
                         std::sqrt(double(n * n - mm * mm));
    }
  }
  double operator()(int n, int m) const { return P[n * (nmax + 1) + m]; }
};
// A_n^m = (-1)^n / sqrt((n-m)!(n+m)!)  (multipole_expansion.h:123-147)
double Anm(int n, int m) {
  m = std::abs(m);
  double f1 = 1, f2 = 1;
  for (int i = n - m; i > 0; --i) f1 *= i;
  for (int i = n + m; i > 0; --i) f2 *= i;
  return ((n & 1) ? -1.0 : 1.0) / std::sqrt(f1 * f2);
}
double ire(int k) {  // Re(i^k), k even
  const int r = ((k % 4) + 4) % 4;
  return r == 0 ? 1.0 : (r == 2 ? -1.0 : 0.0);
}
inline int tri(int n, int m) { return n * (n + 1) / 2 + m; }
inline int sgn(int x) { return (x > 0) - (x < 0); }
// out(n,m) += w * a~ where a~ = (x, sig*y) of input (nn, mi): adds the four real
// entries of the R-linear map into the transposed operator opT[col*kp + row].
inline void put(double *opT, int kt, int kp, int out, int in, double wr, double wi, int sig) {
  opT[(size_t)in * kp + out] += wr;
  opT[(size_t)(kt + in) * kp + out] += -sig * wi;
  opT[(size_t)in * kp + kt + out] += wi;
  opT[(size_t)(kt + in) * kp + kt + out] += sig * wr;
}
void sph(const double d[3], double &rho, double &ca, double &beta) {
  rho = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  ca = d[2] / rho;
  beta = std::atan2(d[1], d[0]);
}
}  // namespace

// MultipoleExpansion::Add(child) (multipole_expansion.cc:184-261)
void op_m2m(int p, const double d[3], double *opT) {
  const int kt = kt_of(p), kp = kp_of(p);
  std::fill(opT, opT + (size_t)kp * kp, 0.0);
  double rho, ca, beta;
  sph(d, rho, ca, beta);
  Leg P(p, ca);
  for (int n = 0; n <= p; ++n)
    for (int m = 0; m <= n; ++m)
      for (int nn = 0; nn <= n; ++nn)
        for (int mm = -nn; mm <= nn; ++mm) {
          if (std::abs(m - mm) > n - nn) continue;
          double rf = P(nn, std::abs(mm)) * std::pow(rho, double(nn)) * Anm(nn, mm) * Anm(n - nn, m - mm) / Anm(n, m);
          rf *= ire(std::abs(m) - std::abs(mm) - std::abs(m - mm));
          put(opT, kt, kp, tri(n, m), tri(n - nn, std::abs(m - mm)), rf * std::cos(mm * beta),
              -rf * std::sin(mm * beta), sgn(m - mm));
        }
}

// LocalExpansion::Add(parent) (local_expansion.cc:107-176, table local_expansion.h:182-215)
void op_l2l(int p, const double d[3], double *opT) {
  const int kt = kt_of(p), kp = kp_of(p);
  std::fill(opT, opT + (size_t)kp * kp, 0.0);
  double rho, ca, beta;
  sph(d, rho, ca, beta);
  Leg P(p, ca);
  for (int n = 0; n <= p; ++n)
    for (int m = 0; m <= n; ++m)
      for (int nn = n; nn <= p; ++nn) {
        const double rf0 = std::pow(rho, double(nn - n));
        for (int mm = -nn; mm <= nn; ++mm) {
          if (std::abs(mm - m) > nn - n) continue;
          double c = Anm(nn - n, mm - m) / Anm(nn, mm) * Anm(n, m) * ire(std::abs(mm) - std::abs(mm - m) - std::abs(m)) *
                     (((nn + n) & 1) ? -1.0 : 1.0);
          const double rf = P(nn - n, std::abs(mm - m)) * rf0 * c;
          put(opT, kt, kp, tri(n, m), tri(nn, std::abs(mm)), rf * std::cos((mm - m) * beta),
              rf * std::sin((mm - m) * beta), sgn(mm));
        }
      }
}

// LocalExpansion::Add(multipole) (local_expansion.cc:179-248, table local_expansion.h:145-180)
void op_m2l(int p, const double d[3], double *opT) {
  const int kt = kt_of(p), kp = kp_of(p);
  std::fill(opT, opT + (size_t)kp * kp, 0.0);
  double rho, ca, beta;
  sph(d, rho, ca, beta);
  Leg P(2 * p, ca);
  for (int n = 0; n <= p; ++n)
    for (int m = 0; m <= n; ++m)
      for (int nn = 0; nn <= p; ++nn) {
        const double rf0 = std::pow(rho, double(-n - nn - 1));
        for (int mm = -nn; mm <= nn; ++mm) {
          double c = Anm(nn, mm) / Anm(n + nn, m - mm) * Anm(n, m) * ire(std::abs(m - mm) - std::abs(m) - std::abs(mm)) /
                     ((nn & 1) ? -1.0 : 1.0);
          const double rf = P(n + nn, std::abs(mm - m)) * rf0 * c;
          put(opT, kt, kp, tri(n, m), tri(nn, std::abs(mm)), rf * std::cos((mm - m) * beta),
              rf * std::sin((mm - m) * beta), mm < 0 ? -1 : 1);
        }
      }
}

// ---------------------------------------------------------------------------
// Far-field plan (generate_multipole_expansions bem_fma.cc:1970-2380 and
// multipole_matr_vect_products bem_fma.cc:2390-3072)
// ---------------------------------------------------------------------------
static void add_class(ChunkList &cl, int op, const std::vector<std::pair<int, int>> &pairs) {
  for (size_t i = 0; i < pairs.size(); i += kM2LChunk) {
    cl.op.push_back(op);
    cl.beg.push_back((int)cl.tgt.size() + (int)i);
    cl.cnt.push_back((int)std::min<size_t>(kM2LChunk, pairs.size() - i));
  }
  for (size_t i = 0; i < pairs.size(); i += kChunk1) {
    cl.op1.push_back(op);
    cl.beg1.push_back((int)cl.tgt.size() + (int)i);
    cl.cnt1.push_back((int)std::min<size_t>(kChunk1, pairs.size() - i));
  }
  for (auto &pr : pairs) {
    cl.tgt.push_back(pr.first);
    cl.src.push_back(pr.second);
  }
}

// Degree n of each row of a real expansion vector [Re(0..Kt-1), Im(0..Kt-1),
// padding] (index n(n+1)/2 + m); padding rows get degree 0.
static std::vector<int> row_degrees(int p, int kp) {
  const int kt = kt_of(p);
  std::vector<int> deg(kp, 0);
  for (int n = 0; n <= p; ++n)
    for (int m = 0; m <= n; ++m) {
      deg[n * (n + 1) / 2 + m] = n;
      deg[kt + n * (n + 1) / 2 + m] = n;
    }
  return deg;
}

// opT[k*kp + row] *= in_scale^deg(k) * out_scale^deg(row)
static void scale_op(double *opT, int kp, const std::vector<int> &deg, double in_scale, double out_scale) {
  for (int k = 0; k < kp; ++k)
    for (int r = 0; r < kp; ++r) opT[(size_t)k * kp + r] *= std::pow(in_scale, deg[k]) * std::pow(out_scale, deg[r]);
}

void build_far_plan(const Tree &t, int p, FarPlan &fp) {
  fp = FarPlan();
  fp.p = p;
  fp.kt = kt_of(p);
  fp.kp = kp_of(p);
  const int L = t.L, kp = fp.kp;
  const std::vector<int> deg = row_degrees(p, kp);
  fp.p2m_leaf = t.leaves;
  for (int b : t.leaves)
    if (t.node_cnt[b] > 0) fp.l2p_leaf.push_back(b);

  // M2M / L2L operators per level and octant
  fp.m2m_ops.assign((size_t)(L + 1) * 8 * kp * kp, 0.0);
  fp.l2l_ops.assign((size_t)(L + 1) * 8 * kp * kp, 0.0);
  for (int l = 1; l <= L; ++l) {
    const double dl = t.delta[0] / std::pow(2.0, l);
    for (int o = 0; o < 8; ++o) {
      double d[3];
      for (int k = 0; k < 3; ++k) d[k] = (kOffs[o][k] != 0 ? 0.5 : -0.5) * dl;
      op_m2m(p, d, &fp.m2m_ops[((size_t)l * 8 + o) * kp * kp]);
      const double dn[3] = {-d[0], -d[1], -d[2]};
      op_l2l(p, dn, &fp.l2l_ops[((size_t)l * 8 + o) * kp * kp]);
      // level-scaled expansions: child ME/delta_l^n -> parent ME/delta_{l-1}^n,
      // parent LE.delta_{l-1}^(n+1) -> child LE.delta_l^(n+1)
      scale_op(&fp.m2m_ops[((size_t)l * 8 + o) * kp * kp], kp, deg, dl, 1.0 / (2.0 * dl));
      scale_op(&fp.l2l_ops[((size_t)l * 8 + o) * kp * kp], kp, deg, 1.0 / (2.0 * dl), dl);
      // the extra power of the box size: child delta / parent delta
      for (size_t i = 0; i < (size_t)kp * kp; ++i) fp.l2l_ops[((size_t)l * 8 + o) * kp * kp + i] *= 0.5;
    }
  }
  // M2M pairs (bem_fma.cc:2318-2339: every block of level l into its parent)
  // and L2L pairs (bem_fma.cc:2600-2602: blocks of dofs_filled[l] from their
  // parent; level 1 takes the root's local expansion, which is zero)
  fp.m2m.level_off.assign(L + 2, 0);
  fp.l2l.level_off.assign(L + 2, 0);
  fp.m2m.level_off1.assign(L + 2, 0);
  fp.l2l.level_off1.assign(L + 2, 0);
  for (int l = 0; l <= L; ++l) {
    fp.m2m.level_off[l] = (int)fp.m2m.op.size();
    fp.l2l.level_off[l] = (int)fp.l2l.op.size();
    fp.m2m.level_off1[l] = (int)fp.m2m.op1.size();
    fp.l2l.level_off1[l] = (int)fp.l2l.op1.size();
    if (l == 0) continue;
    std::vector<std::pair<int, int>> byoct[8];
    for (int b = t.startL[l]; b <= t.endL[l]; ++b) byoct[t.octant[b]].push_back({t.parent[b], b});
    for (int o = 0; o < 8; ++o) add_class(fp.m2m, l * 8 + o, byoct[o]);
    if (l >= 2) {
      std::vector<std::pair<int, int>> lo[8];
      for (int b : t.dofs_filled[l]) lo[t.octant[b]].push_back({b, t.parent[b]});
      for (int o = 0; o < 8; ++o) add_class(fp.l2l, l * 8 + o, lo[o]);
    }
  }
  fp.m2m.level_off[L + 1] = (int)fp.m2m.op.size();
  fp.l2l.level_off[L + 1] = (int)fp.l2l.op.size();
  fp.m2m.level_off1[L + 1] = (int)fp.m2m.op1.size();
  fp.l2l.level_off1[L + 1] = (int)fp.l2l.op1.size();

  // M2L pairs by (level, offset class); M2P pairs.  With ME_n / delta^n and
  // LE_n delta^(n+1) the M2L operator of an offset class is the same at every
  // level (the unit-box one: the Laplace M2L between boxes of size delta
  // scales as delta^-(n+nn+1)), so the classes of all levels share one
  // operator per offset.  Chunks are per (level, class), level after level
  // (a workgroup range then gathers sources of one level), unless merging the
  // levels saves many partial chunks (below).
  fp.m2l_per_level.assign(L + 1, 0);
  fp.m2l.level_off.assign(L + 2, 0);
  fp.m2l.level_off1.assign(L + 2, 0);
  fp.m2p_work_beg.push_back(0);
  std::map<std::array<int, 3>, int> op_of;
  for (int l = 1; l <= L; ++l) {
    fp.m2l.level_off[l] = (int)fp.m2l.op.size();
    fp.m2l.level_off1[l] = (int)fp.m2l.op1.size();
    std::map<std::array<int, 3>, std::vector<std::pair<int, int>>> cls;
    const double dl = t.delta[0] / std::pow(2.0, l);
    for (int b : t.dofs_filled[l]) {
      double cb[3];
      t.center(b, cb);
      for (int s = 0; s < t.nsub[b]; ++s)
        for (int src : t.nil[t.sub_beg[b] + s]) {
          if (src >= t.startL[l] && src <= t.endL[l]) {
            double cs[3];
            t.center(src, cs);
            std::array<int, 3> key;
            for (int k = 0; k < 3; ++k) key[k] = (int)std::lround((cs[k] - cb[k]) / dl);
            cls[key].push_back({b, src});
          } else if (src > t.endL[l]) {
            fp.m2p_tgt.push_back(b);
            fp.m2p_src.push_back(src);
            fp.m2p_work_beg.push_back(fp.m2p_work_beg.back() + t.node_cnt[b]);
            for (int i = 0; i < t.node_cnt[b]; ++i) {
              fp.m2p_item_node.push_back(t.nperm[t.node_beg[b] + i]);
              fp.m2p_item_src.push_back(src);
            }
          }
        }
    }
    for (auto &kv : cls) {
      auto it = op_of.find(kv.first);
      if (it == op_of.end()) {
        const int op = fp.n_m2l_ops++;
        it = op_of.emplace(kv.first, op).first;
      }
      add_class(fp.m2l, it->second, kv.second);
      fp.m2l_per_level[l] += (int64_t)kv.second.size();
    }
  }
  // the operators of the offset classes (unit box), on host threads
  {
    std::vector<std::array<int, 3>> key(fp.n_m2l_ops);
    for (auto &kv : op_of) key[kv.second] = kv.first;
    fp.m2l_ops.assign((size_t)fp.n_m2l_ops * kp * kp, 0.0);
    parallel_for(fp.n_m2l_ops, 1, [&](int64_t op) {
      const double d[3] = {(double)key[op][0], (double)key[op][1], (double)key[op][2]};
      op_m2l(p, d, &fp.m2l_ops[(size_t)op * kp * kp]);
    });
  }
  fp.m2l.level_off[L + 1] = (int)fp.m2l.op.size();
  fp.m2l.level_off1[L + 1] = (int)fp.m2l.op1.size();
  // Where partial chunks per (level, class) are many (small or ragged trees),
  // the classes' pairs of all levels are cut into chunks together instead,
  // listed under level 1: that saves the partial chunks' launches of
  // workgroups.  Where it saves little, the level-major chunks keep each
  // workgroup range's sources on one level (measured: 393,218 DoFs, 0.5%
  // fewer chunks, FP32 M2L 3% slower merged; 64 problems of 1.5k-25k DoFs,
  // solves 4% faster merged; 24,578 DoFs, a fifth fewer chunks, no change).
  {
    std::vector<std::vector<std::pair<int, int>>> by_op(fp.n_m2l_ops);
    for (size_t k = 0; k < fp.m2l.op1.size(); ++k)
      for (int i = fp.m2l.beg1[k]; i < fp.m2l.beg1[k] + fp.m2l.cnt1[k]; ++i)
        by_op[fp.m2l.op1[k]].push_back({fp.m2l.tgt[i], fp.m2l.src[i]});
    size_t merged = 0;
    for (auto &v : by_op) merged += (v.size() + kChunk1 - 1) / kChunk1;
    if (merged < 0.9 * (double)fp.m2l.op1.size()) {
      ChunkList cl;
      cl.level_off.assign(L + 2, 0);
      cl.level_off1.assign(L + 2, 0);
      for (int op = 0; op < fp.n_m2l_ops; ++op) add_class(cl, op, by_op[op]);
      for (int l = 2; l <= L + 1; ++l) {
        cl.level_off[l] = (int)cl.op.size();
        cl.level_off1[l] = (int)cl.op1.size();
      }
      fp.m2l = std::move(cl);
    }
  }
  build_m2m_groups(t, fp);
  build_l2l_groups(t, fp);
}

// C = A B for column-major kp x kp operators (C[k kp + row] = sum_j
// A[j kp + row] B[k kp + j]): B acts first.
static void compose_ops(const double *A, const double *B, int kp, double *C) {
  std::fill(C, C + (size_t)kp * kp, 0.0);
  for (int k = 0; k < kp; ++k)
    for (int j = 0; j < kp; ++j) {
      const double b = B[(size_t)k * kp + j];
      if (b == 0.0) continue;
      for (int row = 0; row < kp; ++row) C[(size_t)k * kp + row] += A[(size_t)j * kp + row] * b;
    }
}

// The upward pass in half the launches: M2M child levels a and a - 1 in one
// launch.  The targets at level a - 1 take their children at level a as
// usual; the targets at level a - 2 take their leaf children at level a - 1
// directly and, for every other child c, c's children at level a through
// the composite operator T_{a-1}(octant c) T_a(octant gc), so that no pair
// of the launch reads an expansion the launch writes.  Child levels run from
// L down to the coarsest one whose parent's expansions something reads (as
// Engine::consumer_levels); group g is m2m_g chunks level_off1[g] ..
// level_off1[g + 1].  The composite operators are appended to m2m_ops.
void build_m2m_groups(const Tree &t, FarPlan &fp) {
  const int L = t.L, kp = fp.kp;
  fp.m2m_g = ChunkList();
  fp.m2m_g.level_off1.assign(1, 0);
  fp.m2m_g.level_off.assign(1, 0);
  int me_lo = L + 1;
  for (int l = 1; l <= L; ++l)
    if (fp.m2l_per_level[l] > 0) me_lo = std::min(me_lo, l);
  for (int b : fp.m2p_src) me_lo = std::min(me_lo, (int)t.level[b]);
  const size_t base = fp.m2m_ops.size() / ((size_t)kp * kp);
  std::vector<double> comp;  // appended composite operators
  int ncomp = 0;
  for (int a = L; a > me_lo; a -= 2) {
    std::map<int, std::vector<std::pair<int, int>>> cls;
    for (int b = t.startL[a]; b <= t.endL[a]; ++b) cls[a * 8 + t.octant[b]].push_back({t.parent[b], b});
    if (a - 1 > me_lo) {
      // composite operators of this level pair, on first use
      std::vector<int> cop(64, -1);
      for (int c = t.startL[a - 1]; c <= t.endL[a - 1]; ++c) {
        if (t.nchild[c] == 0) {
          cls[(a - 1) * 8 + t.octant[c]].push_back({t.parent[c], c});
          continue;
        }
        for (int q = 0; q < t.nchild[c]; ++q) {
          const int gc = t.child[8 * c + q];
          const int key = t.octant[c] * 8 + t.octant[gc];
          if (cop[key] < 0) {
            cop[key] = ncomp++;
            comp.resize(comp.size() + (size_t)kp * kp);
            compose_ops(&fp.m2m_ops[((size_t)(a - 1) * 8 + t.octant[c]) * kp * kp],
                        &fp.m2m_ops[((size_t)a * 8 + t.octant[gc]) * kp * kp], kp,
                        &comp[comp.size() - (size_t)kp * kp]);
          }
          cls[-1 - (int)(base + cop[key])].push_back({t.parent[c], gc});
        }
      }
    }
    for (auto &kv : cls) add_class(fp.m2m_g, kv.first >= 0 ? kv.first : -1 - kv.first, kv.second);
    fp.m2m_g.level_off1.push_back((int)fp.m2m_g.op1.size());
    fp.m2m_g.level_off.push_back((int)fp.m2m_g.op.size());
  }
  fp.m2m_ops.insert(fp.m2m_ops.end(), comp.begin(), comp.end());
}

// The downward pass in half the launches: L2L into target levels l and
// l + 1 in one launch.  A target of level l + 1 takes its parent b's LE and,
// through the composite T_{l+1}(octant gc) T_l(octant b), its grandparent's;
// b's LE then holds its M2L part only, so b itself (a non-leaf of level l)
// takes no L2L: nothing reads its LE afterwards (L2P reads leaves).  Leaves
// of level l take their parent's LE as usual.  No pair reads an expansion
// the launch writes: the reads are of level l - 1 (final) and of non-leaf
// blocks of level l.  Target levels run from max(2, le_lo + 1) (as
// matvec_fused, le_lo the coarsest M2L level) to L; group g is l2l_g chunks
// level_off1[g] .. level_off1[g + 1]; the composites are appended to l2l_ops.
void build_l2l_groups(const Tree &t, FarPlan &fp) {
  const int L = t.L, kp = fp.kp;
  fp.l2l_g = ChunkList();
  fp.l2l_g.level_off1.assign(1, 0);
  fp.l2l_g.level_off.assign(1, 0);
  int le_lo = L + 1;
  for (int l = 1; l <= L; ++l)
    if (fp.m2l_per_level[l] > 0) le_lo = std::min(le_lo, l);
  const size_t base = fp.l2l_ops.size() / ((size_t)kp * kp);
  std::vector<char> filled(t.level.size(), 0);
  for (int l = 0; l <= L; ++l)
    for (int b : t.dofs_filled[l]) filled[b] = 1;
  std::vector<double> comp;
  int ncomp = 0;
  for (int l = std::max(2, le_lo + 1); l <= L; l += 2) {
    std::map<int, std::vector<std::pair<int, int>>> cls;
    const bool pair = l + 1 <= L;
    std::vector<int> cop(64, -1);
    for (int b : t.dofs_filled[l]) {
      if (!pair || t.nchild[b] == 0) {
        cls[l * 8 + t.octant[b]].push_back({b, t.parent[b]});
        continue;
      }
      for (int q = 0; q < t.nchild[b]; ++q) {
        const int gc = t.child[8 * b + q];
        if (!filled[gc]) continue;
        cls[(l + 1) * 8 + t.octant[gc]].push_back({gc, b});
        const int key = t.octant[b] * 8 + t.octant[gc];
        if (cop[key] < 0) {
          cop[key] = ncomp++;
          comp.resize(comp.size() + (size_t)kp * kp);
          compose_ops(&fp.l2l_ops[((size_t)(l + 1) * 8 + t.octant[gc]) * kp * kp],
                      &fp.l2l_ops[((size_t)l * 8 + t.octant[b]) * kp * kp], kp, &comp[comp.size() - (size_t)kp * kp]);
        }
        cls[-1 - (int)(base + cop[key])].push_back({gc, t.parent[b]});
      }
    }
    for (auto &kv : cls) add_class(fp.l2l_g, kv.first >= 0 ? kv.first : -1 - kv.first, kv.second);
    fp.l2l_g.level_off1.push_back((int)fp.l2l_g.op1.size());
    fp.l2l_g.level_off.push_back((int)fp.l2l_g.op.size());
  }
  fp.l2l_ops.insert(fp.l2l_ops.end(), comp.begin(), comp.end());
}

// The image of cell c is the reflection x -> (x, y, 2 z0 - z) of its mapping
// nodes with the two reference directions swapped (local node j of the image
// is node perm[j] of c): the quadrature points are the reflected ones, the
// shape function of each DoF is the same at each of them, JxW is unchanged and
// x_xi cross x_eta of the image is the reflected normal (n_x, n_y, -n_z), which
// is what the reference integrates (bem_problem.cc:1043-1062).
void add_mirror_cells(HostMesh &m, double z0) {
  static const int perm_q1[4] = {0, 2, 1, 3};
  static const int perm_q2[9] = {0, 2, 1, 3, 6, 7, 4, 5, 8};
  const int *perm = m.dpc == 4 ? perm_q1 : perm_q2;
  const int64_t nr = m.ncells_real, dpc = m.dpc;
  m.cdofs.resize(2 * nr * dpc);
  m.geom.resize(2 * nr * dpc * 3);
  for (int64_t c = 0; c < nr; ++c)
    for (int j = 0; j < dpc; ++j) {
      const int64_t src = c * dpc + perm[j], dst = (nr + c) * dpc + j;
      m.cdofs[dst] = m.cdofs[src];
      m.geom[3 * dst] = m.geom[3 * src];
      m.geom[3 * dst + 1] = m.geom[3 * src + 1];
      m.geom[3 * dst + 2] = 2.0 * z0 - m.geom[3 * src + 2];
    }
  m.ncells = 2 * nr;
  m.mirror_z0 = z0;
}

// Columns of the unit-multipole blocks (multipole_integrals,
// bem_fma.cc:1675-1737 keeps one expansion per (block, cell, shape function);
// here the cells' shape functions are merged per DoF, so a leaf's block has one
// column per distinct DoF).
void build_p2m_plan(const HostMesh &m, const QuadData &qd, const Tree &t, FarPlan &fp) {
  const int dpc = m.dpc, nq = qd.nq;
  const int nleaf = (int)fp.p2m_leaf.size();
  // per leaf (host threads): its sorted DoF columns; then the offsets; then
  // every quad point's entries at their final places (the same layout as a
  // sequential pass over the leaves)
  std::vector<std::vector<int32_t>> cols(nleaf);
  parallel_for(nleaf, 64, [&](int64_t li) {
    const int b = fp.p2m_leaf[li];
    const int beg = t.quad_beg[b], cnt = t.quad_cnt[b];
    std::vector<int32_t> &c = cols[li];
    c.reserve((size_t)cnt * dpc);
    for (int i = 0; i < cnt; ++i) {
      const int64_t cell = t.qperm[beg + i] / nq;
      for (int j = 0; j < dpc; ++j) c.push_back(m.cdofs[cell * dpc + j]);
    }
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
  });
  fp.p2m_col_beg.assign(nleaf + 1, 0);
  std::vector<int64_t> qbeg(nleaf + 1, 0);
  for (int li = 0; li < nleaf; ++li) {
    fp.p2m_col_beg[li + 1] = fp.p2m_col_beg[li] + (int)cols[li].size();
    qbeg[li + 1] = qbeg[li] + t.quad_cnt[fp.p2m_leaf[li]];
  }
  fp.p2m_col_dof.assign(fp.p2m_col_beg[nleaf], 0);
  fp.p2m_qpt.assign(qbeg[nleaf], 0);
  fp.p2m_qleaf.assign(qbeg[nleaf], 0);
  fp.p2m_qcol.assign((size_t)qbeg[nleaf] * dpc, 0);
  parallel_for(nleaf, 64, [&](int64_t li) {
    const int b = fp.p2m_leaf[li];
    const int beg = t.quad_beg[b], cnt = t.quad_cnt[b];
    const std::vector<int32_t> &c = cols[li];
    const int c0 = fp.p2m_col_beg[li];
    std::copy(c.begin(), c.end(), fp.p2m_col_dof.begin() + c0);
    for (int i = 0; i < cnt; ++i) {
      const int64_t q = qbeg[li] + i;
      const int g = t.qperm[beg + i];
      const int64_t cell = g / nq;
      fp.p2m_qpt[q] = g;
      fp.p2m_qleaf[q] = (int32_t)li;
      for (int j = 0; j < dpc; ++j) {
        const int d = m.cdofs[cell * dpc + j];
        fp.p2m_qcol[q * dpc + j] = c0 + (int)(std::lower_bound(c.begin(), c.end(), d) - c.begin());
      }
    }
  });
}

}  // namespace pbfmm
