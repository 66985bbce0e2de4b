// TEST INFRASTRUCTURE ONLY -- the CPU oracle for the BEM-FMM path.
//
// A plain FP64 C++ restatement of pi-BEM's BEMFMA (source/bem_fma.cc) and of
// the FMA branch of BEMProblem (source/bem_problem.cc), following the
// reference's data structures (std::set / std::map lists) and loop order.
// It is loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg only, as the checker; nothing in pi-bem_amd/ links it.
//
// Parity status per function:
//   expansions                     pinned (reference goldens, orc_math.h)
//   octree + lists                 pinned by brute force in tests/test_oracle_tree.py
//                                  (partition of nodes/quad points, leaf rule,
//                                  M2L pairs non-adjacent under TOLL, double-node
//                                  intList augmentation, near + far + bigger
//                                  cover every quad point exactly once per node)
//   direct_integrals               pinned: regular part against brute force
//                                  kernel sums; singular rules against closed
//                                  forms at corner, edge and centre points, the
//                                  Telles cubic's defining identities and the
//                                  Duffy split's polynomial exactness
//                                  (tests/test_oracle_quadrature.py); whole
//                                  operator against the dense assembly
//   multipole_matr_vect_products   pinned against the dense collocation
//                                  matvec with the same singular rule
//                                  (truncation error decays with p)
//   vmult / rhs / alpha / GMRES    pinned against the analytic harmonic
//                                  solution (Green's identity residual)
//   deal.II's QTelles / QSplit(QDuffy) as such: not vendored; restated from
//                                  Telles 1987 and Duffy 1982 and pinned by the
//                                  identities above (bit-exact node order of a
//                                  deal.II build stays unpinned)
#include <omp.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <string>

#include "orc_math.h"

namespace orc {

constexpr double TOLL = 0.001;  // bem_fma.cc:1

struct Block {
  int level = 0;
  int64_t parent = 0;
  double pmin[3] = {0, 0, 0};
  double delta = 0;
  std::vector<int64_t> children;
  std::vector<int64_t> nodes;
  std::map<int64_t, std::vector<int>> quads;  // cell -> quad point ids
  std::vector<std::set<int64_t>> nn{1}, il{1}, nil{1};
  void center(double c[3]) const {
    for (int k = 0; k < 3; ++k) c[k] = pmin[k] + delta / 2.0;
  }
};

struct CellME {
  int64_t cell;
  std::vector<Exp> e1, e2;  // per shape function: dipole (Ker1), monopole (Ker2)
};

struct Oracle {
  int p = 6, levels = 10, maxnodes = 20, qorder = 4, sqorder = 5;
  int sing_rule = 0;  // near field: 0 Telles (bem_fma.cc:127-131), 1 Duffy split (bem_problem.cc:213-229)
  int64_t N = 0, ncells = 0;
  int64_t nreal = 0;  // cells [nreal, ncells) are double-body mirror images, never singular
  int dpc = 4;
  std::vector<double> xyz, geom;
  std::vector<int64_t> cdofs;
  std::vector<std::set<int64_t>> dbl;
  std::vector<double> dir;  // 1 = Dirichlet dof
  // quadrature data per cell
  int nq = 0;
  std::vector<double> qp, qn, qw, qs;
  // tree
  std::vector<Block> blocks;
  int64_t num_blocks = 0;
  std::vector<int64_t> startL, endL, childless;
  std::vector<std::vector<int64_t>> dofs_filled, parentList;
  std::vector<std::vector<int64_t>> qp2b, dof2elems;
  // near field (CSR, shared pattern)
  std::vector<int64_t> rowptr, col;
  std::vector<double> vN, vD, vP;
  // multipole integrals, per leaf (indexed like childless)
  std::vector<std::vector<CellME>> elem;
  std::map<int64_t, size_t> leaf_index;
  // expansions
  std::vector<Exp> me1, me2, le1, le2;
  // operator state
  std::vector<double> alpha;
  bool have_dirichlet = false;
  std::string err;
};

// ---------------------------------------------------------------------------
// FE values at arbitrary reference points of one cell (FEValues with the
// isoparametric MappingFEField; normal = x_xi x x_eta / |.|).
// ---------------------------------------------------------------------------
static void fe_values(const Oracle &o, int64_t cell, const std::vector<double> &pts,
                      const std::vector<double> &wts, std::vector<double> &X,
                      std::vector<double> &Nr, std::vector<double> &JxW, std::vector<double> &Sh) {
  FEQ fe(o.dpc);
  const size_t n = wts.size();
  X.assign(3 * n, 0.0);
  Nr.assign(3 * n, 0.0);
  JxW.assign(n, 0.0);
  Sh.assign(n * o.dpc, 0.0);
  const double *G = &o.geom[cell * o.dpc * 3];
  double sh[9], gx[9], gy[9];
  for (size_t q = 0; q < n; ++q) {
    fe.shape(pts[2 * q], pts[2 * q + 1], sh);
    fe.grad(pts[2 * q], pts[2 * q + 1], gx, gy);
    double x[3] = {0, 0, 0}, a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
    for (int j = 0; j < o.dpc; ++j)
      for (int k = 0; k < 3; ++k) {
        x[k] += sh[j] * G[3 * j + k];
        a[k] += gx[j] * G[3 * j + k];
        b[k] += gy[j] * G[3 * j + k];
      }
    double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    double J = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    for (int k = 0; k < 3; ++k) {
      X[3 * q + k] = x[k];
      Nr[3 * q + k] = c[k] / J;
    }
    JxW[q] = J * wts[q];
    for (int j = 0; j < o.dpc; ++j) Sh[q * o.dpc + j] = sh[j];
  }
}

// LaplaceKernel::kernels (include/laplace_kernel.h:55-74)
static inline void lap_kernels(const double R[3], double D[3], double &d) {
  double r = std::sqrt(R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
  double r2 = r * r;
  d = 1. / (r * 4 * PI);
  for (int k = 0; k < 3; ++k) D[k] = R[k] / (-4 * PI * r2 * r);
}

// ---------------------------------------------------------------------------
// generate_octree_blocking (bem_fma.cc:3614-4801)
// ---------------------------------------------------------------------------
static bool adjacent(const Block &b1, const Block &b2) {
  double c1 = b1.delta / 2., c2 = b2.delta / 2.;
  double P1[3], Q1[3], P2[3], Q2[3];
  for (int k = 0; k < 3; ++k) {
    P1[k] = b1.pmin[k];
    Q1[k] = 2. * c1 + b1.pmin[k];
    P2[k] = b2.pmin[k];
    Q2[k] = 2. * c2 + b2.pmin[k];
  }
  auto ov = [&](int k) { return (P1[k] - TOLL <= Q2[k]) && (Q1[k] + TOLL >= P2[k]); };
  auto face = [&](int k) { return (std::fabs(P1[k] - Q2[k]) <= TOLL) || (std::fabs(Q1[k] - P2[k]) <= TOLL); };
  bool r = false;
  if (face(0) && ov(1) && ov(2)) r = true;
  if (face(1) && ov(0) && ov(2)) r = true;
  if (face(2) && ov(1) && ov(0)) r = true;
  return r;
}

static void generate_octree(Oracle &o) {
  const int L = o.levels;
  // quadrature data (bem_fma.cc:3724-3752)
  std::vector<double> gx, gw, pts, wts;
  gauss01(o.qorder, gx, gw);
  tensor2(gx, gw, gx, gw, pts, wts);
  o.nq = (int)wts.size();
  o.qp.assign(o.ncells * o.nq * 3, 0);
  o.qn.assign(o.ncells * o.nq * 3, 0);
  o.qw.assign(o.ncells * o.nq, 0);
  o.qs.assign(o.ncells * o.nq * o.dpc, 0);
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < o.ncells; ++c) {
    std::vector<double> X, Nr, J, S;
    fe_values(o, c, pts, wts, X, Nr, J, S);
    std::copy(X.begin(), X.end(), o.qp.begin() + c * o.nq * 3);
    std::copy(Nr.begin(), Nr.end(), o.qn.begin() + c * o.nq * 3);
    std::copy(J.begin(), J.end(), o.qw.begin() + c * o.nq);
    std::copy(S.begin(), S.end(), o.qs.begin() + c * o.nq * o.dpc);
  }

  double maxc = 0;
  for (int64_t i = 0; i < o.N; ++i)
    for (int k = 0; k < 3; ++k) maxc = std::max(maxc, std::fabs(o.xyz[3 * i + k]));
  for (int64_t i = 3 * o.nreal * o.dpc; i < 3 * o.ncells * o.dpc; ++i) maxc = std::max(maxc, std::fabs(o.geom[i]));

  o.blocks.clear();
  o.blocks.reserve(1024);
  o.startL.assign(L + 1, 0);
  o.endL.assign(L + 1, 0);
  o.parentList.assign(L + 1, {});
  o.parentList[0].push_back(0);
  o.childless.clear();
  o.dofs_filled.assign(L + 1, {});
  o.qp2b.assign(o.ncells * o.nq, {});
  o.dof2elems.assign(o.N, {});

  Block root;
  for (int k = 0; k < 3; ++k) root.pmin[k] = -1.1 * maxc;
  double delta = 2.2 * maxc;
  root.delta = delta;
  for (int64_t c = 0; c < o.ncells; ++c) {
    for (int j = 0; j < o.nq; ++j) {
      root.quads[c].push_back(j);
      o.qp2b[c * o.nq + j].push_back(0);
    }
    if (c < o.nreal)
      for (int j = 0; j < o.dpc; ++j) o.dof2elems[o.cdofs[c * o.dpc + j]].push_back(c);
  }
  for (int64_t i = 0; i < o.N; ++i) root.nodes.push_back(i);
  o.blocks.push_back(root);
  int64_t blocksCount = 0;

  static const double offs[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                    {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  auto child_of = [&](const double pm[3], double dl, const double x[3]) {
    int c;
    if (x[2] <= pm[2] + dl) {
      if (x[1] <= pm[1] + dl) c = (x[0] <= pm[0] + dl) ? 0 : 1;
      else c = (x[0] <= pm[0] + dl) ? 3 : 2;
    } else {
      if (x[1] <= pm[1] + dl) c = (x[0] <= pm[0] + dl) ? 4 : 5;
      else c = (x[0] <= pm[0] + dl) ? 7 : 6;
    }
    return c;
  };

  for (int level = 1; level <= L; ++level) {
    delta /= 2.;
    for (size_t kk = 0; kk < o.parentList[level - 1].size(); ++kk) {
      int64_t jj = o.parentList[level - 1][kk];
      Block ch[8];
      double pm[3] = {o.blocks[jj].pmin[0], o.blocks[jj].pmin[1], o.blocks[jj].pmin[2]};
      for (int c = 0; c < 8; ++c) {
        ch[c].level = level;
        ch[c].parent = jj;
        ch[c].delta = delta;
        for (int k = 0; k < 3; ++k) ch[c].pmin[k] = pm[k] + (offs[c][k] != 0 ? delta : 0.);
      }
      for (int64_t id : o.blocks[jj].nodes) ch[child_of(pm, delta, &o.xyz[3 * id])].nodes.push_back(id);
      for (auto &cq : o.blocks[jj].quads)
        for (int q : cq.second)
          ch[child_of(pm, delta, &o.qp[(cq.first * o.nq + q) * 3])].quads[cq.first].push_back(q);
      for (int c = 0; c < 8; ++c) {
        if (ch[c].nodes.size() + ch[c].quads.size() > 0) {
          blocksCount += 1;
          o.blocks.push_back(ch[c]);
          o.blocks[jj].children.push_back(blocksCount);
          for (auto &cq : o.blocks[blocksCount].quads)
            for (int q : cq.second) o.qp2b[cq.first * o.nq + q].push_back(blocksCount);
        }
      }
    }
    o.startL[level] = o.endL[level - 1] + 1;
    o.endL[level] = blocksCount;
    for (int64_t jj = o.startL[level]; jj <= o.endL[level]; ++jj) {
      Block &b = o.blocks[jj];
      double bn = 0.0;
      for (int64_t id : b.nodes) bn += 1.0 / double(o.dbl[id].size());
      if (std::round(bn) <= o.maxnodes) {
        o.childless.push_back(jj);
        for (auto &cq : b.quads)
          for (int q : cq.second)
            for (int j = level + 1; j <= L; ++j) o.qp2b[cq.first * o.nq + q].push_back(jj);
      } else {
        o.parentList[level].push_back(jj);
      }
      if (bn > 0) o.dofs_filled[level].push_back(jj);
    }
  }
  for (int64_t jj : o.parentList[L]) o.childless.push_back(jj);
  o.num_blocks = blocksCount + 1;

  // near neighbours (bem_fma.cc:4271-4579)
  for (int64_t ii = o.startL[1]; ii <= o.endL[1]; ++ii)
    for (int64_t jj = o.startL[1]; jj <= o.endL[1]; ++jj) o.blocks[ii].nn[0].insert(jj);
  for (int level = 2; level <= L; ++level) {
    for (int64_t kk = o.startL[level]; kk <= o.endL[level]; ++kk) {
      Block &b1 = o.blocks[kk];
      b1.nn[0].insert(kk);
      std::set<int64_t> pnn = o.blocks[b1.parent].nn[0];
      for (int64_t pos : pnn) {
        const Block &b2 = o.blocks[pos];
        if (b2.children.empty() && adjacent(b1, b2)) b1.nn[0].insert(pos);
        for (int64_t c : b2.children)
          if (adjacent(b1, o.blocks[c])) b1.nn[0].insert(c);
      }
      if (b1.children.empty()) {
        const int ns = L - level + 1;
        b1.nn.resize(ns);
        b1.il.resize(ns);
        b1.nil.resize(ns);
        for (int s = 1; s < ns; ++s) {
          std::set<int64_t> up = b1.nn[s - 1];
          for (int64_t pos : up) {
            const Block &b2 = o.blocks[pos];
            if (b2.children.empty()) b1.nn[s].insert(pos);
            for (int64_t c : b2.children)
              if (adjacent(b1, o.blocks[c])) b1.nn[s].insert(c);
          }
        }
      }
    }
  }
  // interaction lists (bem_fma.cc:4610-4711)
  for (int64_t ii = o.startL[1]; ii <= o.endL[1]; ++ii)
    for (int64_t jj = o.startL[1]; jj <= o.endL[1]; ++jj) o.blocks[ii].il[0].insert(jj);
  for (int level = 2; level <= L; ++level) {
    for (int64_t jj = o.startL[level]; jj <= o.endL[level]; ++jj) {
      Block &b1 = o.blocks[jj];
      for (size_t s = 0; s < b1.nn.size(); ++s) {
        for (int64_t pos : b1.nn[s]) b1.il[s].insert(pos);
        for (int64_t node : b1.nodes)
          for (int64_t d : o.dbl[node])
            for (int64_t cell : o.dof2elems[d])
              for (int j = 0; j < o.nq; ++j) b1.il[s].insert(o.qp2b[cell * o.nq + j][level + s]);
      }
      for (size_t s = 0; s < b1.nn.size(); ++s) {
        std::set<int64_t> ilist = b1.il[s];
        std::set<int64_t> pil = (s == 0) ? o.blocks[b1.parent].il[0] : b1.il[s - 1];
        for (int64_t pos : pil) {
          const Block &b2 = o.blocks[pos];
          if (b2.children.empty()) {
            if (ilist.count(pos) == 0) b1.nil[s].insert(pos);
          } else {
            for (int64_t c : b2.children)
              if (ilist.count(c) == 0) b1.nil[s].insert(c);
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// direct_integrals (bem_fma.cc:107-1600)
// ---------------------------------------------------------------------------
struct RowAcc {
  std::map<int64_t, std::array<double, 3>> v;  // col -> N, D, prec
};

static void integrate_regular(const Oracle &o, int64_t node, int64_t cell, const std::set<int> &qset,
                              double *Nj, double *Dj) {
  for (int j = 0; j < o.dpc; ++j) Nj[j] = Dj[j] = 0;
  for (int q : qset) {
    const int64_t g = cell * o.nq + q;
    double R[3], D[3], s;
    for (int k = 0; k < 3; ++k) R[k] = o.qp[3 * g + k] - o.xyz[3 * node + k];
    lap_kernels(R, D, s);
    const double dn = D[0] * o.qn[3 * g] + D[1] * o.qn[3 * g + 1] + D[2] * o.qn[3 * g + 2];
    for (int j = 0; j < o.dpc; ++j) {
      Nj[j] += dn * o.qs[g * o.dpc + j] * o.qw[g];
      Dj[j] += s * o.qs[g * o.dpc + j] * o.qw[g];
    }
  }
}

// Telles rule of singular_quadrature_order towards unit support point js
// (bem_fma.cc:127-131, 838-904) or the Duffy split rule (bem_problem.cc:213-229).
static void integrate_singular(const Oracle &o, int64_t node, int64_t cell, int js, int rule,
                               double *Nj, double *Dj) {
  FEQ fe(o.dpc);
  std::vector<double> pts, wts;
  if (rule == 0) {
    std::vector<double> x1, w1, x2, w2;
    telles01(o.sqorder, fe.usp[js][0], x1, w1);
    telles01(o.sqorder, fe.usp[js][1], x2, w2);
    tensor2(x1, w1, x2, w2, pts, wts);
  } else {
    duffy_split(o.sqorder, fe.usp[js][0], fe.usp[js][1], pts, wts);
  }
  std::vector<double> X, Nr, J, S;
  fe_values(o, cell, pts, wts, X, Nr, J, S);
  for (int j = 0; j < o.dpc; ++j) Nj[j] = Dj[j] = 0;
  for (size_t q = 0; q < wts.size(); ++q) {
    double R[3], D[3], s;
    for (int k = 0; k < 3; ++k) R[k] = X[3 * q + k] - o.xyz[3 * node + k];
    lap_kernels(R, D, s);
    const double dn = D[0] * Nr[3 * q] + D[1] * Nr[3 * q + 1] + D[2] * Nr[3 * q + 2];
    for (int j = 0; j < o.dpc; ++j) {
      Nj[j] += dn * S[q * o.dpc + j] * J[q];
      Dj[j] += s * S[q * o.dpc + j] * J[q];
    }
  }
}

static void direct_integrals(Oracle &o) {
  std::vector<RowAcc> rows(o.N);
  auto add_row = [&](RowAcc &r, int64_t cell, const double *Nj, const double *Dj) {
    for (int j = 0; j < o.dpc; ++j) {
      int64_t c = o.cdofs[cell * o.dpc + j];
      auto &e = r.v[c];
      e[0] += Nj[j];
      e[1] += Dj[j];
      e[2] += (o.dir[c] > 0.8) ? -Dj[j] : Nj[j];
    }
  };
  // leaves: near quad points of the deepest intList (bem_fma.cc:670-913)
  const int64_t nleaf = (int64_t)o.childless.size();
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t kk = 0; kk < nleaf; ++kk) {
    const Block &b1 = o.blocks[o.childless[kk]];
    if (b1.nodes.empty()) continue;
    std::map<int64_t, std::set<int>> dq;
    for (int64_t pos : b1.il.back())
      for (auto &cq : o.blocks[pos].quads)
        for (int q : cq.second) dq[cq.first].insert(q);
    for (int64_t node : b1.nodes) {
      RowAcc &r = rows[node];
      for (auto &cq : dq) {
        const int64_t cell = cq.first;
        int js = -1;
        for (int j = 0; j < o.dpc && cell < o.nreal; ++j)
          if (o.dbl[node].count(o.cdofs[cell * o.dpc + j])) {
            js = j;
            break;
          }
        double Nj[9], Dj[9];
        if (js < 0) integrate_regular(o, node, cell, cq.second, Nj, Dj);
        else integrate_singular(o, node, cell, js, o.sing_rule, Nj, Dj);
        add_row(r, cell, Nj, Dj);
      }
    }
  }
  // bigger blocks in every level's nonIntList (bem_fma.cc:997-1153)
  for (int level = 1; level <= o.levels; ++level) {
    const int64_t sb = o.startL[level];
    const auto &df = o.dofs_filled[level];
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t k = 0; k < (int64_t)df.size(); ++k) {
      const Block &b1 = o.blocks[df[k]];
      std::map<int64_t, std::set<int>> dq;
      for (size_t s = 0; s < b1.nn.size(); ++s)
        for (auto it = b1.nil[s].begin(); it != b1.nil[s].lower_bound(sb); ++it)
          for (auto &cq : o.blocks[*it].quads)
            for (int q : cq.second) dq[cq.first].insert(q);
      if (dq.empty()) continue;
      for (int64_t node : b1.nodes) {
        RowAcc &r = rows[node];
        for (auto &cq : dq) {
          double Nj[9], Dj[9];
          integrate_regular(o, node, cq.first, cq.second, Nj, Dj);
          add_row(r, cq.first, Nj, Dj);
        }
      }
    }
  }
  o.rowptr.assign(o.N + 1, 0);
  for (int64_t i = 0; i < o.N; ++i) o.rowptr[i + 1] = o.rowptr[i] + (int64_t)rows[i].v.size();
  const int64_t nnz = o.rowptr[o.N];
  o.col.assign(nnz, 0);
  o.vN.assign(nnz, 0);
  o.vD.assign(nnz, 0);
  o.vP.assign(nnz, 0);
  for (int64_t i = 0; i < o.N; ++i) {
    int64_t k = o.rowptr[i];
    for (auto &e : rows[i].v) {
      o.col[k] = e.first;
      o.vN[k] = e.second[0];
      o.vD[k] = e.second[1];
      o.vP[k] = e.second[2];
      ++k;
    }
  }
}

// ---------------------------------------------------------------------------
// multipole_integrals (bem_fma.cc:1608-1840)
// ---------------------------------------------------------------------------
static void multipole_integrals(Oracle &o) {
  const int64_t nleaf = (int64_t)o.childless.size();
  o.elem.assign(nleaf, {});
  o.leaf_index.clear();
  for (int64_t k = 0; k < nleaf; ++k) o.leaf_index[o.childless[k]] = k;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t kk = 0; kk < nleaf; ++kk) {
    const Block &b = o.blocks[o.childless[kk]];
    double c[3];
    b.center(c);
    for (auto &cq : b.quads) {
      CellME ce;
      ce.cell = cq.first;
      for (int j = 0; j < o.dpc; ++j) {
        ce.e1.emplace_back(o.p, c);
        ce.e2.emplace_back(o.p, c);
      }
      for (int q : cq.second) {
        const int64_t g = cq.first * o.nq + q;
        for (int j = 0; j < o.dpc; ++j) {
          const double str = o.qs[g * o.dpc + j] * o.qw[g] / 4 / PI;
          me_add_normder(ce.e1[j], str, &o.qp[3 * g], &o.qn[3 * g]);
          me_add_point(ce.e2[j], str, &o.qp[3 * g]);
        }
      }
      o.elem[kk].push_back(std::move(ce));
    }
  }
}

// ---------------------------------------------------------------------------
// generate_multipole_expansions (bem_fma.cc:1970-2380)
// ---------------------------------------------------------------------------
static void generate_me(Oracle &o, const double *phi, const double *dphi) {
  o.me1.assign(o.num_blocks, Exp());
  o.me2.assign(o.num_blocks, Exp());
  for (int64_t b = 0; b < o.num_blocks; ++b) {
    double c[3];
    o.blocks[b].center(c);
    o.me1[b] = Exp(o.p, c);
    o.me2[b] = Exp(o.p, c);
  }
  const int64_t nleaf = (int64_t)o.childless.size();
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t kk = 0; kk < nleaf; ++kk) {
    const int64_t id = o.childless[kk];
    for (const CellME &ce : o.elem[kk])
      for (int j = 0; j < o.dpc; ++j) {
        const int64_t dof = o.cdofs[ce.cell * o.dpc + j];
        me_add_scaled(o.me2[id], ce.e2[j], dphi[dof]);
        me_add_scaled(o.me1[id], ce.e1[j], phi[dof]);
      }
  }
  for (int level = o.levels; level > 0; --level) {
    const int64_t s = o.startL[level], e = o.endL[level];
    if (e < s) continue;
    std::vector<Exp> t1(e - s + 1), t2(e - s + 1);
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t b = s; b <= e; ++b) {
      double pc[3];
      o.blocks[o.blocks[b].parent].center(pc);
      t1[b - s] = Exp(o.p, pc);
      t2[b - s] = Exp(o.p, pc);
      me_add_child(t1[b - s], o.me1[b]);
      me_add_child(t2[b - s], o.me2[b]);
    }
    for (int64_t b = s; b <= e; ++b) {
      const int64_t par = o.blocks[b].parent;
      me_add_child(o.me1[par], t1[b - s]);
      me_add_child(o.me2[par], t2[b - s]);
    }
  }
}

static void csr_vmult(const Oracle &o, const std::vector<double> &v, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < o.N; ++i) {
    double s = 0;
    for (int64_t k = o.rowptr[i]; k < o.rowptr[i + 1]; ++k) s += v[k] * x[o.col[k]];
    y[i] = s;
  }
}

// ---------------------------------------------------------------------------
// multipole_matr_vect_products (bem_fma.cc:2390-3072)
// ---------------------------------------------------------------------------
static void products(Oracle &o, const double *phi, const double *dphi, double *outN, double *outD) {
  csr_vmult(o, o.vN, phi, outN);
  csr_vmult(o, o.vD, dphi, outD);
  o.le1.assign(o.num_blocks, Exp());
  o.le2.assign(o.num_blocks, Exp());
  for (int64_t b = 0; b < o.num_blocks; ++b) {
    double c[3];
    o.blocks[b].center(c);
    o.le1[b] = Exp(o.p, c);
    o.le2[b] = Exp(o.p, c);
  }
  for (int level = 1; level <= o.levels; ++level) {
    const int64_t sb = o.startL[level], eb = o.endL[level];
    if (eb < sb) continue;
    const auto &df = o.dofs_filled[level];
    const int64_t nb = (int64_t)df.size();
    std::vector<Exp> t1(nb), t2(nb);
    std::vector<std::vector<double>> c1(nb), c2(nb);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t k = 0; k < nb; ++k) {
      const int64_t id = df[k];
      const Block &b = o.blocks[id];
      double c[3];
      b.center(c);
      t1[k] = Exp(o.p, c);
      t2[k] = Exp(o.p, c);
      c1[k].assign(b.nodes.size(), 0.0);
      c2[k].assign(b.nodes.size(), 0.0);
      le_add_parent(t1[k], o.le1[b.parent]);
      le_add_parent(t2[k], o.le2[b.parent]);
      for (size_t s = 0; s < b.nn.size(); ++s) {
        const auto &nil = b.nil[s];
        for (auto it = nil.lower_bound(sb); it != nil.upper_bound(eb); ++it) {
          le_add_multipole(t1[k], o.me1[*it]);
          le_add_multipole(t2[k], o.me2[*it]);
        }
        for (auto it = nil.upper_bound(eb); it != nil.end(); ++it)
          for (size_t ii = 0; ii < b.nodes.size(); ++ii) {
            const double *x = &o.xyz[3 * b.nodes[ii]];
            c1[k][ii] += me_evaluate(o.me1[*it], x);
            c2[k][ii] += me_evaluate(o.me2[*it], x);
          }
      }
    }
    for (int64_t k = 0; k < nb; ++k) {
      const int64_t id = df[k];
      le_add_parent(o.le1[id], t1[k]);
      le_add_parent(o.le2[id], t2[k]);
      const Block &b = o.blocks[id];
      for (size_t ii = 0; ii < b.nodes.size(); ++ii) {
        outD[b.nodes[ii]] += c2[k][ii];
        outN[b.nodes[ii]] += c1[k][ii];
      }
    }
  }
  const int64_t nleaf = (int64_t)o.childless.size();
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t kk = 0; kk < nleaf; ++kk) {
    const int64_t id = o.childless[kk];
    for (int64_t node : o.blocks[id].nodes) {
      const double *x = &o.xyz[3 * node];
      outD[node] += le_evaluate(o.le2[id], x);
      outN[node] += le_evaluate(o.le1[id], x);
    }
  }
}

// ---------------------------------------------------------------------------
// BEMProblem, FMA branch: compute_alpha (bem_problem.cc:1138-1162), vmult
// (1164-1217), compute_rhs (1219-1260); ConstrainedOperator
// (include/constrained_matrix.h:91-147).
// ---------------------------------------------------------------------------
static double l2(const std::vector<double> &v) {
  double s = 0;
  for (double x : v) s += x * x;
  return std::sqrt(s);
}

static void fmm_apply(Oracle &o, const std::vector<double> &phi, const std::vector<double> &dphi,
                      std::vector<double> &N, std::vector<double> &D) {
  N.assign(o.N, 0.0);
  D.assign(o.N, 0.0);
  generate_me(o, phi.data(), dphi.data());
  products(o, phi.data(), dphi.data(), N.data(), D.data());
}

static void compute_alpha(Oracle &o) {
  std::vector<double> ones(o.N, -1.0), zeros(o.N, 0.0), dummy;
  fmm_apply(o, ones, zeros, o.alpha, dummy);
}

static void vmult(Oracle &o, const double *src, double *dst) {
  std::vector<double> sp(src, src + o.N), sd(src, src + o.N);
  if (!o.have_dirichlet) {
    double nrm = l2(sp);
    for (double &x : sp) x -= nrm;
  }
  for (int64_t i = 0; i < o.N; ++i) {
    sp[i] *= (1.0 - o.dir[i]);
    sd[i] *= o.dir[i];
  }
  std::vector<double> N, D;
  fmm_apply(o, sp, sd, N, D);
  for (int64_t i = 0; i < o.N; ++i) dst[i] = -D[i] + N[i] + o.alpha[i] * sp[i];
  if (!o.have_dirichlet) {
    std::vector<double> d(dst, dst + o.N);
    double nrm = l2(d);
    for (int64_t i = 0; i < o.N; ++i) dst[i] -= nrm;
  }
}

static void compute_rhs(Oracle &o, const double *g, double *dst) {
  std::vector<double> sp(o.N), sd(o.N);
  for (int64_t i = 0; i < o.N; ++i) {
    sp[i] = g[i] * o.dir[i];
    sd[i] = g[i] * (1.0 - o.dir[i]);
  }
  std::vector<double> N, D;
  fmm_apply(o, sp, sd, N, D);
  for (int64_t i = 0; i < o.N; ++i) dst[i] = -(N[i] + o.alpha[i] * sp[i]) + D[i];
}

struct Constraints {
  std::vector<int64_t> ptr, idx;
  std::vector<double> val, inhom;
  std::vector<uint8_t> is_c;
};

static void cc_vmult(Oracle &o, const Constraints &c, const double *src, double *dst) {
  vmult(o, src, dst);
  if (c.is_c.empty()) return;
  for (int64_t i = 0; i < o.N; ++i)
    if (c.is_c[i]) {
      double v = src[i];
      for (int64_t k = c.ptr[i]; k < c.ptr[i + 1]; ++k) v -= c.val[k] * src[c.idx[k]];
      dst[i] = v;
    }
}

// Restarted GMRES(m), modified Gram-Schmidt, no preconditioner; stops on the
// true relative residual ||b - A x|| / ||b|| <= tol (deal.II's SolverGMRES is
// not vendored: its left-preconditioned absolute criterion is not restated).
static int gmres(Oracle &o, const Constraints &c, const double *b, double *x, double tol, int m,
                 int maxit, int *iters, double *relres) {
  const int64_t n = o.N;
  std::vector<double> r(n), w(n);
  double bn = 0;
  for (int64_t i = 0; i < n; ++i) bn += b[i] * b[i];
  bn = std::sqrt(bn);
  if (bn == 0) bn = 1;
  int it = 0;
  double rr = 1;
  std::vector<std::vector<double>> V(m + 1, std::vector<double>(n));
  std::vector<double> H((m + 1) * m), cs(m), sn(m), g(m + 1);
  while (it < maxit) {
    cc_vmult(o, c, x, w.data());
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - w[i];
    double beta = 0;
    for (double v : r) beta += v * v;
    beta = std::sqrt(beta);
    rr = beta / bn;
    if (rr <= tol) break;
    for (int64_t i = 0; i < n; ++i) V[0][i] = r[i] / beta;
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    for (; k < m && it < maxit; ++k, ++it) {
      cc_vmult(o, c, V[k].data(), w.data());
      for (int j = 0; j <= k; ++j) {
        double h = 0;
        for (int64_t i = 0; i < n; ++i) h += w[i] * V[j][i];
        H[j * m + k] = h;
        for (int64_t i = 0; i < n; ++i) w[i] -= h * V[j][i];
      }
      double hn = 0;
      for (double v : w) hn += v * v;
      hn = std::sqrt(hn);
      H[(k + 1) * m + k] = hn;
      for (int64_t i = 0; i < n; ++i) V[k + 1][i] = (hn > 0) ? w[i] / hn : 0.0;
      for (int j = 0; j < k; ++j) {
        double a = H[j * m + k], bb = H[(j + 1) * m + k];
        H[j * m + k] = cs[j] * a + sn[j] * bb;
        H[(j + 1) * m + k] = -sn[j] * a + cs[j] * bb;
      }
      double a = H[k * m + k], bb = H[(k + 1) * m + k];
      double den = std::sqrt(a * a + bb * bb);
      cs[k] = a / den;
      sn[k] = bb / den;
      H[k * m + k] = den;
      H[(k + 1) * m + k] = 0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      if (std::fabs(g[k + 1]) / bn <= tol * 0.5) {
        ++k;
        ++it;
        break;
      }
    }
    std::vector<double> y(k);
    for (int i = k - 1; i >= 0; --i) {
      double s = g[i];
      for (int j = i + 1; j < k; ++j) s -= H[i * m + j] * y[j];
      y[i] = s / H[i * m + i];
    }
    for (int j = 0; j < k; ++j)
      for (int64_t i = 0; i < n; ++i) x[i] += y[j] * V[j][i];
  }
  cc_vmult(o, c, x, w.data());
  double res = 0;
  for (int64_t i = 0; i < n; ++i) res += (b[i] - w[i]) * (b[i] - w[i]);
  rr = std::sqrt(res) / bn;
  if (iters) *iters = it;
  if (relres) *relres = rr;
  return 0;
}

// Dense one-body collocation assembly (bem_problem.cc:477-950): every cell
// integrated against every collocation node, regular Gauss rule, singular
// rule (0 Telles as the FMA near field, 1 Duffy split as get_singular_quadrature)
// where the cell holds a double of the node.
static void dense_assemble(Oracle &o, int rule, double *Nm, double *Dm) {
  std::vector<double> gx, gw, pts, wts;
  gauss01(o.qorder, gx, gw);
  tensor2(gx, gw, gx, gw, pts, wts);
  std::set<int> all;
  for (int q = 0; q < o.nq; ++q) all.insert(q);
#pragma omp parallel for schedule(dynamic, 8)
  for (int64_t i = 0; i < o.N; ++i) {
    double *nr = Nm + i * o.N, *dr = Dm + i * o.N;
    std::fill(nr, nr + o.N, 0.0);
    std::fill(dr, dr + o.N, 0.0);
    for (int64_t cell = 0; cell < o.ncells; ++cell) {
      int js = -1;
      for (int j = 0; j < o.dpc && cell < o.nreal; ++j)
        if (o.dbl[i].count(o.cdofs[cell * o.dpc + j])) {
          js = j;
          break;
        }
      double Nj[9], Dj[9];
      if (js < 0) integrate_regular(o, i, cell, all, Nj, Dj);
      else integrate_singular(o, i, cell, js, rule, Nj, Dj);
      for (int j = 0; j < o.dpc; ++j) {
        nr[o.cdofs[cell * o.dpc + j]] += Nj[j];
        dr[o.cdofs[cell * o.dpc + j]] += Dj[j];
      }
    }
  }
}

}  // namespace orc

// ===========================================================================
// C API (ctypes) -- test infrastructure
// ===========================================================================
using namespace orc;

static void pack(const Exp &e, double *out) {
  for (int i = 0; i < e.kt(); ++i) {
    out[2 * i] = e.a[i].real();
    out[2 * i + 1] = e.a[i].imag();
  }
}
static Exp unpack(int p, const double *c, const double *in) {
  Exp e(p, c);
  for (int i = 0; i < e.kt(); ++i) e.a[i] = cd(in[2 * i], in[2 * i + 1]);
  e.is_zero = false;
  return e;
}

extern "C" {

void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_max_threads() { return omp_get_max_threads(); }

void orc_legendre(int nmax, double x, double *P, double *dP) {
  LegTable t;
  t.fill(nmax, x, dP != nullptr);
  for (int n = 0; n <= nmax; ++n)
    for (int m = 0; m <= nmax; ++m) {
      P[n * (nmax + 1) + m] = (m <= n) ? t.p(n, m) : 0.0;
      if (dP) dP[n * (nmax + 1) + m] = (m <= n) ? t.dp(n, m) : 0.0;
    }
}

void orc_me_p2m(int p, const double *c, int n, const double *q, const double *y, double *out) {
  Exp e(p, c);
  for (int i = 0; i < n; ++i) me_add_point(e, q[i], y + 3 * i);
  pack(e, out);
}
void orc_me_p2m_dipole(int p, const double *c, int n, const double *q, const double *y,
                       const double *nrm, double *out) {
  Exp e(p, c);
  for (int i = 0; i < n; ++i) me_add_normder(e, q[i], y + 3 * i, nrm + 3 * i);
  pack(e, out);
}
void orc_me_m2m(int p, const double *cc, const double *child, const double *pc, double *out) {
  Exp ch = unpack(p, cc, child), par(p, pc);
  me_add_child(par, ch);
  pack(par, out);
}
double orc_me_eval(int p, const double *c, const double *coef, const double *x) {
  Exp e = unpack(p, c, coef);
  return me_evaluate(e, x);
}
void orc_le_m2l(int p, const double *mc, const double *me, const double *lc, double *out) {
  Exp m = unpack(p, mc, me), l(p, lc);
  le_add_multipole(l, m);
  pack(l, out);
}
void orc_le_l2l(int p, const double *pc, const double *pl, const double *cc, double *out) {
  Exp par = unpack(p, pc, pl), ch(p, cc);
  le_add_parent(ch, par);
  pack(ch, out);
}
double orc_le_eval(int p, const double *c, const double *coef, const double *x) {
  Exp e = unpack(p, c, coef);
  return le_evaluate(e, x);
}
int orc_gauss01(int n, double *x, double *w) {
  std::vector<double> a, b;
  gauss01(n, a, b);
  std::copy(a.begin(), a.end(), x);
  std::copy(b.begin(), b.end(), w);
  return n;
}
int orc_telles01(int n, double x0, double *x, double *w) {
  std::vector<double> a, b;
  telles01(n, x0, a, b);
  std::copy(a.begin(), a.end(), x);
  std::copy(b.begin(), b.end(), w);
  return (int)a.size();
}
int orc_duffy(int n, double sx, double sy, double *pts, double *w) {
  std::vector<double> a, b;
  duffy_split(n, sx, sy, a, b);
  std::copy(a.begin(), a.end(), pts);
  std::copy(b.begin(), b.end(), w);
  return (int)b.size();
}

void *orc_create(int p, int levels, int maxnodes, int qorder, int sqorder) {
  Oracle *o = new Oracle();
  o->p = p;
  o->levels = levels;
  o->maxnodes = maxnodes;
  o->qorder = qorder;
  o->sqorder = sqorder;
  return o;
}
void orc_destroy(void *h) { delete (Oracle *)h; }
// The near field's singular rule (before direct_integrals): 0 QTelles as
// BEMFMA::direct_integrals (bem_fma.cc:127-131), 1 QSplit(QDuffy) as
// BEMProblem::get_singular_quadrature (bem_problem.cc:213-229).
int orc_set_sing_rule(void *h, int rule) {
  if (rule != 0 && rule != 1) return -1;
  ((Oracle *)h)->sing_rule = rule;
  return 0;
}

int orc_init(void *h, int64_t n, const double *xyz, int64_t ncells, int dpc, const int64_t *cdofs,
             const double *geom, const int64_t *dptr, const int64_t *didx, const uint8_t *dirm) {
  Oracle &o = *(Oracle *)h;
  if (dpc != 4 && dpc != 9) return -1;
  o.N = n;
  o.ncells = ncells;
  o.nreal = ncells;
  o.dpc = dpc;
  o.xyz.assign(xyz, xyz + 3 * n);
  o.cdofs.assign(cdofs, cdofs + ncells * dpc);
  o.geom.assign(ncells * dpc * 3, 0.0);
  for (int64_t c = 0; c < ncells; ++c)
    for (int j = 0; j < dpc; ++j)
      for (int k = 0; k < 3; ++k)
        o.geom[(c * dpc + j) * 3 + k] = geom ? geom[(c * dpc + j) * 3 + k] : xyz[3 * cdofs[c * dpc + j] + k];
  o.dbl.assign(n, {});
  for (int64_t i = 0; i < n; ++i) {
    if (dptr)
      for (int64_t k = dptr[i]; k < dptr[i + 1]; ++k) o.dbl[i].insert(didx[k]);
    o.dbl[i].insert(i);
  }
  o.dir.assign(n, 0.0);
  o.have_dirichlet = false;
  for (int64_t i = 0; i < n; ++i) {
    o.dir[i] = dirm[i] ? 1.0 : 0.0;
    if (dirm[i]) o.have_dirichlet = true;
  }
  return 0;
}

// Double body (BEMProblem::_assemble_system_double_body,
// bem_problem.cc:957-1135): every real cell also acts through its mirror image
// about z = z0, with reflected quadrature points, the normal's z component
// negated and the regular rule only.  The image is the reflected mapping nodes
// with the reference directions swapped (Q1 {0,2,1,3}, Q2 {0,2,1,3,6,7,4,5,8}),
// which gives the reflected points with the same shape values, JxW and the
// reflected normal.  Before generate_octree_blocking.
int orc_set_double_body(void *h, double z0) {
  Oracle &o = *(Oracle *)h;
  static const int perm_q1[4] = {0, 2, 1, 3};
  static const int perm_q2[9] = {0, 2, 1, 3, 6, 7, 4, 5, 8};
  const int *perm = o.dpc == 4 ? perm_q1 : perm_q2;
  const int64_t nr = o.nreal, dpc = o.dpc;
  o.cdofs.resize(2 * nr * dpc);
  o.geom.resize(2 * nr * dpc * 3);
  for (int64_t c = 0; c < nr; ++c)
    for (int j = 0; j < dpc; ++j) {
      const int64_t src = c * dpc + perm[j], dst = (nr + c) * dpc + j;
      o.cdofs[dst] = o.cdofs[src];
      o.geom[3 * dst] = o.geom[3 * src];
      o.geom[3 * dst + 1] = o.geom[3 * src + 1];
      o.geom[3 * dst + 2] = 2.0 * z0 - o.geom[3 * src + 2];
    }
  o.ncells = 2 * nr;
  return 0;
}

int orc_generate_octree_blocking(void *h) {
  generate_octree(*(Oracle *)h);
  return 0;
}
int orc_direct_integrals(void *h) {
  direct_integrals(*(Oracle *)h);
  return 0;
}
int orc_multipole_integrals(void *h) {
  multipole_integrals(*(Oracle *)h);
  return 0;
}
int orc_generate_multipole_expansions(void *h, const double *phi, const double *dphi) {
  generate_me(*(Oracle *)h, phi, dphi);
  return 0;
}
int orc_multipole_matr_vect_products(void *h, const double *phi, const double *dphi, double *outN,
                                     double *outD) {
  Oracle &o = *(Oracle *)h;
  products(o, phi, dphi, outN, outD);
  return 0;
}

// counts[0..7]: num_blocks, levels, n_childless, n_sub (sum of sublevels),
// n_il, n_nil, n_nodes_total, n_quad_entries_total
void orc_tree_counts(void *h, int64_t *cnt) {
  Oracle &o = *(Oracle *)h;
  cnt[0] = o.num_blocks;
  cnt[1] = o.levels;
  cnt[2] = (int64_t)o.childless.size();
  int64_t ns = 0, nil = 0, nnil = 0, nn = 0, nqe = 0;
  for (auto &b : o.blocks) {
    ns += (int64_t)b.nn.size();
    for (auto &s : b.il) nil += (int64_t)s.size();
    for (auto &s : b.nil) nnil += (int64_t)s.size();
    nn += (int64_t)b.nodes.size();
    for (auto &cq : b.quads) nqe += (int64_t)cq.second.size();
  }
  cnt[3] = ns;
  cnt[4] = nil;
  cnt[5] = nnil;
  cnt[6] = nn;
  cnt[7] = nqe;
}

void orc_tree_export(void *h, int32_t *level, int64_t *parent, double *pmin, double *delta,
                     int64_t *startL, int64_t *endL, int64_t *childless, int64_t *sub_ptr,
                     int64_t *il_ptr, int64_t *il_idx, int64_t *nil_ptr, int64_t *nil_idx,
                     int64_t *node_ptr, int64_t *node_idx) {
  Oracle &o = *(Oracle *)h;
  for (int l = 0; l <= o.levels; ++l) {
    startL[l] = o.startL[l];
    endL[l] = o.endL[l];
  }
  for (size_t k = 0; k < o.childless.size(); ++k) childless[k] = o.childless[k];
  int64_t s = 0, a = 0, b = 0, nd = 0;
  sub_ptr[0] = 0;
  il_ptr[0] = 0;
  nil_ptr[0] = 0;
  node_ptr[0] = 0;
  for (int64_t i = 0; i < o.num_blocks; ++i) {
    const Block &bl = o.blocks[i];
    level[i] = bl.level;
    parent[i] = bl.parent;
    for (int k = 0; k < 3; ++k) pmin[3 * i + k] = bl.pmin[k];
    delta[i] = bl.delta;
    for (size_t t = 0; t < bl.nn.size(); ++t) {
      for (int64_t x : bl.il[t]) il_idx[a++] = x;
      for (int64_t x : bl.nil[t]) nil_idx[b++] = x;
      ++s;
      il_ptr[s] = a;
      nil_ptr[s] = b;
    }
    sub_ptr[i + 1] = s;
    for (int64_t x : bl.nodes) node_idx[nd++] = x;
    node_ptr[i + 1] = nd;
  }
}

int64_t orc_csr_nnz(void *h) { return ((Oracle *)h)->rowptr.back(); }
void orc_csr_export(void *h, int64_t *rowptr, int64_t *col, double *vN, double *vD, double *vP) {
  Oracle &o = *(Oracle *)h;
  std::copy(o.rowptr.begin(), o.rowptr.end(), rowptr);
  std::copy(o.col.begin(), o.col.end(), col);
  std::copy(o.vN.begin(), o.vN.end(), vN);
  std::copy(o.vD.begin(), o.vD.end(), vD);
  if (vP) std::copy(o.vP.begin(), o.vP.end(), vP);
}

// expansions of every block, 2*Kt doubles per block per kernel
void orc_get_expansions(void *h, int which, double *k1, double *k2) {
  Oracle &o = *(Oracle *)h;
  const auto &A = which == 0 ? o.me1 : o.le1;
  const auto &B = which == 0 ? o.me2 : o.le2;
  const int kt = (o.p + 1) * (o.p + 2) / 2;
  for (int64_t i = 0; i < o.num_blocks; ++i) {
    pack(A[i], k1 + 2 * kt * i);
    pack(B[i], k2 + 2 * kt * i);
  }
}

void orc_quad_export(void *h, double *qp, double *qn, double *qw, double *qs) {
  Oracle &o = *(Oracle *)h;
  std::copy(o.qp.begin(), o.qp.end(), qp);
  std::copy(o.qn.begin(), o.qn.end(), qn);
  std::copy(o.qw.begin(), o.qw.end(), qw);
  std::copy(o.qs.begin(), o.qs.end(), qs);
}

int orc_compute_alpha(void *h, double *alpha) {
  Oracle &o = *(Oracle *)h;
  compute_alpha(o);
  std::copy(o.alpha.begin(), o.alpha.end(), alpha);
  return 0;
}
int orc_vmult(void *h, const double *src, double *dst) {
  vmult(*(Oracle *)h, src, dst);
  return 0;
}
int orc_compute_rhs(void *h, const double *g, double *dst) {
  compute_rhs(*(Oracle *)h, g, dst);
  return 0;
}

// solve_system (bem_problem.cc:1266-1347): alpha, rhs, constraints
// (caller-supplied rows: dst(i) = src(i) - sum c_ij src(j), rhs(i) = inhom(i)),
// GMRES, then scatter sol into phi (Neumann dofs) / dphi (Dirichlet dofs).
int orc_solve_system(void *h, const double *tmp_rhs, const int64_t *cptr, const int64_t *cidx,
                     const double *cval, const double *cinh, const uint8_t *is_c, double *phi,
                     double *dphi, double tol, int restart, int maxit, int *iters, double *relres) {
  Oracle &o = *(Oracle *)h;
  compute_alpha(o);
  std::vector<double> rhs(o.N);
  compute_rhs(o, tmp_rhs, rhs.data());
  Constraints c;
  if (is_c) {
    c.ptr.assign(cptr, cptr + o.N + 1);
    c.idx.assign(cidx, cidx + cptr[o.N]);
    c.val.assign(cval, cval + cptr[o.N]);
    c.inhom.assign(cinh, cinh + o.N);
    c.is_c.assign(is_c, is_c + o.N);
    for (int64_t i = 0; i < o.N; ++i)
      if (is_c[i]) rhs[i] = cinh[i];
  }
  std::vector<double> sol(rhs);
  gmres(o, c, rhs.data(), sol.data(), tol, restart, maxit, iters, relres);
  for (int64_t i = 0; i < o.N; ++i) {
    if (o.dir[i] == 0) phi[i] = sol[i];
    else dphi[i] = sol[i];
  }
  return 0;
}

int orc_dense_assemble(void *h, int rule, double *Nm, double *Dm) {
  Oracle &o = *(Oracle *)h;
  // the quadrature data come with the octree; build it when not done yet
  if (o.nq == 0 || (int64_t)o.qw.size() != o.ncells * o.nq) generate_octree(o);
  dense_assemble(o, rule, Nm, Dm);
  return 0;
}

}  // extern "C"
