// Internal declarations of libpbfmm: host-side setup structures and the
// device-side state shared by the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/pbfmm.h"

namespace pbfmm {

constexpr double kPi = 3.14159265358979323846;
constexpr double kToll = 0.001;  // box adjacency tolerance of bem_fma.cc:1

inline int kt_of(int p) { return (p + 1) * (p + 2) / 2; }

// Runs body(i) for i in [0, n) on up to 16 host threads, `grain` indices at
// a time (dynamic); the caller's results must not depend on the order.
void parallel_for(int64_t n, int64_t grain, const std::function<void(int64_t)> &body);
// Pairs per GEMM chunk of the batched translations (64 GEMM columns): 32 pairs
// when both kernels' expansions travel (2 columns per pair), 64 pairs when
// one combined expansion does.
constexpr int kM2LChunk = 32;
constexpr int kChunk1 = 64;
// padded length of one real expansion vector [Re(0..Kt-1), Im(0..Kt-1)]
inline int kp_of(int p) { return ((2 * kt_of(p) + 15) / 16) * 16; }

// ---------------------------------------------------------------------------
// Mesh and regular quadrature data (host, FP64)
// ---------------------------------------------------------------------------
struct HostMesh {
  int64_t n = 0, ncells = 0;
  // cells [ncells_real, ncells) are mirror images of cells [0, ncells_real)
  // (double body): same DoFs, always integrated with the regular rule
  int64_t ncells_real = 0;
  double mirror_z0 = 0.0;
  int dpc = 4;
  std::vector<double> xyz;      // 3n
  std::vector<double> geom;     // ncells*dpc*3 mapping nodes
  std::vector<int32_t> cdofs;   // ncells*dpc
  std::vector<int32_t> dbl_ptr, dbl_idx;  // double-node sets, each holding the dof itself
  std::vector<uint8_t> dir;
  std::vector<int32_t> d2c_ptr, d2c_idx;  // dof -> real cells (ascending)
};
// Appends the mirror image of every real cell about the plane z = z0
// (BEMProblem::_assemble_system_double_body, bem_problem.cc:957-1135).
void add_mirror_cells(HostMesh &m, double z0);

struct Quadrature2D {
  std::vector<double> pts, w;  // 2*n, n on the unit square, x fastest
};

// Gauss-Legendre on [0,1] (Golub-Welsch free Newton iteration), Telles cubic
// transformation, tensor products, the Duffy split rule; Lagrange Q1/Q2 in
// FE_Q numbering.
void gauss_legendre01(int n, std::vector<double> &x, std::vector<double> &w);
void telles_rule01(int n, double x0, std::vector<double> &x, std::vector<double> &w);
Quadrature2D tensor_rule(const std::vector<double> &x1, const std::vector<double> &w1,
                         const std::vector<double> &x2, const std::vector<double> &w2);
Quadrature2D duffy_split_rule(int n, double sx, double sy);
void lagrange_q(int dpc, double xi, double eta, double *N, double *dNx, double *dNy);
void unit_support_point(int dpc, int j, double &xi, double &eta);

struct QuadData {
  int nq = 0;
  std::vector<double> x, n, jxw, sh;  // per global quad id g = cell*nq+q
};
void compute_quad_data(const HostMesh &m, const Quadrature2D &rule, QuadData &qd);

// ---------------------------------------------------------------------------
// Octree with the reference list semantics, in flat arrays
// ---------------------------------------------------------------------------
struct Tree {
  int L = 0;
  int nb = 0;
  std::vector<int32_t> level, parent, nchild, child, octant;  // child: 8 per block
  std::vector<double> pmin, delta;                            // 3*nb, nb
  std::vector<int32_t> startL, endL;                          // L+1
  std::vector<int32_t> leaves;                                // childlessList order
  std::vector<std::vector<int32_t>> dofs_filled;              // per level
  std::vector<int32_t> nperm, node_beg, node_cnt;
  std::vector<int32_t> qperm, quad_beg, quad_cnt;
  std::vector<int32_t> qblock;  // (L+1) x nquad
  std::vector<int32_t> nsub, sub_beg;
  std::vector<std::vector<int32_t>> nn, il, nil;  // index sub_beg[b]+s
  void center(int b, double c[3]) const {
    for (int k = 0; k < 3; ++k) c[k] = pmin[3 * b + k] + delta[b] / 2.0;
  }
};
void build_tree(const HostMesh &m, const QuadData &qd, int levels, int max_nodes, Tree &t);

// ---------------------------------------------------------------------------
// Near field: leaf-blocked storage.  All rows (nodes) of one leaf share the
// same column set, so values are a dense rows x cols block per leaf and the
// column indices are stored once per leaf.
// ---------------------------------------------------------------------------
struct NearPlan {
  int nleaf = 0;                    // leaves holding nodes
  std::vector<int32_t> leaf_block;  // block id
  std::vector<int32_t> row_beg, rows;   // rows (node ids) per leaf
  std::vector<int32_t> col_beg, cols;   // columns per leaf
  std::vector<int64_t> val_beg;         // offset of the rows x cols block
  std::vector<int32_t> ent_beg;         // entries per leaf
  std::vector<int32_t> ent_cell;
  std::vector<uint64_t> ent_mask;       // quad points of the cell in this entry
  std::vector<uint8_t> ent_near;        // 1: leaf near list (may be singular)
  std::vector<int32_t> ent_colpos;      // dpc per entry
  std::vector<int64_t> work_beg;        // prefix of rows*entries per leaf
  std::vector<int32_t> row_leaf;        // leaf-ordered row -> leaf
  int64_t nnz = 0;
};
void build_near_plan(const HostMesh &m, const QuadData &qd, const Tree &t, NearPlan &np);

// ---------------------------------------------------------------------------
// Far field plan and translation operators
// ---------------------------------------------------------------------------
// Pairs (tgt, src) of one kind of translation grouped by operator class and
// cut into chunks of <= kM2LChunk pairs (op, beg, cnt; level_off[l] ..
// level_off[l+1] are the chunks of level l, M2M: child level, L2L: target
// level) and, over the same pairs, into chunks of <= kChunk1 pairs (*1).
struct ChunkList {
  std::vector<int32_t> op, beg, cnt;  // per chunk: operator index, first pair, pair count
  std::vector<int32_t> op1, beg1, cnt1;
  std::vector<int32_t> tgt, src;      // pairs
  std::vector<int32_t> level_off, level_off1;  // L+2 entries
};

struct FarPlan {
  int p = 0, kt = 0, kp = 0;
  std::vector<int32_t> p2m_leaf;  // all childless blocks
  // P2M as a product with unit multipoles: per leaf (index into p2m_leaf) the
  // sorted DoFs of the cells that have quad points in the leaf (columns
  // p2m_col_beg[li] .. p2m_col_beg[li+1]); per quad point of the leaves, in
  // leaf order, its id, its leaf and the column of each of its cell's DoFs.
  std::vector<int32_t> p2m_col_beg, p2m_col_dof;
  std::vector<int32_t> p2m_qpt, p2m_qleaf, p2m_qcol;
  ChunkList m2m;                  // tgt = parent, src = child; op = level*8 + octant
  ChunkList m2l;                  // op = offset class (all levels)
  ChunkList l2l;                  // tgt = child, src = parent; op = level*8 + octant
  // M2M in launches of two child levels (build_m2m_groups): level_off1[g]
  // .. level_off1[g + 1] are group g's chunks; ops index m2m_ops, whose
  // composite two-level operators follow the (L + 1) x 8 single ones
  ChunkList m2m_g;
  // L2L in launches of two target levels (build_l2l_groups), likewise over
  // l2l_ops; the non-leaf targets of a group's first level keep their M2L
  // part only (their children take the parent's part through composites)
  ChunkList l2l_g;
  std::vector<int64_t> m2l_per_level;
  int n_m2l_ops = 0;
  // M2P pairs (target block, smaller far source block)
  std::vector<int32_t> m2p_tgt, m2p_src;
  std::vector<int64_t> m2p_work_beg;  // prefix of node counts
  std::vector<int32_t> m2p_item_node, m2p_item_src;  // per (pair, target node) evaluation
  // L2P
  std::vector<int32_t> l2p_leaf;  // childless blocks holding nodes
  // operators, transposed (column-major: op[k*kp + row]), FP64, acting on
  // level-scaled expansions: a block of level l with box size delta_l stores
  // ME_n^m / delta_l^n and LE_n^m * delta_l^(n+1), so that every operator
  // entry is O(1) in box units (the unscaled entries of a p = 8 M2L between
  // boxes of size 2e-3 reach 1e40, past FP32's range) and the M2L operator of
  // an offset class is the same at every level
  std::vector<double> m2m_ops;  // (L+1) levels x 8 octants
  std::vector<double> l2l_ops;  // (L+1) x 8
  std::vector<double> m2l_ops;  // n_m2l_ops: one per offset class (unit box)
};
void build_far_plan(const Tree &t, int p, FarPlan &fp);
void build_m2m_groups(const Tree &t, FarPlan &fp);
void build_l2l_groups(const Tree &t, FarPlan &fp);
void build_p2m_plan(const HostMesh &m, const QuadData &qd, const Tree &t, FarPlan &fp);
// R-linear translation operators on [Re, Im] vectors (padded to kp).
void op_m2m(int p, const double d[3], double *opT);  // d = child - parent
void op_l2l(int p, const double d[3], double *opT);  // d = parent - child
void op_m2l(int p, const double d[3], double *opT);  // d = source - target

}  // namespace pbfmm
