// A forest of independent problems as one operator: the host plans of the
// parts (mesh, quadrature data, octree arrays the device uses, near-field and
// far-field plans) concatenated with index offsets.  Chunk lists are regrouped
// level by level, so one M2M or L2L launch per level, and one launch for each
// other phase, covers every part; the operator is block diagonal, as the parts
// never interact.
#include <stdexcept>
#include <string>

#include "engine.h"

namespace pbfmm {

namespace {

template <typename T>
void append_off(std::vector<T> &dst, const std::vector<T> &src, T off, bool keep_negative = false) {
  dst.reserve(dst.size() + src.size());
  for (const T &v : src) dst.push_back(keep_negative && v < 0 ? v : v + off);
}
template <typename T>
void append(std::vector<T> &dst, const std::vector<T> &src) {
  dst.insert(dst.end(), src.begin(), src.end());
}
// prefix array (n+1 entries starting at 0) of a part appended to a merged one
template <typename T>
void append_prefix(std::vector<T> &dst, const std::vector<T> &src) {
  if (dst.empty()) dst.push_back(0);
  const T base = dst.back();
  for (size_t i = 1; i < src.size(); ++i) dst.push_back(src[i] + base);
}

void merge_chunks(const std::vector<const ChunkList *> &cs, const std::vector<int32_t> &op_off,
                  const std::vector<int32_t> &boff, int L, ChunkList &out) {
  out = ChunkList();
  const size_t S = cs.size();
  std::vector<int32_t> pair_off(S);
  for (size_t s = 0; s < S; ++s) {
    pair_off[s] = (int32_t)out.tgt.size();
    append_off(out.tgt, cs[s]->tgt, boff[s]);
    append_off(out.src, cs[s]->src, boff[s]);
  }
  out.level_off.assign(L + 2, 0);
  out.level_off1.assign(L + 2, 0);
  for (int l = 0; l <= L; ++l) {
    out.level_off[l] = (int32_t)out.op.size();
    out.level_off1[l] = (int32_t)out.op1.size();
    for (size_t s = 0; s < S; ++s) {
      const ChunkList &c = *cs[s];
      if ((int)c.level_off.size() != L + 2 || (int)c.level_off1.size() != L + 2)
        throw std::runtime_error("forest: chunk list without level offsets");
      for (int k = c.level_off[l]; k < c.level_off[l + 1]; ++k) {
        out.op.push_back(c.op[k] + op_off[s]);
        out.beg.push_back(c.beg[k] + pair_off[s]);
        out.cnt.push_back(c.cnt[k]);
      }
      for (int k = c.level_off1[l]; k < c.level_off1[l + 1]; ++k) {
        out.op1.push_back(c.op1[k] + op_off[s]);
        out.beg1.push_back(c.beg1[k] + pair_off[s]);
        out.cnt1.push_back(c.cnt1[k]);
      }
    }
  }
  out.level_off[L + 1] = (int32_t)out.op.size();
  out.level_off1[L + 1] = (int32_t)out.op1.size();
}

}  // namespace

void merge_forest(const std::vector<const Engine *> &parts, Engine &e, std::vector<int64_t> &dof_off) {
  if (parts.empty()) throw std::runtime_error("forest: no parts");
  const Engine &a = *parts[0];
  for (const Engine *q : parts) {
    if (q->p != a.p || q->L != a.L || q->mesh.dpc != a.mesh.dpc || q->qd.nq != a.qd.nq ||
        q->sing_order != a.sing_order || q->sing_rule != a.sing_rule)
      throw std::runtime_error("forest: parts differ in order, levels, element or quadrature");
    if (q->mesh.ncells != q->mesh.ncells_real)
      throw std::runtime_error("forest: double-body parts are not supported");
    if (!q->have_dirichlet)
      throw std::runtime_error("forest: every part needs Dirichlet DoFs (pure Neumann parts are not supported)");
  }
  const size_t S = parts.size();
  const int L = a.L, dpc = a.mesh.dpc, nq = a.qd.nq;
  e.p = a.p;
  e.kt = a.kt;
  e.kp = a.kp;
  e.L = L;
  e.dpc = dpc;
  e.nq = nq;
  e.sing_order = a.sing_order;
  e.sing_rule = a.sing_rule;
  e.have_dirichlet = true;
  HostMesh &m = e.mesh;
  m = HostMesh();
  m.dpc = dpc;
  e.qd = QuadData();
  e.qd.nq = nq;
  Tree &t = e.tree;
  t = Tree();
  t.L = L;
  NearPlan &np = e.near;
  np = NearPlan();
  FarPlan &fp = e.far;
  fp = FarPlan();
  fp.p = a.far.p;
  fp.kt = a.far.kt;
  fp.kp = a.far.kp;
  fp.m2l_per_level.assign(L + 1, 0);
  dof_off.assign(S + 1, 0);
  std::vector<int32_t> boff(S), m2m_off(S), m2l_off(S);
  std::vector<const ChunkList *> cm2m(S), cm2l(S), cl2l(S);
  for (size_t s = 0; s < S; ++s) {
    const Engine &q = *parts[s];
    const int32_t doff = (int32_t)m.n, coff = (int32_t)m.ncells, bo = t.nb;
    const int32_t qoff = coff * nq;
    boff[s] = bo;
    dof_off[s] = doff;
    // mesh
    append(m.xyz, q.mesh.xyz);
    append(m.geom, q.mesh.geom);
    append_off(m.cdofs, q.mesh.cdofs, doff);
    if (m.dbl_ptr.empty()) m.dbl_ptr.push_back(0);
    {
      const int32_t base = m.dbl_ptr.back();
      for (size_t i = 1; i < q.mesh.dbl_ptr.size(); ++i) m.dbl_ptr.push_back(q.mesh.dbl_ptr[i] + base);
    }
    append_off(m.dbl_idx, q.mesh.dbl_idx, doff);
    append(m.dir, q.mesh.dir);
    append_prefix(m.d2c_ptr, q.mesh.d2c_ptr);
    append_off(m.d2c_idx, q.mesh.d2c_idx, coff);
    m.n += q.mesh.n;
    m.ncells += q.mesh.ncells;
    m.ncells_real += q.mesh.ncells_real;
    // quadrature data (global quad id = cell * nq + q)
    append(e.qd.x, q.qd.x);
    append(e.qd.n, q.qd.n);
    append(e.qd.jxw, q.qd.jxw);
    append(e.qd.sh, q.qd.sh);
    // octree arrays the device reads
    append(t.level, q.tree.level);
    append_off(t.parent, q.tree.parent, bo, true);
    append(t.nchild, q.tree.nchild);
    append_off(t.child, q.tree.child, bo, true);
    append(t.octant, q.tree.octant);
    append(t.pmin, q.tree.pmin);
    append(t.delta, q.tree.delta);
    append_off(t.leaves, q.tree.leaves, bo);
    append_off(t.node_beg, q.tree.node_beg, (int32_t)t.nperm.size());
    append(t.node_cnt, q.tree.node_cnt);
    append_off(t.nperm, q.tree.nperm, doff);
    append_off(t.quad_beg, q.tree.quad_beg, (int32_t)t.qperm.size());
    append(t.quad_cnt, q.tree.quad_cnt);
    append_off(t.qperm, q.tree.qperm, qoff);
    t.nb += q.tree.nb;
    // near field
    const int32_t leaf_off = np.nleaf;
    append_off(np.leaf_block, q.near.leaf_block, bo);
    append_prefix(np.row_beg, q.near.row_beg);
    append_off(np.rows, q.near.rows, doff);
    append_prefix(np.col_beg, q.near.col_beg);
    append_off(np.cols, q.near.cols, doff);
    append_off(np.val_beg, q.near.val_beg, np.nnz);
    append_prefix(np.ent_beg, q.near.ent_beg);
    append_off(np.ent_cell, q.near.ent_cell, coff);
    append(np.ent_mask, q.near.ent_mask);
    append(np.ent_near, q.near.ent_near);
    append(np.ent_colpos, q.near.ent_colpos);
    append_prefix(np.work_beg, q.near.work_beg);
    append_off(np.row_leaf, q.near.row_leaf, leaf_off);
    np.nleaf += q.near.nleaf;
    np.nnz += q.near.nnz;
    // far field
    const int32_t p2m_leaf_off = (int32_t)fp.p2m_leaf.size();
    const int32_t col_off = (int32_t)fp.p2m_col_dof.size();
    append_off(fp.p2m_leaf, q.far.p2m_leaf, bo);
    append_prefix(fp.p2m_col_beg, q.far.p2m_col_beg);
    append_off(fp.p2m_col_dof, q.far.p2m_col_dof, doff);
    append_off(fp.p2m_qpt, q.far.p2m_qpt, qoff);
    append_off(fp.p2m_qleaf, q.far.p2m_qleaf, p2m_leaf_off);
    append_off(fp.p2m_qcol, q.far.p2m_qcol, col_off);
    m2m_off[s] = (int32_t)(fp.m2m_ops.size() / ((size_t)fp.kp * fp.kp));
    // the single-level operators only: the batch runs M2M and L2L level by
    // level, without the parts' composite two-level ones (build_m2m_groups,
    // build_l2l_groups)
    const size_t nsingle = (size_t)(L + 1) * 8 * fp.kp * fp.kp;
    fp.m2m_ops.insert(fp.m2m_ops.end(), q.far.m2m_ops.begin(),
                      q.far.m2m_ops.begin() + std::min(q.far.m2m_ops.size(), nsingle));
    fp.l2l_ops.insert(fp.l2l_ops.end(), q.far.l2l_ops.begin(),
                      q.far.l2l_ops.begin() + std::min(q.far.l2l_ops.size(), nsingle));
    m2l_off[s] = fp.n_m2l_ops;
    append(fp.m2l_ops, q.far.m2l_ops);
    fp.n_m2l_ops += q.far.n_m2l_ops;
    for (int l = 0; l <= L; ++l) fp.m2l_per_level[l] += q.far.m2l_per_level[l];
    append_off(fp.m2p_tgt, q.far.m2p_tgt, bo);
    append_off(fp.m2p_src, q.far.m2p_src, bo);
    append_prefix(fp.m2p_work_beg, q.far.m2p_work_beg);
    append_off(fp.m2p_item_node, q.far.m2p_item_node, doff);
    append_off(fp.m2p_item_src, q.far.m2p_item_src, bo);
    append_off(fp.l2p_leaf, q.far.l2p_leaf, bo);
    cm2m[s] = &q.far.m2m;
    cm2l[s] = &q.far.m2l;
    cl2l[s] = &q.far.l2l;
  }
  dof_off[S] = m.n;
  if (np.row_beg.empty()) np.row_beg.push_back(0);
  if (np.col_beg.empty()) np.col_beg.push_back(0);
  if (np.ent_beg.empty()) np.ent_beg.push_back(0);
  if (np.work_beg.empty()) np.work_beg.push_back(0);
  if (fp.p2m_col_beg.empty()) fp.p2m_col_beg.push_back(0);
  if (fp.m2p_work_beg.empty()) fp.m2p_work_beg.push_back(0);
  // M2M and L2L operators are indexed level * 8 + octant within a part
  merge_chunks(cm2m, m2m_off, boff, L, fp.m2m);
  merge_chunks(cm2l, m2l_off, boff, L, fp.m2l);
  merge_chunks(cl2l, m2m_off, boff, L, fp.l2l);
  e.n = m.n;
}

namespace {
template <typename V, typename T>
void in_range(const char *what, const V &v, T lo, T hi, bool allow_neg1 = false) {
  for (const auto &x : v)
    if (!((x >= lo && x < hi) || (allow_neg1 && x == -1)))
      throw std::runtime_error(std::string("plan check: ") + what + " index out of range");
}
template <typename V>
void prefix_ok(const char *what, const V &v, size_t n_expect, int64_t last) {
  if (v.size() != n_expect || (n_expect && (v.front() != 0 || (int64_t)v.back() != last)))
    throw std::runtime_error(std::string("plan check: ") + what + " prefix array malformed");
  for (size_t i = 1; i < v.size(); ++i)
    if (v[i] < v[i - 1]) throw std::runtime_error(std::string("plan check: ") + what + " not monotone");
}
void size_is(const char *what, size_t got, size_t want) {
  if (got != want) throw std::runtime_error(std::string("plan check: ") + what + " has the wrong size");
}
void chunks_ok(const char *what, const ChunkList &c, int nops, int nb) {
  const int np = (int)c.tgt.size();
  size_is(what, c.src.size(), np);
  in_range(what, c.tgt, 0, nb);
  in_range(what, c.src, 0, nb);
  in_range(what, c.op, 0, nops);
  in_range(what, c.op1, 0, nops);
  for (size_t k = 0; k < c.op.size(); ++k)
    if (c.beg[k] < 0 || c.cnt[k] < 0 || c.beg[k] + c.cnt[k] > np || c.cnt[k] > kM2LChunk)
      throw std::runtime_error(std::string("plan check: ") + what + " chunk out of range");
  for (size_t k = 0; k < c.op1.size(); ++k)
    if (c.beg1[k] < 0 || c.cnt1[k] < 0 || c.beg1[k] + c.cnt1[k] > np || c.cnt1[k] > kChunk1)
      throw std::runtime_error(std::string("plan check: ") + what + " chunk out of range");
  if (c.level_off.empty() || c.level_off.back() != (int)c.op.size() || c.level_off1.back() != (int)c.op1.size())
    throw std::runtime_error(std::string("plan check: ") + what + " level offsets malformed");
  for (size_t l = 1; l < c.level_off.size(); ++l)
    if (c.level_off[l] < c.level_off[l - 1] || c.level_off1[l] < c.level_off1[l - 1])
      throw std::runtime_error(std::string("plan check: ") + what + " level offsets not monotone");
}
}  // namespace

void validate_plans(const Engine &e) {
  const HostMesh &m = e.mesh;
  const int64_t n = m.n, nc = m.ncells, dpc = m.dpc, nq = e.qd.nq, NQ = nc * nq;
  const Tree &t = e.tree;
  const int nb = t.nb;
  size_is("xyz", m.xyz.size(), 3 * n);
  size_is("geom", m.geom.size(), nc * dpc * 3);
  size_is("cell dofs", m.cdofs.size(), nc * dpc);
  in_range("cell dofs", m.cdofs, (int64_t)0, n);
  prefix_ok("double nodes", m.dbl_ptr, n + 1, (int64_t)m.dbl_idx.size());
  in_range("double nodes", m.dbl_idx, (int64_t)0, n);
  size_is("dirichlet mask", m.dir.size(), n);
  prefix_ok("dof to cells", m.d2c_ptr, n + 1, (int64_t)m.d2c_idx.size());
  in_range("dof to cells", m.d2c_idx, (int64_t)0, nc);
  size_is("quad points", e.qd.x.size(), 3 * NQ);
  size_is("quad normals", e.qd.n.size(), 3 * NQ);
  size_is("quad weights", e.qd.jxw.size(), NQ);
  size_is("quad shapes", e.qd.sh.size(), NQ * dpc);
  size_is("tree parent", t.parent.size(), nb);
  in_range("tree parent", t.parent, -1, nb, true);
  size_is("tree children", t.child.size(), (size_t)8 * nb);
  in_range("tree children", t.child, -1, nb, true);
  size_is("tree pmin", t.pmin.size(), (size_t)3 * nb);
  size_is("tree delta", t.delta.size(), nb);
  size_is("tree octant", t.octant.size(), nb);
  size_is("tree nchild", t.nchild.size(), nb);
  in_range("tree nodes", t.nperm, (int64_t)0, n);
  in_range("tree quads", t.qperm, (int64_t)0, NQ);
  for (int b = 0; b < nb; ++b) {
    if (t.node_beg[b] < 0 || t.node_cnt[b] < 0 || t.node_beg[b] + t.node_cnt[b] > (int64_t)t.nperm.size() ||
        t.quad_beg[b] < 0 || t.quad_cnt[b] < 0 || t.quad_beg[b] + t.quad_cnt[b] > (int64_t)t.qperm.size() ||
        t.nchild[b] < 0 || t.nchild[b] > 8)
      throw std::runtime_error("plan check: tree block ranges out of range");
  }
  const NearPlan &np = e.near;
  const int nl = np.nleaf;
  size_is("near leaves", np.leaf_block.size(), nl);
  in_range("near leaves", np.leaf_block, 0, nb);
  prefix_ok("near rows", np.row_beg, (size_t)nl + 1, (int64_t)np.rows.size());
  prefix_ok("near cols", np.col_beg, (size_t)nl + 1, (int64_t)np.cols.size());
  prefix_ok("near entries", np.ent_beg, (size_t)nl + 1, (int64_t)np.ent_cell.size());
  size_is("near values", np.val_beg.size(), nl);
  in_range("near rows", np.rows, (int64_t)0, n);
  in_range("near cols", np.cols, (int64_t)0, n);
  in_range("near cells", np.ent_cell, (int64_t)0, nc);
  size_is("near entry masks", np.ent_mask.size(), np.ent_cell.size());
  size_is("near entry flags", np.ent_near.size(), np.ent_cell.size());
  size_is("near entry columns", np.ent_colpos.size(), np.ent_cell.size() * dpc);
  size_is("near work", np.work_beg.size(), (size_t)nl + 1);
  size_is("near row leaf", np.row_leaf.size(), np.rows.size());
  in_range("near row leaf", np.row_leaf, 0, std::max(nl, 1));
  int64_t vend = 0;
  for (int l = 0; l < nl; ++l) {
    const int64_t R = np.row_beg[l + 1] - np.row_beg[l], C = np.col_beg[l + 1] - np.col_beg[l];
    if (np.val_beg[l] != vend) throw std::runtime_error("plan check: near value offsets malformed");
    vend += R * C;
    for (int k = np.ent_beg[l] * dpc; k < np.ent_beg[l + 1] * dpc; ++k)
      if (np.ent_colpos[k] < 0 || np.ent_colpos[k] >= C) throw std::runtime_error("plan check: near column position");
    if (np.work_beg[l + 1] - np.work_beg[l] != R * (np.ent_beg[l + 1] - np.ent_beg[l]))
      throw std::runtime_error("plan check: near work offsets malformed");
  }
  if (vend != np.nnz) throw std::runtime_error("plan check: near value count");
  const FarPlan &fp = e.far;
  const int kp = fp.kp, npl = (int)fp.p2m_leaf.size();
  in_range("p2m leaves", fp.p2m_leaf, 0, nb);
  prefix_ok("p2m columns", fp.p2m_col_beg, (size_t)npl + 1, (int64_t)fp.p2m_col_dof.size());
  in_range("p2m columns", fp.p2m_col_dof, (int64_t)0, n);
  in_range("p2m quad points", fp.p2m_qpt, (int64_t)0, NQ);
  size_is("p2m quad leaves", fp.p2m_qleaf.size(), fp.p2m_qpt.size());
  size_is("p2m quad columns", fp.p2m_qcol.size(), fp.p2m_qpt.size() * dpc);
  for (size_t i = 0; i < fp.p2m_qpt.size(); ++i) {
    const int li = fp.p2m_qleaf[i];
    if (li < 0 || li >= npl) throw std::runtime_error("plan check: p2m quad leaf");
    for (int j = 0; j < dpc; ++j) {
      const int c = fp.p2m_qcol[i * dpc + j];
      if (c < fp.p2m_col_beg[li] || c >= fp.p2m_col_beg[li + 1]) throw std::runtime_error("plan check: p2m column");
    }
  }
  const size_t op_sz = (size_t)kp * kp;
  if (fp.m2m_ops.size() % op_sz || fp.l2l_ops.size() % op_sz || fp.m2l_ops.size() != (size_t)fp.n_m2l_ops * op_sz)
    throw std::runtime_error("plan check: operator storage");
  chunks_ok("m2m", fp.m2m, (int)(fp.m2m_ops.size() / op_sz), nb);
  chunks_ok("l2l", fp.l2l, (int)(fp.l2l_ops.size() / op_sz), nb);
  chunks_ok("m2l", fp.m2l, fp.n_m2l_ops, nb);
  size_is("m2p items", fp.m2p_item_src.size(), fp.m2p_item_node.size());
  in_range("m2p nodes", fp.m2p_item_node, (int64_t)0, n);
  in_range("m2p sources", fp.m2p_item_src, 0, nb);
  in_range("l2p leaves", fp.l2p_leaf, 0, nb);
}

}  // namespace pbfmm
