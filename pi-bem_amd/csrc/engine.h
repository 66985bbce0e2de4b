// Device-resident state of one handle and the host orchestration of the
// BEM-FMM kernels (engine.hip).
#pragma once
#include <stdexcept>

#include "pbfmm_impl.h"

namespace pbfmm {

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T *as() const { return reinterpret_cast<T *>(p); }
};

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Chunk list of one batched translation (see FarPlan::ChunkList) on the device.
struct ChunkDev {
  DevBuf op, beg, cnt, op1, beg1, cnt1, tgt, src;
};


// Per-precision copies of the far-field operators and expansion storage.
template <typename T>
struct FarDev {
  using value_type = T;
  DevBuf m2m, l2l, m2l;  // transposed operators
  DevBuf me, le;         // [nb][2 kernels][kp]
  DevBuf qx, qn, qw, qsh;  // regular quadrature data (FP64 copy only)
  DevBuf vals;             // near-field values (N, D) pairs, leaf-blocked
  DevBuf unit;             // unit multipoles: [column][2 kernels][kp]
  // the operator's single-value copies: N on Neumann / -D on Dirichlet
  // columns, U1 on Neumann / -U2 on Dirichlet DoFs ([column][kp])
  DevBuf opvals, opunit;
  DevBuf opvals_p;  // FP32 only: opvals in row pairs (k_round_f32_pairs)
  // opvals in column vectors of 16 bytes (k_round_cols), per-leaf offsets
  DevBuf opvals_q, qval_beg;
  DevBuf slots;  // slot M2L: [pair][2 Kt], on first use
};

// Mixed precision: the relaxed scheme spends its first steps (about a
// third) on the FP64 operator; refinement runs its steps on the FP32 operator
// and, with the first inner solve's Krylov space recycled into the second,
// about as many of them as an FP64 solve, plus two FP64 residuals.  With N
// steps, refinement is cheaper once the FP64 / FP32 cost ratio exceeds about
// N / (N - 3): 1.1 for N ~ 40.  Measured: 24,578 DoFs, leaf 64, ratio 1.02:
// relaxed 5.00 ms, refinement 5.21 ms; 393,218 DoFs, leaf 20, ratio 1.78:
// refinement 85.6 ms, relaxed 98.0 ms.
constexpr double kRelaxCostRatio = 1.15;

struct Engine {
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int device = 0;
  int p = 6, kt = 0, kp = 0, L = 10;
  int kpc = 0;     // compact unit-multipole length of the operator (set with the unit multipoles)
  DevBuf d_cmap;   // its rows' places in a kp-row expansion (-1: padding)
  int dpc = 4, nq = 16, sing_order = 5;
  int sing_rule = 0;  // 0 Telles tensor rule, 1 Duffy split (PBFMM_SING_*)
  int64_t n = 0;
  // host plans
  HostMesh mesh;
  QuadData qd;
  Tree tree;
  NearPlan near;
  FarPlan far;
  bool have_dirichlet = false;
  bool neumann_mean = false;  // pure Neumann: A0 x + mean(x) 1 (set during GMRES)
  DevBuf d_ones;
  // device: mesh
  DevBuf d_xyz, d_geom, d_cdofs, d_dbl_ptr, d_dbl_idx, d_dir;
  // device: tree
  DevBuf d_pmin, d_delta, d_parent, d_nchild, d_child, d_octant, d_node_beg, d_node_cnt, d_nperm,
      d_quad_beg, d_quad_cnt, d_qperm;
  // device: plans
  DevBuf d_p2m_leaf, d_m2p_node, d_m2p_isrc;
  DevBuf d_p2m_col_beg, d_p2m_col_dof;
  DevBuf d_is_leaf;  // per block: 1 for a leaf (P2M stores its ME)
  ChunkDev m2m_cl, m2l_cl, l2l_cl;
  ChunkDev m2m_g_cl;  // M2M two child levels per launch (FarPlan::m2m_g)
  ChunkDev l2l_g_cl;  // L2L two target levels per launch (FarPlan::l2l_g)
  DevBuf d_pval_beg;  // FP32 operator values: per-leaf offsets of the row-pair layout
  DevBuf d_row_leaf, d_row_beg, d_rows, d_col_beg, d_cols, d_val_beg, d_ent_beg, d_ent_cell,
      d_ent_mask, d_ent_near, d_ent_colpos, d_leaf_block;
  DevBuf d_sing;  // singular rules: per js [w, N(dpc), dNx(dpc), dNy(dpc)] per point
  int sing_nq = 0;
  FarDev<double> f64;
  FarDev<float> f32;
  bool have_f32 = false;
  // operator vectors
  DevBuf d_alpha, d_sp, d_sd, d_outN, d_outD, d_in, d_out, d_red;
  DevBuf d_cgs;  // partial sums of the Arnoldi step's dots (kDotParts per column) and norms (one per block)
  int max_leaf_rows = 1 << 30;  // largest block of the preconditioner
  // pinned host copy of the Hessenberg slots (asynchronous device-to-host
  // copies) and the event that marks a copy done
  double *h_pinned = nullptr;
  size_t h_pinned_n = 0;
  hipEvent_t h_ready = nullptr;
  // coarsest levels whose expansions are read (MEs: M2L and M2P sources; LEs:
  // M2L targets); M2M into coarser levels and L2L out of them are skipped in
  // the operator pipeline.  -1 until computed from the plan.
  int me_lo = -1, le_lo = -1;
  void consumer_levels();
  DevBuf d_c_ptr, d_c_idx, d_c_val, d_c_inh, d_c_is;
  bool have_constraints = false;
  std::vector<int32_t> hc_ptr, hc_idx;  // host copies of the constraint rows
  std::vector<double> hc_val, hc_inh;
  std::vector<uint8_t> hc_is;
  // near-field block-Jacobi preconditioner: blocks of the near leaves that
  // share a parent block (at most prec_rows rows, leaf after leaf; a larger
  // leaf alone), per block the inverse of its rows x rows near-field block
  // (inverted in FP64; d_pinv[0] FP64 in column pairs for FP64 solves,
  // d_pinv[1] FP32 in column quads for mixed solves, at d_pvec_beg[copy]
  // [block]: pinv_vec_index; pinv_beg[block]: the R x R blocks' own
  // offsets, as assembled); prow_beg / prows / prow_blk: the
  // blocks' row ranges, rows in block order, block of each of them
  // Row cap of a preconditioner block (k_bjacobi_wg takes up to 128, one or
  // two rows per lane; pbfmm_set_preconditioner_rows).  Larger blocks cut
  // GMRES steps and cost bytes per application (n x rows values).  Measured
  // at 393,218 DoFs (Q2, order 8), GMRES to 1e-8, mixed / FP64 ms (steps):
  // leaf size 64, 32 / 48 / 64 / 96 / 128 rows: 55.4 (59) / 54.5 (58) / 54.6
  // (58) / 52.4 (55) / 52.0 (54) and 91.7 / 90.0 / 90.1 / 87.5 / 86.5; leaf
  // 20 (siblings hold at most ~64 rows): 91.6 / 86.5 / 84.4 / 84.5 / 84.6.
  int prec_rows = 128;
  bool have_prec = false;
  bool prec_f32 = false;  // set for the duration of a mixed solve: apply d_pinv[1]
  int nprec = 0;
  std::vector<int64_t> pinv_beg;
  std::vector<int32_t> prow_beg, prows;
  DevBuf d_pinv_beg, d_pvec_beg[2], d_pinv[2], d_ptmp, d_ptmp2, d_prow_beg, d_prows, d_prow_blk;
  void build_preconditioner();
  std::vector<int32_t> prec_pos;  // block-ordered row -> leaf-ordered near row
  void prec_blocks(std::vector<double> &blk);
  void apply_prec(const double *v, double *z, int R = 1);  // R right-hand sides, n apart
  DevBuf d_rhs_sum;  // per right-hand side sums of the lockstep operator's mean term
  // GMRES workspace
  DevBuf d_V, d_w, d_r, d_x, d_h;
  DevBuf d_rec, d_zr, d_qr;  // recycled Krylov space: old basis and C scratch, coefficients, Q
  bool recycle = true;       // refinement's second inner solve recycles the first's space
  int gm_restart = 0;
  // timings of the last matvec (ms)
  double last_ms[9] = {0};
  bool time_phases = false;
  hipEvent_t ev[12] = {};
  bool slot_marked = false;  // the last phase-timed M2L ran as slots + sum (event 11 between)
  void mark(int i);
  void collect_phase_times();

  void init_device(int dev, void *stream);
  ~Engine();
  void upload_mesh_and_tree();
  void upload_near_plan();
  void upload_quad(bool want_f32);
  void upload_far(bool want_f32);
  void build_unit_multipoles(bool want_f32);
  void assemble_near(bool want_f32);
  template <typename T>
  void matvec(const double *phi, const double *dphi, double *outN, double *outD);
  // the BEMProblem operator's FMM: out = N.(x.neu) - D.(x.dir) with one
  // combined expansion per block and single-value near field / unit multipoles
  template <typename T>
  void matvec_combined(const double *x, double *out);
  // BEMProblem::vmult's FMM and epilogue in fewer launches (used by vmult
  // while phase timing is off): dst = out + m2p + L2P + alpha x, with the
  // SpMV and M2P run beside translation launches per side mode (0 none,
  // 1 SpMV beside an M2M level and M2P beside an L2L level, 2 SpMV beside an
  // M2M level and M2P beside M2L, 3 both beside M2L)
  template <typename T>
  void matvec_fused(const double *x, double *out, double *dst, int side);
  // side fusion mode per precision ([0] FP64, [1] FP32): -1 until timed on
  // first use, then the fastest of the plans kPlans
  // Plans are bit sets: bits 0-1 the side placement above (0..3); bit 2
  // (plans 4, 5 with placements 0, 1) M2L into per-pair slots and a
  // per-target sum (k_translate<., ., 3, 1> + k_m2l_reduce) instead of atomic
  // adds into the targets; bit 3 (plans 8, 12 with placement 0) the SpMV on a
  // second stream beside the whole FMM, joined before L2P; bit 4 (plans 24,
  // 28) that fork after the M2M levels instead; bit 5 (plans 56, 60) M2P
  // after the SpMV on the second stream too; bit 6 (plans 88, 92) M2P before
  // the SpMV there instead.
  static constexpr int kPlans[] = {0, 1, 2, 3, 4, 5, 8, 12, 24, 28, 56, 60, 88, 92};
  static bool valid_plan(int m) {
    for (int v : kPlans)
      if (v == m) return true;
    return false;
  }
  hipStream_t side_stream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // slot M2L: targets, their slot ranges (CSR) and each pair's slot, the
  // slots of one target consecutive (its pairs in ascending order)
  DevBuf d_red_tgt, d_red_beg, d_slot_pos;
  int n_red_tgt = 0;
  template <typename T>
  T *m2l_slots();
  template <typename T>
  void launch_m2l_slots(const T *ops, const T *IN, T *LE);
  int side_fusion[2] = {-1, -1};
  // One Arnoldi step's operator (preconditioner + vmult, step_op) captured
  // into a HIP graph per step precision ([0] FP64, [1] FP32) on the fixed
  // buffers d_vn (the step's input, the last basis vector) -> d_wk, used when
  // replaying it beat direct launches (graph_use: -1 not timed, 0 direct,
  // 1 graph), recaptured when a choice or buffer it holds changes.
  int graph_use[2] = {-1, -1};
  bool graph_enabled = true;  // pbfmm_set_graph_mode 0 turns graphs off
  hipGraphExec_t op_graph[2] = {nullptr, nullptr};
  std::vector<int64_t> op_graph_key[2];
  DevBuf d_vn, d_wk;
  float graph_ms[2][2] = {{0, 0}, {0, 0}};  // timed ms per step: direct, graph
  int64_t constraint_version = 0;
  void step_op(const double *vin, double *w, int precision);
  void step_op_graph(int precision);
  std::vector<int64_t> graph_key(int precision) const;
  void drop_graphs();
  int use_side_fusion(bool f32);
  // near-field SpMV kernel per precision ([0] FP64, [1] FP32) and storage
  // (pairs, single values): 0 workgroup per leaf with an LDS copy of the
  // columns, 1 wavefront per leaf with 2 rows in flight, 2 with 4 rows;
  // chosen by timing all three at set-up
  int spmv_variant[2][2] = {{0, 0}, {0, 0}};  // [precision][NK - 1]
  template <typename T, int NK>
  void tune_spmv();
  template <typename T, int NK>
  void launch_spmv(int variant, const double *phi, const double *dphi, double *outN, double *outD, int R = 1,
                   hipStream_t st = nullptr);
  template <typename T, int NK>
  void generate(const double *phi, const double *dphi);
  template <typename T, int NK>
  void products(const double *phi, const double *dphi, double *outN, double *outD);
  // BEMProblem operator (device vectors of n doubles)
  void vmult(const double *src, double *dst, int precision);
  void compute_alpha();
  void compute_rhs(const double *g, double *rhs);
  int gmres(const double *rhs, double *x, double tol, int restart, int maxit, int precision,
            int *iters, double *relres);
  // mixed precision: the relaxed scheme (FP64 operator, then FP32, in one
  // Krylov space) when the FP64 operator is not much dearer than the FP32
  // one, else iterative refinement; the cost ratio is timed once
  double op_cost_ratio = 0;
  // 0 chosen by the timed cost ratio, 1 relaxed, 2 refinement (recycling),
  // 3 without recycling, 4 recycling with an FP32 basis in the inner solves
  int mixed_scheme = 0;
  // GMRES (gmres_core): the second Gram-Schmidt projection of a step runs
  // when the first left |w1| < reorth_eta |w0| (pbfmm_set_gmres_reorth);
  // >= 1: always (CGS2), 0: never (CGS).  At 393,218 DoFs the first
  // projection leaves 0.12-0.95 of |w0|, and a second one changed |w| by
  // less than 1e-6.
  static constexpr double kReorthEta = 0.05;
  double reorth_eta = kReorthEta;
  double inner_reduction = 0;  // refinement's inner residual reduction (0: the default)
  double relax_switch = 0;     // relaxed scheme's FP32 switch over tol ||b|| (0: the default)
  double inner_target(double tol, double bn, double rn) const;
  double relax_sw() const;
  DevBuf d_cost;
  bool mixed_relaxed();
  double norm2(const double *v);
  void sync();

  // ---- forest of independent parts (forest.cpp) ----
  std::vector<int64_t> seg_off;  // DoF range of each part (n_parts + 1)
  DevBuf d_seg_off, d_seg, d_segv, d_fb;
  // reduction chunks aligned to the parts: chunk c covers rch_beg[c] ..
  // rch_beg[c+1], the chunks of part s are sch_off[s] .. sch_off[s+1]
  int n_rch = 0;
  DevBuf d_rch_beg, d_sch_off, d_seg_active, d_hs, d_ys;
  void forest_segments();        // device copies of seg_off and the DoF -> part map
  void seg_norms(const double *v, std::vector<double> &out);
  // GMRES with every part's own relative residual <= tol: one Arnoldi
  // process per part inside shared basis vectors (segmented dots, norms and
  // projections), one operator application for all parts per step; relres
  // gets the true FP64 relative residual of each part
  int gmres_forest(const double *b, double *x, double tol, int restart, int maxit, int precision, int *iters,
                   double *relres);

  // ---- several right-hand sides of one operator in lockstep ----
  // Vectors hold R right-hand sides one after another (rhs r at r * n).
  // Translations run once for all of them over pair lists expanded to
  // (rhs, pair) with block ids rhs * nb + block, on expansions stored
  // rhs-major ([rhs][block][kp]); the other phases loop over the rhs.
  int mr_R = 0, mr_nch = 0;
  ChunkList mr_m2m_h, mr_m2l_h, mr_l2l_h;
  ChunkDev mr_m2m, mr_m2l, mr_l2l;
  DevBuf mr_me64, mr_le64, mr_me32, mr_le32, mr_out;
  DevBuf mr_seg_off, mr_seg, mr_rch, mr_sch, mr_act, mr_segv;
  DevBuf mr_V, mr_w, mr_r, mr_d, mr_t1, mr_t2, mr_hs, mr_ys, mr_red;
  int mr_restart = 0;
  void multi_setup(int R);
  template <typename T>
  void matvec_combined_multi(const double *x, double *out, int R);
  void vmult_multi(const double *src, double *dst, int R, int precision);
  void apply_prec_multi(const double *v, double *z, int R);
  int gmres_multi(const double *b, double *x, int R, double tol, int restart, int maxit, int precision,
                  int *iters, double *relres);

  // ---- post-solve (postsolve.hip): BEMProblem::compute_normals/gradients,
  // pressures and body integrals, over the real cells only ----
  struct PostSolve {
    bool ready = false;
    int nq = 0, nq1 = 0;
    int64_t ncells = 0, nnz = 0;
    DevBuf tab, tab1;                       // 2-D rule [w, N, dNx, dNy] per point; 1-D rule [t, w]
    DevBuf geom, cdofs, xyz;                // real cells; support points
    DevBuf d2c_ptr, d2c_cell, d2c_loc;      // dof -> (cell, local index), cells ascending
    DevBuf m_ptr, m_col, m_val, m_diag, m_pos;  // scalar mass matrix (CSR), pos[c][i][j]
    DevBuf loc;                             // per-cell local matrices / vectors
    DevBuf b, x, r, z, pv, qv, cg, part;    // CG on three components at once
    DevBuf in1, in2, out, aux, mask;        // staging of host arrays
    DevBuf grad;                            // last compute_gradients result (3n)
    bool have_grad = false;
    int last_iters = 0;
  } ps;
  void ps_setup(int quad_order);
  // L2 projection onto the vector element: mode 0 cell normals, 1 full
  // gradient n dphi_dn + surface gradient of phi, 2 surface gradient of in1.
  // Device vectors; out has 3n entries, component-major.
  void ps_project(int mode, const double *in1, const double *in2, double *out);
  void ps_hydrostatic(double gravity, double density, double z0, double *p);
  void ps_hydrodynamic(double density, const double *wind, const double *grad, double *p);
  // sums over masked real cells: [0..2] force, [3..5] moment about cog,
  // [6..8] int x p, [9] int p, [10..12] -int x_d n_d (volume)
  void ps_integrals(const uint8_t *mask, const double *p, const double cog[3], double out[13]);
  // force and moment of ps_integrals into out[0..5] (device), partial sums in
  // the device's fixed scratch array: no allocation (mask NULL: all cells)
  void ps_force_moment_noalloc(const uint8_t *mask, const double *p, const double cog[3], double *out);
  void ps_elevation(const int32_t *faces, int64_t nfaces, const double *p, double rho_g, double *out);
};

// Fixed device scratch of the allocation-free body integrals (postsolve.hip):
// its address on device dev is looked up once, with dev current.
void ps_scratch_init(int dev);
int64_t ps_scratch_max_cells();

// Host plans of independent problems merged into one block-diagonal
// operator (forest.cpp); dof_off receives the DoF range of each part.
void merge_forest(const std::vector<const Engine *> &parts, Engine &e, std::vector<int64_t> &dof_off);
// Bounds of every index array of a set-up plan against what it indexes;
// throws std::runtime_error on the first violation.
void validate_plans(const Engine &e);

}  // namespace pbfmm
