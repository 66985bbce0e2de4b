/*
 * pbfmm.h -- C ABI of the MI355X BEM-FMM library (libpbfmm.so).
 *
 * Drop-in boundary for pi-BEM's fast multipole path: the BEMFMA<3> method set
 * (include/bem_fma.h:74-202) and the FMA branch of BEMProblem<3>
 * (include/bem_problem.h:131-242), re-expressed with plain pointers and sizes
 * so that a thin C++ adapter class BEMFMA<3> can forward to it from the
 * existing Driver.  Every call returns 0 on success and a negative status on
 * failure; pbfmm_last_error() then holds the message (the reference reports
 * the same conditions through AssertThrow / ExcMessage exceptions).
 *
 * Vectors are host arrays of n_dofs doubles in the caller's DoF numbering
 * unless a name ends in _dev (device pointers on the handle's device).  The
 * library owns all device memory; calls are ordered on the handle's stream;
 * a handle is used from one host thread at a time (the reference methods are
 * not reentrant either: bem_fma.h:256-293 mutates mutable members).
 */
#ifndef PBFMM_H
#define PBFMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBFMM_OK 0
#define PBFMM_ERR_ARG -1        /* invalid argument (null pointer, size, range) */
#define PBFMM_ERR_STATE -2      /* call out of order (e.g. products before octree) */
#define PBFMM_ERR_HIP -3        /* HIP runtime error; the GPU path is mandatory */
#define PBFMM_ERR_ALLOC -4      /* device or host allocation failed */
#define PBFMM_ERR_UNSUPPORTED -5

#define PBFMM_FP64 0  /* every phase in FP64 */
#define PBFMM_MIXED 1 /* FP32 near field (regular), expansions, M2L, SpMV; FP64
                         singular quadrature, alpha, Arnoldi, residual and
                         iterative refinement */

/* Singular quadrature of the near field (cells holding the collocation node
 * or one of its double nodes). */
#define PBFMM_SING_TELLES 0 /* QTelles<2>(sing_quad_order, support point): the
                               rule of BEMFMA::direct_integrals (bem_fma.cc:127-131) */
#define PBFMM_SING_DUFFY 1  /* QSplit<2>(QDuffy(sing_quad_order, 1), support point):
                               the rule of the dense assembly,
                               BEMProblem::get_singular_quadrature
                               (bem_problem.cc:213-229) */

typedef struct pbfmm_handle pbfmm_handle;

/* Parameters of BEMFMA<3>::declare_parameters (bem_fma.cc:64-85) and of the
 * quadrature rules handed to init_fma (bem_fma.cc:41-62). */
typedef struct pbfmm_params {
  int32_t num_octree_levels;  /* "Number of Octree Levels", default 10 */
  int32_t max_nodes_per_leaf; /* "Maximum Number of Collocation Points per
                                 Childless Block", default 20 */
  int32_t trunc_order;        /* "FMA Truncation Order", default 6, 1..10 */
  int32_t quad_order;         /* QGauss order of the regular rule, default 4 */
  int32_t sing_quad_order;    /* Gauss order of the singular rule, default 5 */
  int32_t precision;          /* PBFMM_FP64 or PBFMM_MIXED (solve_system) */
  int32_t device;             /* HIP device ordinal */
  void *stream;               /* hipStream_t, NULL = library-owned stream */
  int32_t preconditioner;     /* 0 none, 1 near-field block Jacobi (default) */
  int32_t sing_rule;          /* PBFMM_SING_TELLES (default, as BEMFMA) or
                                 PBFMM_SING_DUFFY (as the dense solve) */
} pbfmm_params;

/* Defaults of bem_fma.cc:64-85 and bem_problem.cc:268-269. */
void pbfmm_default_params(pbfmm_params *p);

/* BEMFMA<3>::BEMFMA (bem_fma.h:94) + declare/parse_parameters (bem_fma.h:112-115). */
int pbfmm_create(const pbfmm_params *p, pbfmm_handle **h);
/* BEMFMA<3>::~BEMFMA (bem_fma.cc:31-39). */
int pbfmm_destroy(pbfmm_handle *h);
const char *pbfmm_last_error(const pbfmm_handle *h);

/* BEMFMA<3>::init_fma (bem_fma.h:103-108, bem_fma.cc:41-62) with the mesh
 * handed over as arrays instead of DoFHandler / Mapping:
 *   support_xyz   3*n_dofs   DoF support points (map_dofs_to_support_points)
 *   cell_dofs     n_cells*dpc  DoFs per cell in deal.II FE_Q numbering, dpc 4 (Q1) or 9 (Q2)
 *   cell_geom     n_cells*dpc*3 mapping nodes (MappingFEField), NULL = support points
 *   dbl_ptr/idx   CSR of double_nodes_set (each set may omit the DoF itself)
 *   dirichlet_mask n_dofs, 1 = Dirichlet DoF (BEMProblem::dirichlet_nodes) */
int pbfmm_init_fma(pbfmm_handle *h, int64_t n_dofs, const double *support_xyz, int64_t n_cells,
                   int32_t dpc, const int64_t *cell_dofs, const double *cell_geom,
                   const int64_t *dbl_ptr, const int64_t *dbl_idx, const uint8_t *dirichlet_mask);

/* The double-body operator of BEMProblem<3> (the reference's default problem
 * type, bem_problem.h:437-445; dense assembly _assemble_system_double_body,
 * bem_problem.cc:957-1135): every cell also acts through its mirror image about
 * the plane z = z0, whose quadrature points are reflected, whose normals have
 * their z component negated and which is always integrated with the regular
 * rule; collocation points stay on the original surface.  The images become
 * part of the octree, near field and multipoles.  Call after init_fma and
 * before generate_octree_blocking; enable = 0 restores the one-body operator. */
int pbfmm_set_double_body(pbfmm_handle *h, int32_t enable, double z0);

/* BEMFMA<3>::generate_octree_blocking (bem_fma.h:180, bem_fma.cc:3614-4801). */
int pbfmm_generate_octree_blocking(pbfmm_handle *h);

/* BEMFMA<3>::compute_m2l_flags (bem_fma.h:173, body bem_fma.cc:1846-1956,
 * commented out in the reference).  Single process: writes the reference's
 * m2l_operations_per_level, levels+1 entries, into m2l_count_per_level: per
 * level, the nonIntList entries of all sublevels summed over the blocks that
 * hold nodes (same-level entries are M2L translations, finer ones M2P
 * evaluations, coarser ones direct integrals).  Needs the octree only. */
int pbfmm_compute_m2l_flags(pbfmm_handle *h, int64_t *m2l_count_per_level);

/* BEMFMA<3>::direct_integrals (bem_fma.h:127, bem_fma.cc:107-1600): the
 * near-field N and D matrices on their shared pattern, assembled on the GPU. */
int pbfmm_direct_integrals(pbfmm_handle *h);

/* BEMFMA<3>::multipole_integrals (bem_fma.h:137, bem_fma.cc:1608-1840): the
 * translation plans and operators, and the unit multipoles of every leaf
 * (the reference's elemME per (leaf, cell, shape function), merged here per
 * (leaf, DoF)), computed on the GPU in FP64. */
int pbfmm_multipole_integrals(pbfmm_handle *h);

/* BEMFMA<3>::generate_multipole_expansions (bem_fma.h:149-150,
 * bem_fma.cc:1970-2380): leaf P2M of phi (dipole, Ker1) and dphi_dn
 * (monopole, Ker2), then M2M up the tree.  FP64. */
int pbfmm_generate_multipole_expansions(pbfmm_handle *h, const double *phi, const double *dphi_dn);

/* BEMFMA<3>::multipole_matr_vect_products (bem_fma.h:167-170,
 * bem_fma.cc:2390-3072): near SpMV, L2L, M2L, M2P, L2P.  Overwrites
 * matrVectProdN and matrVectProdD.  Uses the expansions of the last
 * generate_multipole_expansions call.  FP64. */
int pbfmm_multipole_matr_vect_products(pbfmm_handle *h, const double *phi, const double *dphi_dn,
                                       double *matrVectProdN, double *matrVectProdD);

/* BEMProblem<3>::compute_alpha, FMA branch (bem_problem.cc:1138-1162,
 * upstream calls at 1159-1160): alpha = N.(-1) via the FMM.  Stored in the
 * handle and copied to alpha (may be NULL). */
int pbfmm_compute_alpha(pbfmm_handle *h, double *alpha);

/* AffineConstraints built by BEMProblem::compute_constraints
 * (bem_problem.cc:1376-1653), as CSR rows over constrained DoFs:
 * row i (is_constrained[i]=1) reads dst(i) = src(i) - sum_k c_val[k] src(c_idx[k])
 * in vmult (constrained_matrix.h:91-135) and rhs(i) = c_inhom[i]
 * (distribute_rhs, constrained_matrix.h:137-147).  NULL pointers clear them. */
int pbfmm_set_constraints(pbfmm_handle *h, const int64_t *c_ptr, const int64_t *c_idx,
                          const double *c_val, const double *c_inhom, const uint8_t *is_constrained);

/* BEMProblem::compute_constraints (bem_problem.cc:1376-1653), one process,
 * no hanging nodes: builds the double-node constraint rows from tmp_rhs (the
 * boundary data, host, n_dofs) and installs them as pbfmm_set_constraints
 * does.  For each set of coincident DoFs the first Dirichlet DoF (else the
 * first DoF) is the master i; Neumann partners of a Neumann master get
 * phi(it) = phi(i), Neumann partners of a Dirichlet master take tmp_rhs(i) as
 * inhomogeneity, and Dirichlet partners of a Dirichlet master get both normal
 * derivatives from the L2-projected normals and surface gradients of tmp_rhs
 * (pbfmm_compute_normals, pbfmm_compute_surface_gradients) when
 * continuous_gradient is set ("Continuos gradient across edges", default
 * true, bem_problem.cc:277).  The reference's parallel-normal test never
 * fires (it divides n_i.n_it by itself, bem_problem.cc:1490), and the
 * formula then divides by 1 - (n_i.n_it)^2 = 0; here projected normals within
 * 1e-4 of parallel give dphi_dn(it) = sign(n_i.n_it) dphi_dn(i) instead. */
int pbfmm_compute_constraints(pbfmm_handle *h, const double *tmp_rhs, int32_t continuous_gradient);
/* The installed constraint rows, in pbfmm_set_constraints' layout (c_ptr
 * n_dofs + 1, c_idx / c_val nnz, c_inhom / is_constrained n_dofs): with only
 * nnz non-NULL, writes the number of entries. */
int pbfmm_get_constraints(pbfmm_handle *h, int64_t *nnz, int64_t *c_ptr, int64_t *c_idx, double *c_val,
                          double *c_inhom, uint8_t *is_constrained);

/* BEMFMA<3>::FMA_preconditioner (bem_fma.h:201-202, bem_fma.cc:3078-3411):
 * the near-field matrix with -D on Dirichlet columns and N on Neumann columns
 * (init_preconditioner, bem_fma.cc:919-975), alpha added on the diagonal of
 * unconstrained Neumann rows, constrained rows replaced by their constraint
 * (with the operator's sign, -c_ij).  The reference factors it with ILU(0);
 * here the rows of the leaves that share a parent block (at most 128 rows per
 * block by default, pbfmm_set_preconditioner_rows; larger parents cut into
 * even parts) form dense blocks that are
 * inverted (block Jacobi), applied as right preconditioner by the GMRES of
 * solve_system and gmres_dev.  alpha NULL uses the handle's alpha (computed
 * if needed). */
int pbfmm_fma_preconditioner(pbfmm_handle *h, const double *alpha);

/* ConstrainedOperator::vmult over BEMProblem<3>::vmult, FMA branch
 * (constrained_matrix.h:91-135, bem_problem.cc:1164-1217):
 * dst = N(src.neu) - D(src.dir) + alpha.(src.neu), pure-Neumann l2 shifts. */
int pbfmm_vmult(pbfmm_handle *h, const double *src, double *dst);

/* BEMProblem<3>::compute_rhs, FMA branch (bem_problem.cc:1219-1260). */
int pbfmm_compute_rhs(pbfmm_handle *h, const double *tmp_rhs, double *rhs);

/* BEMProblem<3>::solve_system (bem_problem.cc:1266-1347) with the FMA
 * branch wired: alpha, rhs, constraints, GMRES(restart) on the GPU, then
 * phi(i) = sol(i) at Neumann DoFs and dphi_dn(i) = sol(i) at Dirichlet DoFs
 * (other entries untouched).  Stops on the true FP64 relative residual
 * ||b - A x|| / ||b|| <= rel_tol.  PBFMM_MIXED runs FP32-operator GMRES
 * inside FP64 iterative refinement. */
int pbfmm_solve_system(pbfmm_handle *h, const double *tmp_rhs, double *phi, double *dphi_dn,
                       double rel_tol, int32_t restart, int32_t max_it, int32_t *iters,
                       double *relres);

/* Device-pointer matvec used by benchmarks: the FMM operator
 * (generate + products) on resident vectors; precision PBFMM_FP64 or
 * PBFMM_MIXED (FP32 operator, FP64 in/out). */
int pbfmm_matvec_dev(pbfmm_handle *h, const double *d_phi, const double *d_dphi, double *d_outN,
                     double *d_outD, int32_t precision);

/* Device-pointer BEMProblem<3>::vmult (same operator as pbfmm_vmult, with
 * constraints) on resident vectors: the FMM of N.src - D.src runs with one
 * combined expansion per block.  precision PBFMM_FP64, or PBFMM_MIXED for the
 * FP32 operator (FP64 in/out); src and dst must differ. */
int pbfmm_vmult_dev(pbfmm_handle *h, const double *d_src, double *d_dst, int32_t precision);

/* pbfmm_vmult_dev on nrhs columns (device pointers, column r at d_src + r*ld
 * and d_dst + r*ld, ld >= n_dofs; the two ranges must not overlap).  One
 * operator application serves all columns: the translation pair lists are
 * expanded over them, so each level is one launch for every column.  Column
 * r equals pbfmm_vmult_dev of column r (FP64 to rounding; FP32 within its
 * rounding, as the atomics sum in another order).  Pure Neumann problems
 * keep vmult's norm shifts, which apply one column at a time.  The
 * reference applies its operator to one vector per BEMProblem::vmult
 * (bem_problem.cc:1164-1217); the design contract's nrhs argument asks for
 * this form. */
int pbfmm_vmult_dev_multi(pbfmm_handle *h, const double *d_src, double *d_dst, int32_t nrhs, int64_t ld,
                          int32_t precision);

/* Device-pointer solve: rhs_dev from pbfmm_compute_rhs semantics already
 * applied (constrained rows set), x_dev initial guess in, solution out.  The
 * two vectors may not overlap (PBFMM_ERR_ARG): the solve reads rhs again
 * after it has updated x. */
int pbfmm_gmres_dev(pbfmm_handle *h, const double *d_rhs, double *d_x, double rel_tol,
                    int32_t restart, int32_t max_it, int32_t precision, int32_t *iters,
                    double *relres);

/* Several right-hand sides of one operator (device pointers, column r at
 * d_rhs + r*ld and d_x + r*ld, ld >= n_dofs; packed internally when
 * ld > n_dofs), solved in lockstep: every right-hand side runs its own
 * Arnoldi process and stopping test (true relative residual <= rel_tol, as
 * gmres_dev), and every operator application serves all of them, with one
 * translation launch per level for all right-hand sides (pair lists expanded
 * over them) and the near field, unit multipoles, operators and
 * preconditioner shared; the rhs and x ranges may not overlap.  iters gets the Arnoldi steps each right-hand side
 * used and relres its true relative residual (nrhs entries, may be NULL).  No
 * counterpart in the reference, which solves one right-hand side per
 * BEMProblem::solve. */
int pbfmm_gmres_dev_multi(pbfmm_handle *h, const double *d_rhs, double *d_x, int32_t nrhs, int64_t ld,
                          double rel_tol, int32_t restart, int32_t max_it, int32_t precision, int32_t *iters,
                          double *relres);

/* ---- introspection (tests, profiling) ---------------------------------- */
/* counts: [0] num_blocks [1] levels [2] n_childless [3] n_sub (sum of
 * sublevels) [4] n_intlist [5] n_nonintlist [6] n_node_entries [7] nnz near */
int pbfmm_tree_counts(pbfmm_handle *h, int64_t *counts);
/* Same layout as the counts above; lists per (block, sublevel) ascending. */
int pbfmm_tree_export(pbfmm_handle *h, int32_t *level, int64_t *parent, double *pmin,
                      double *delta, int64_t *startL, int64_t *endL, int64_t *childless,
                      int64_t *sub_ptr, int64_t *il_ptr, int64_t *il_idx, int64_t *nil_ptr,
                      int64_t *nil_idx, int64_t *node_ptr, int64_t *node_idx);
/* Near field as CSR in DoF numbering, columns ascending (precision selects
 * the FP64 or the mixed-mode FP32 values, widened). */
int pbfmm_nearfield_export(pbfmm_handle *h, int32_t precision, int64_t *rowptr, int64_t *col,
                           double *valN, double *valD);
/* Expansions of every block after the last FP64
 * pbfmm_generate_multipole_expansions / pbfmm_multipole_matr_vect_products
 * (the device entry points skip levels whose expansions nothing reads):
 * which 0 = multipole, 1 = local; Kt=(p+1)(p+2)/2 complex coefficients per block and kernel,
 * interleaved (re, im), index n(n+1)/2+m as in multipole_expansion.h:101-117. */
int pbfmm_get_expansions(pbfmm_handle *h, int32_t which, double *ker1, double *ker2);
/* Per-phase device times of the last matvec in ms, measured with HIP events
 * on the handle's stream while phase timing is enabled (it adds one host
 * synchronisation per matvec):
 * [0] spmv [1] p2m [2] m2m [3] l2l [4] m2l [5] m2p [6] l2p [7] total
 * [8] the per-target sum of the slot M2L (0 when M2L adds into its targets;
 *     [4] then holds the translations alone): 9 doubles. */
int pbfmm_set_phase_timing(pbfmm_handle *h, int32_t enable);
int pbfmm_last_timings(pbfmm_handle *h, double *ms);

/* Work counts of the current plan, for roofline accounting:
 * [0] n_dofs [1] near-field values (rows x cols of the leaf blocks)
 * [2] near-field column entries [3] near-field rows [4] leaves
 * [5] quad points [6] M2L pairs [7] M2P (pair, node) evaluations
 * [8] M2M child translations [9] L2L translations [10] num_blocks
 * [11] trunc_order [12] padded expansion length kp
 * [13] unit-multipole columns (sum over leaves of the DoFs feeding P2M)
 * [14] [15] near-field SpMV kernels chosen at set-up for FP64 / FP32:
 *      kernel of the operator's single values + 8 * kernel of the (N, D)
 *      pair storage (as pbfmm_set_spmv_kernel) */
int pbfmm_work_counts(pbfmm_handle *h, int64_t *counts);

/* The block-Jacobi preconditioner of pbfmm_fma_preconditioner before
 * inversion (the matrix FMA_preconditioner assembles, bem_fma.cc:3078-3411,
 * restricted to its diagonal blocks), for tests: sizes[0] blocks, [1] rows in
 * all, [2] values in all; with row_beg (blocks + 1), rows and values non-NULL
 * also block b's rows rows[row_beg[b] .. row_beg[b + 1]) and its R x R
 * matrix, row-major, in values (blocks one after another).  Builds the
 * preconditioner if needed. */
int pbfmm_preconditioner_blocks(pbfmm_handle *h, int64_t *sizes, int32_t *row_beg, int32_t *rows, double *values);
/* Row cap of the block-Jacobi blocks (8..128, default 128): the near leaves
 * of one parent block are grouped up to this many rows (a larger leaf stays
 * alone).  Rebuilds the preconditioner if it exists.  No counterpart in the
 * reference (its ILU has no blocks). */
int pbfmm_set_preconditioner_rows(pbfmm_handle *h, int32_t rows);

/* Mixed-precision solver scheme (no counterpart in the reference, whose
 * solve is FP64 throughout): 0 (default) chooses by the ratio of the FP64 to
 * the FP32 operator's cost, timed once on this handle; 1 relaxed (FP64
 * operator until the residual is below 2e5 * rel_tol * ||b||, then the FP32
 * operator, in one Krylov space); 2 iterative refinement (FP32-operator GMRES
 * on FP64 residuals, the second inner solve recycling the first one's Krylov
 * space); 3 the same without recycling; 4 scheme 2 with the inner solves'
 * basis vectors stored in FP32 (their arithmetic in FP64; an option, since
 * the FP64 Arnoldi of the other schemes is what the mixed mode promises).
 * Every scheme ends on the true FP64 residual.  Mixed solves apply the
 * block-Jacobi inverses stored in FP32, FP64 solves the FP64 copy. */
int pbfmm_set_mixed_scheme(pbfmm_handle *h, int32_t scheme);

/* The two thresholds of the mixed-precision schemes (no counterpart in the
 * reference): inner_reduction, a fixed residual reduction asked of each
 * FP32-operator solve of the iterative refinement (an inner solve then stops
 * at max(rel_tol ||b|| / 2, inner_reduction ||r||); by default the reduction
 * still needed is split into the fewest equal stages of at most 1e-5), and
 * relax_switch, the relaxed scheme's switch to the FP32 operator once the
 * residual is below relax_switch * rel_tol * ||b|| (default 2e5).  A value
 * <= 0 restores the default.  For tuning. */
int pbfmm_set_mixed_tolerances(pbfmm_handle *h, double inner_reduction, double relax_switch);

/* Gram-Schmidt of the GMRES solves (pbfmm_gmres_dev, pbfmm_solve_system): an
 * Arnoldi step projects the new vector against the basis once and projects
 * it a second time only when the first projection left less than a fraction
 * eta of its norm (default 0.05; eta >= 1 always projects twice, classical
 * Gram-Schmidt with reorthogonalisation; 0 never does; < 0 restores the
 * default).  Inner solves with FP32 basis vectors (mixed scheme 4) always
 * project twice.  Every solve still ends on the true relative residual.  No
 * counterpart in the reference (deal.II's SolverGMRES orthogonalises by its
 * own rule). */
int pbfmm_set_gmres_reorth(pbfmm_handle *h, double eta);

/* Overrides the near-field SpMV kernel that set-up's autotuner chose for one
 * precision (PBFMM_FP64 or PBFMM_MIXED for the FP32 values), for both
 * storages: 0 workgroup per leaf, 1 wavefront per leaf with 2 rows in flight,
 * 2 with 4 rows; FP32 also 3..5: kernels 0..2 on the operator's single values
 * kept in row pairs (two rows per 8-byte load; the (N, D) pair storage then
 * uses kernel - 3), and 6: workgroup per leaf on the single values kept in
 * column vectors, a row per lane (four columns per 16-byte load; the pair
 * storage uses kernel 0); FP64 also 3: the same with two columns per
 * 16-byte load.  For tests and tuning; call after
 * pbfmm_direct_integrals.  No counterpart in the reference (Trilinos
 * chooses its own SpMV). */
int pbfmm_set_spmv_kernel(pbfmm_handle *h, int32_t precision, int32_t kernel);

/* Overrides the timed launch plan of the operator for one precision
 * (0 none, 1 SpMV beside an M2M level and M2P beside an L2L level,
 * 2 SpMV beside an M2M level and M2P beside M2L, 3 both beside M2L; 4 and 5:
 * plans 0 and 1 with M2L storing each pair's result into a slot of its own,
 * summed per target by a second kernel, instead of adding it to the target
 * with atomics; 8 and 12: plans 0 and 4 with the SpMV on a second stream of
 * the library, beside the whole FMM; 24 and 28: plans 8 and 12 with the
 * second stream forked after the M2M levels, beside M2L onwards; 56 and 60:
 * plans 24 and 28 with M2P following the SpMV on the second stream; 88 and
 * 92: with M2P before the SpMV there; -1 times them all again on the next
 * application).  For tests and tuning.  No
 * counterpart in the reference. */
int pbfmm_set_side_fusion(pbfmm_handle *h, int32_t precision, int32_t mode);

/* The choices this handle's timings made so far (-1 = not timed yet):
 * [0] [1] near-field SpMV kernel of the operator's single values, FP64 / FP32
 *         (as pbfmm_set_spmv_kernel)
 * [2] [3] launch plan of the operator, FP64 / FP32: 0 none,
 *         1 SpMV beside an M2M level and M2P beside an L2L level, 2 SpMV
 *         beside an M2M level and M2P beside M2L, 3 both beside M2L, 4 and 5
 *         plans 0 and 1 with M2L into per-pair slots summed per target, 8
 *         and 12 plans 0 and 4 with the SpMV on a second stream, 24 and 28
 *         the same forked after the M2M levels, 56 and 60 plans 24 and 28
 *         with M2P on the second stream after the SpMV, 88 and 92 before it
 * [4]     mixed-precision scheme: 1 relaxed, 2 iterative refinement
 * [5]     timed FP64 / FP32 operator cost ratio x 1000
 * [6] [7] the GMRES step's operator, FP64 / FP32 steps: 1 replayed as a HIP
 *         graph, 0 launched directly (as pbfmm_set_graph_mode)
 * state holds 8 values.  No counterpart in the reference. */
int pbfmm_tuning_state(pbfmm_handle *h, int32_t *state);

/* HIP graphs for the GMRES step's operator (no counterpart in the
 * reference): the preconditioner and the operator application of one Arnoldi
 * step of pbfmm_gmres_dev / pbfmm_solve_system run from a fixed input into a
 * fixed output, so the launches can be captured once and replayed.  -1
 * (default): the first steps of each precision capture the graph and time
 * three replays against three direct runs, and the faster is kept; 0 direct
 * launches only; 1 graph replays.  A graph is captured again whenever a
 * choice or buffer it holds changes.  Batch handles and many right-hand
 * sides launch directly. */
int pbfmm_set_graph_mode(pbfmm_handle *h, int32_t mode);
/* The timing behind the graph choice, ms per step: [0] [1] FP64 direct /
 * graph, [2] [3] FP32 direct / graph (0 until timed). */
int pbfmm_graph_timings(pbfmm_handle *h, float *ms);

/* ---- Post-solve (the step after the solve; real cells only) ----
 * Vector fields are component-major: v[d * n_dofs + i] is component d at DoF
 * i, which is the reference's vector DoF numbering after component_wise
 * renumbering (vec_original_to_sub_wise[n_dofs * d + i], boundary_conditions.cc:226-236).
 * The L2 projections solve the consistent mass matrix of the scalar element,
 * built with the QGauss(quad_order) rule of the handle, to a relative
 * residual of 1e-13 per component; they need init_fma only. */

/* BEMProblem::compute_normals (bem_problem.cc:1997-2079): L2 projection of
 * the cell normals t1 x t2 / |t1 x t2| onto the vector element (not
 * normalised afterwards; vector_normals_solution). */
int pbfmm_compute_normals(pbfmm_handle *h, double *normals);
/* BEMProblem::compute_gradients (bem_problem.cc:1746-1876): L2 projection of
 * n * dphi_dn + surface gradient of phi (vector_gradients_solution).  The
 * result is also kept for pbfmm_hydrodynamic_pressure. */
int pbfmm_compute_gradients(pbfmm_handle *h, const double *phi, const double *dphi_dn, double *gradients);
/* BEMProblem::compute_surface_gradients (bem_problem.cc:1878-1995): L2
 * projection of the surface gradient of tmp_rhs. */
int pbfmm_compute_surface_gradients(pbfmm_handle *h, const double *tmp_rhs, double *gradients);
/* Conjugate-gradient steps of the last projection. */
int pbfmm_projection_iterations(pbfmm_handle *h, int32_t *iters);
/* BEMProblem::hydrostatic_pressure (bem_problem.cc:2081-2117):
 * p_i = -density * gravity * (z_i - z0) at every DoF. */
int pbfmm_hydrostatic_pressure(pbfmm_handle *h, double gravity, double density, double z0, double *pressure);
/* BEMProblem::hydrodynamic_pressure (bem_problem.cc:2119-2167):
 * p_i = -density * sum_d u_d (u_d / 2 - W_d), with u the gradients of the last
 * pbfmm_compute_gradients (PBFMM_ERR_STATE before it) and W the wind function
 * evaluated at the support points (3 n_dofs, component-major). */
int pbfmm_hydrodynamic_pressure(pbfmm_handle *h, double density, const double *wind, double *pressure);
/* Body integrals over the cells with cell_mask[c] != 0 (Body::hasMaterial of
 * the cell's material id; NULL selects every cell), n_cells entries:
 * BEMProblem::center_of_pressure (bem_problem.cc:2169-2219): int x p / int p,
 * center left unchanged when int p is 0;
 * BEMProblem::pressure_force_and_moment (bem_problem.cc:2221-2272):
 * force = int p n, moment = int p (cog - x) x n;
 * BEMProblem::volume_integral (bem_problem.cc:2274-2309): -int x_d n_d. */
int pbfmm_center_of_pressure(pbfmm_handle *h, const uint8_t *cell_mask, const double *pressure, double *center);
int pbfmm_pressure_force_and_moment(pbfmm_handle *h, const uint8_t *cell_mask, const double *pressure,
                                    const double *cog, double *force, double *moment);
int pbfmm_volume_integral(pbfmm_handle *h, const uint8_t *cell_mask, double *volume);
/* pbfmm_pressure_force_and_moment without any device allocation.  No
 * counterpart in the reference: for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops).  Device pointers: d_cell_mask
 * (n_cells bytes, NULL selects every real cell), d_pressure (n_dofs) and
 * d_out (6 doubles: force, then moment about cog); cog is a host array of 3.
 * The per-workgroup partial sums live in a fixed device array of the library
 * (6.8 MB on every GPU it runs on, held until the process ends), which limits
 * the mesh to 16,777,216 cells (PBFMM_ERR_ARG beyond).  The handle's
 * post-solve data must exist (any post-solve call above sets it up;
 * PBFMM_ERR_STATE before).  Asynchronous on the handle's stream: d_out holds
 * the result once that stream has reached this call.  Every call on one GPU
 * shares that array, so calls that overlap on one GPU (other streams, other
 * host threads) race; calls in order on one stream do not. */
int pbfmm_pressure_force_and_moment_noalloc(pbfmm_handle *h, const uint8_t *d_cell_mask, const double *d_pressure,
                                            const double *cog, double *d_out);
/* BEMProblem::free_surface_elevation (bem_problem.cc:2311-2370): for each
 * (cell, line) pair of cell_lines (line in deal.II numbering 0..3: x = 0,
 * x = 1, y = 0, y = 1 of the unit cell; the reference visits waterline
 * boundary lines of body cells in ascending cell, then line, order) and each
 * of the quad_order Gauss points along it, elevation receives
 * (x, y, p / (density * gravity)): 3 * quad_order doubles per pair. */
int pbfmm_free_surface_elevation(pbfmm_handle *h, double gravity, double density, const int32_t *cell_lines,
                                 int64_t n_lines, const double *pressure, double *elevation);

/* ---- Batches of independent problems ----
 * No counterpart in the reference, which sets up and solves one problem per
 * run; a batch composes its per-problem steps (BEMFMA set-up,
 * bem_fma.cc:41-4801; BEMProblem::solve_system, bem_problem.cc:1266-1347) over
 * many problems as one block-diagonal operator, so that every kernel launch
 * covers all of them.
 *
 * pbfmm_batch_create: parts are handles after init_fma (their octrees and
 * near/far plans are built here, in parallel on the host, if not built yet);
 * they must share trunc_order, num_octree_levels, quad orders, precision,
 * device and element degree, and may not be double bodies or pure Neumann
 * problems.  The new handle is set up (near field, multipoles) and accepts
 * every call of the one-problem path on the concatenated vectors (part s at
 * DoFs dof_off[s] .. dof_off[s+1]); the per-mesh set-up and tree-export calls
 * return PBFMM_ERR_UNSUPPORTED on it.  The parts stay valid and independent. */
int pbfmm_batch_create(pbfmm_handle *const *parts, int32_t n_parts, pbfmm_handle **batch);
/* DoF offsets of the parts: n_parts + 1 entries. */
int pbfmm_batch_offsets(pbfmm_handle *batch, int64_t *dof_off);
/* GMRES on device vectors with each part's own true relative residual
 * ||b_s - A_s x_s|| / ||b_s|| <= rel_tol (written to relres, n_parts
 * entries): every part runs its own Arnoldi process (own dot products,
 * Hessenberg matrix, rotations and stopping test) inside shared basis vectors,
 * the operator is applied to all parts at once, and a part that has converged
 * stops taking part; iters counts operator applications. */
int pbfmm_batch_gmres_dev(pbfmm_handle *batch, const double *d_rhs, double *d_x, double rel_tol, int32_t restart,
                          int32_t max_it, int32_t precision, int32_t *iters, double *relres);

#ifdef __cplusplus
}
#endif
#endif /* PBFMM_H */
