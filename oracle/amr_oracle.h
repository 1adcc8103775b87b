/*
 * amr_oracle.h -- CPU restatement of Chombo 3.2's AMRPoissonOp / MultiGrid /
 * AMRMultiGrid / BiCGStab FP64 path.
 *
 * THIS IS TEST INFRASTRUCTURE (the oracle).  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it.  The product library never
 * links, loads or calls anything in this directory.
 *
 * Hierarchy description (shared with the product's C ABI, see include/amrmg.h):
 *   ispec[0]           = nlev (number of AMR levels, refinement ratio 2)
 *   ispec[1..6]        = level-0 domain box lo0 lo1 lo2 hi0 hi1 hi2
 *   ispec[7..9]        = BC type on lo faces  (0 Neumann, 1 Dirichlet)
 *   ispec[10..12]      = BC type on hi faces
 *   ispec[13..13+nlev) = number of boxes on each level
 *   then 6 ints per box (lo0 lo1 lo2 hi0 hi1 hi2), level by level, each level
 *   sorted like a Chombo DisjointBoxLayout (Box::operator<).
 *   rspec[0] = dx0, rspec[1] = alpha, rspec[2] = beta,
 *   rspec[3..5] = BC values lo faces, rspec[6..8] = BC values hi faces.
 *
 * Field data crossing this interface is "valid-packed": for each level, for each
 * box in layout order, the box's valid cells in Fortran order (i fastest).
 */
#ifndef AMR_ORACLE_H
#define AMR_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct OHier OHier;

/* solver parameters: sparams = {pre, post, bottom, numMG, iterMax, imin},
   rparams = {eps, hang, normThresh} */
OHier* oracle_create(const int* ispec, const double* rspec);
/* AMRMultiGrid::solve(phi, rhs, l_max, l_base): levels l_base..l_max only, with
   phi of level l_base-1 as coarse-fine boundary data (phi/rhs packed over all levels) */
int oracle_solve_levels(OHier* h, const int* sparams, const double* rparams, const double* rhs,
                        double* phi, int zero_phi, int l_max, int l_base, double* hist, int hist_cap, int* info);
/* the same with refinement ratios (2 or 4) between consecutive levels; NULL: all 2 */
OHier* oracle_create_r(const int* ispec, const double* rspec, const int* ratios);
/* BCValueFunc with a caller pointer; fn NULL restores the constant face values.
   value holds one entry per component. */
typedef void (*oracle_bc_value_fn)(double* pos, int* dir, int* side, double* value, void* user);
int    oracle_set_bc_values(OHier* h, oracle_bc_value_fn fn, void* user);
/* components (1..16) of the fields of oracle_solve / _solve_levels / _amr_residual /
   _amr_operator / _level_norm / _level_dot: valid-packed box by box, in each box
   component by component.  Every component shares the operator and its boundary
   conditions (ConstDiriNeumBC gives each the same face value); the dot products and
   norms of the solver run over all components in FArrayBox order.  The other level
   functions take one component.  Not with variable coefficients. */
int    oracle_set_ncomp(OHier* h, int ncomp);
void   oracle_destroy(OHier* h);
long   oracle_level_cells(const OHier* h, int lev);
int    oracle_num_mg_depths(const OHier* h);
const char* oracle_last_error(void);

/* AMRMultiGrid<LevelData<FArrayBox>>::solve with a BiCGStab bottom solver.
   phi in: initial guess (used only when zero_phi == 0); out: solution.
   hist receives the max-norm residual after each outer iteration (hist[0] = initial),
   up to hist_cap entries.  info = {iterations, exitStatus, bottom calls}. */
int oracle_solve(OHier* h, const int* sparams, const double* rparams,
                 const double* rhs, double* phi, int zero_phi,
                 double* hist, int hist_cap, int* info);

/* AMRMultiGrid::computeAMRResidual over levels 0..nlev-1 (inhomogeneous BCs unless
   homog): res = rhs - L(phi) with CF interpolation and refluxing; covered cells of
   res zeroed.  Returns the max norm.  phi is updated only in its ghost cells, which
   are not part of the valid-packed output. */
double oracle_amr_residual(OHier* h, const double* phi, const double* rhs,
                           double* res, int homog);

/* AMRMultiGrid::computeAMROperator (no zeroing of covered cells). */
int oracle_amr_operator(OHier* h, const double* phi, double* lphi, int homog);

/* Single-level kernels on the MG level `depth` of AMR level 0 (depth 0 = level 0
   itself), each preceded by the reference's ghost refresh (exchange + BC, and
   homogeneous CF interpolation where the reference does it). */
int oracle_level_apply(OHier* h, int depth, const double* phi, double* lphi, int homog);
int oracle_level_residual(OHier* h, int depth, const double* phi, const double* rhs,
                          double* res, int homog);
int oracle_level_relax(OHier* h, int depth, double* phi, const double* rhs, int iters);
int oracle_level_restrict(OHier* h, int depth, const double* phi, const double* rhs,
                          double* resc);
int oracle_level_prolong(OHier* h, int depth, double* phi, const double* ec);
long oracle_mg_cells(const OHier* h, int depth);

/* one colour pass of GSRB on the raw kernel, ghosts supplied by the caller.
   phi_g: ghosted-packed (each box grown by 1), rhs: valid-packed. */
int oracle_kernel_gsrb(OHier* h, int lev, double* phi_g, const double* rhs,
                       int color);

/* AMRPoissonOp::homogeneousCFInterp on AMR level lev; phi_g ghosted-packed in/out */
int oracle_homog_cf(OHier* h, int lev, double* phi_g);
/* QuadCFInterp::coarseFineInterp(phi[lev], phi[lev-1]); phi_g ghosted-packed for
   level lev (in/out), coarse valid-packed */
int oracle_quad_cf(OHier* h, int lev, double* phi_g, const double* coarse);
/* AMRPoissonOp::AMRRestrictS + AVERAGE: resc (layout coarsen(grids[lev],2)) */
int oracle_average(OHier* h, int lev, const double* fine, double* coarse_avg);
long oracle_ghosted_cells(const OHier* h, int lev);

/* BiCGStab on the bottom MG level of AMR level 0, exactly as MultiGrid::cycle
   invokes it (bottom relaxations first).  metric = convergence metric, eps. */
int oracle_bottom_solve(OHier* h, double* corr, const double* res, int nbottom,
                        double metric, double eps, int* info);

/* VCAMRPoissonOp2: L phi = alpha a phi - beta div(b grad phi) on every AMR level.
   a valid-packed over all levels; b_d level after level, per box in layout order,
   the box's faces normal to d (the box grown by one in d) in Fortran order.  Every
   later call on h uses the variable-coefficient operator (MG depths of level 0 get
   the coefficients averaged from level 0; reflux uses -b * gradient fluxes). */
int  oracle_vc_set(OHier* h, const double* a, const double* b0, const double* b1, const double* b2);
long oracle_face_cells(const OHier* h, int lev, int dir);
/* norm(BoxLayoutData, interval, p) (BoxLayoutData.cpp:23-92) on MG depth d, p = 0, 1, 2 */
int  oracle_level_norm(OHier* h, int d, const double* x, int p, double* out);
/* LevelDataOps::dotProduct (per box FArrayBox::dotProduct, LevelDataOps.H:109-136) on MG depth d */
int  oracle_level_dot(OHier* h, int d, const double* a, const double* b, double* out);
/* AMRPoissonOp::s_relaxMode for every operator: 1 levelGSRB (default), 4 levelJacobi */
int  oracle_set_relax_mode(OHier* h, int mode);
int  oracle_set_vc_average(OHier* h, int type);
/* setAlphaAndBeta on every operator (AMR levels and MG depths), lambda recomputed */
int  oracle_set_alpha_beta(OHier* h, double alpha, double beta);
int  oracle_vc_lambda(OHier* h, int depth, double* lambda_out);

/* BiCGStabSolver's public parameters (BiCGStabSolver.H:105-153) for every later bottom
   solve (oracle_bottom_solve and the multigrid bottom of oracle_solve); defaults
   imax 80, numRestarts 5, reps 1e-12, hang 1e-8, small 1e-30.  oracle_bottom_solve's
   info = {exit status, iterations}. */
int  oracle_set_bottom_params(OHier* h, int imax, int numRestarts, double reps, double hang, double small);
/* applyOp / residual / relax of AMR level lev's operator (its own CFRegion,
   homogeneousCFInterp with dxCrse = dx of level lev-1); valid-packed level data */
int  oracle_amrlevel_apply(OHier* h, int lev, const double* phi, double* out, int homog);
int  oracle_amrlevel_residual(OHier* h, int lev, const double* phi, const double* rhs, double* out, int homog);
int  oracle_amrlevel_relax(OHier* h, int lev, double* phi, const double* rhs, int iters);
/* QuadCFStencil branch statistics of level lev's coarse-fine stencils (13 counts,
   documented at the definition) */
int  oracle_cf_branch_counts(OHier* h, int lev, long* counts);

/* OpenMP threads of every later call (n > 0), and the current count */
void oracle_set_threads(int n);
int  oracle_get_threads(void);

/* number of cell-relaxations (one red+black GSRB iteration of one cell) since reset */
long long oracle_cell_relax_count(void);
void oracle_reset_counters(void);

#ifdef __cplusplus
}
#endif
#endif
