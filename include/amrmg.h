/*
 * amrmg.h -- C ABI of the MI355X-native AMR multigrid Poisson/Helmholtz library
 * (libamrmg.so, built from chombo_3.2_amd/csrc).
 *
 * The reference (Chombo 3.2) crosses its language boundary once per box per
 * kernel: C++ calls ChomboFortran subroutines through FORTRAN_NAME/FORT_* with
 * CHF_FRA/CHF_BOX pointer arguments (lib/src/BoxTools/FORT_PROTO.H:293-447,
 * lib/util/chfpp/subrout.pm:1049-1068).  This library replaces that layer: each
 * entry point below runs the reference operation on a whole level (all boxes) on
 * the GPU, and the solver entry points run the reference's AMRMultiGrid /
 * MultiGrid / BiCGStab control flow on device-resident levels.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative code on error; the
 *    message is available from amrmg_last_error() (thread-local).  No C++
 *    exception crosses this boundary.
 *  - A hierarchy is described by (ispec, rspec):
 *      ispec[0]            nlev (AMR levels, refinement ratio 2)
 *      ispec[1..6]         level-0 domain box lo0 lo1 lo2 hi0 hi1 hi2
 *      ispec[7..9]         BC type on the lo faces (0 Neumann, 1 Dirichlet, 2 Dirichlet
 *                          quadratic, 3 periodic)
 *      ispec[10..12]       BC type on the hi faces
 *      ispec[13..13+nlev)  number of boxes per level
 *      then 6 ints per box (lo0 lo1 lo2 hi0 hi1 hi2), level by level, each level
 *      sorted by Chombo's Box::operator< (as a closed DisjointBoxLayout is)
 *      rspec = {dx0, alpha, beta, bcval_lo[3], bcval_hi[3]}
 *    (AMRPoissonOpFactory::define, lib/src/AMRElliptic/AMRPoissonOp.cpp:1857-1903,
 *     with ConstDiriNeumBC, lib/src/AMRElliptic/BCFunc.cpp:58-158).
 *  - Host field buffers are "valid-packed": level by level, box by box in layout
 *    order, the box's valid cells in Fortran order (i fastest).  "Ghosted-packed"
 *    buffers hold each box grown by one cell instead.
 *  - solver parameters: sparams = {pre, post, bottom, numMG, iterMax, imin},
 *    rparams = {eps, hang, normThresh}
 *    (AMRMultiGrid::setSolverParameters, lib/src/AMRElliptic/AMRMultiGrid.H:528-553).
 *  - precision: 0 = FP64 throughout; 1 = mixed (FP32 inner V-cycles, FP64 outer
 *    residual / update / bottom solve).
 */
#ifndef AMRMG_H
#define AMRMG_H
#ifdef __cplusplus
extern "C" {
#endif

/* BC type codes in ispec: 0 Neumann (ConstBCFunction type 0), 1 Dirichlet with the
   linear rule 2v - near (ConstBCFunction type 1), 2 Dirichlet with DiriBC's order-2
   rule (8/3)v + (1/3)far - 2near (BCFunc.cpp:34-40, 291-354); values are constant per face.
   3 periodic, on both faces of a direction (ProblemDomain's isPeriodic; the ghosts come
   from the boxes' periodic images as Copier::exchangeDefine takes them, Copier.cpp:1464-1541);
   every level above the base must stay at least 2 coarse cells away from periodic faces. */
typedef struct amrmg_ctx amrmg_ctx;
typedef struct amrmg_solver amrmg_solver;

const char* amrmg_last_error(void);

/* device context (one HIP stream, default priority) */
int amrmg_ctx_create(int device, amrmg_ctx** out);
int amrmg_ctx_destroy(amrmg_ctx* ctx);

/* Relaxation on levels stored as one rectangle: 1 (default) runs each GSRB
   iteration as one launch that forms the coarse-fine and boundary ghosts of both
   colours in registers; 0 runs the reference's sequence (ghost refresh, red,
   ghost refresh, black) as four launches.  Results are bitwise the same. */
int amrmg_ctx_set_fused_relax(amrmg_ctx* ctx, int on);

/* AMRPoissonOpFactory::define + AMRMultiGrid::define + init with a BiCGStab
   bottom solver (replaces AMRPoissonOp.cpp:1857-2070, AMRMultiGrid.H:1212-1328) */
int amrmg_solver_create(amrmg_ctx* ctx, const int* ispec, const double* rspec, int precision,
                        amrmg_solver** out);
/* the same with the refinement ratio between each pair of consecutive levels:
   ref_ratios[l-1] (2 or 4) between levels l-1 and l, nlev-1 entries (NULL: 2 everywhere),
   as AMRPoissonOpFactory::define's refRatios (AMRPoissonOp.cpp:1857-1903).  A level of
   ratio 4 smooths as AMRMultiGrid::relax does there (AMRMultiGrid.H:633-666): m_bottom
   relaxations per visit (one when its multigrid bottom has one cell), not m_pre / m_post. */
int amrmg_solver_create_ratios(amrmg_ctx* ctx, const int* ispec, const double* rspec, const int* ref_ratios,
                               int precision, amrmg_solver** out);
/* The same for fields of ncomp components (1..16): AMRPoissonOp on LevelData<FArrayBox>
   with nComp = ncomp (AMRPoissonOp.cpp, every ChF kernel looping over components,
   e.g. AMRPoissonOpF.ChF:158).  Every component shares the operator and the BC types;
   ConstDiriNeumBC gives each component the same face value, a value function fills
   one value per component (amrmg_solver_set_bc_values).  The solver's dot products and
   norms run over all components in FArrayBox order (LevelDataOps::dotProduct,
   norm(BoxLayoutData, p)), so the bottom BiCGStab couples the components as the
   reference's does.  Fields are valid-packed box by box, each box component by
   component.  FP64 (precision 0) and constant coefficients only; the level, FAB and
   bottom entry points below take one component; the heat steppers, the batched and
   serving paths take single-component problems. */
int amrmg_solver_create_ncomp(amrmg_ctx* ctx, const int* ispec, const double* rspec, const int* ref_ratios,
                              int ncomp, int precision, amrmg_solver** out);
int amrmg_solver_destroy(amrmg_solver* s);
int amrmg_solver_set_params(amrmg_solver* s, const int* sparams, const double* rparams);
/* AMRPoissonOp::setAlphaAndBeta (AMRPoissonOp.cpp:1103-1108, unit aCoef and bCoef) on
   every operator of the solver (AMR levels and multigrid levels): the same as defining the solver
   with these coefficients. */
int amrmg_solver_set_alpha_beta(amrmg_solver* s, double alpha, double beta);
/* BCValueFunc (BCFunc.H:205-208) with a caller pointer: writes value[0], the boundary
   value at face position pos[3] of the domain face (dir 0..2, side 0 low / 1 high). */
typedef void (*amrmg_bc_value_fn)(double* pos, int* dir, int* side, double* value, void* user);
/* DiriBC / NeumBC with a BCValueHolder (BCFunc.cpp:217-258, 291-354) in place of the
   constant face values of ispec/rspec, on every operator of the solver: each
   inhomogeneous ghost on a Dirichlet or Neumann face takes fn's value at the centre of
   its domain face, pos = dx*(iv+0.5) of the valid cell next to it moved half a cell
   outward (getDomainFacePosition, BCFunc.cpp:187-202, no origin offset).  fn is called
   here, once per face cell of every level and multigrid depth, and must be a pure
   function of its arguments; fn = NULL restores the constant values.  The BC types
   stay those of ispec.  value holds one entry per component of the solver's fields
   (BCValueFunc's a_values[icomp]). */
int amrmg_solver_set_bc_values(amrmg_solver* s, amrmg_bc_value_fn fn, void* user);
/* Variable coefficients: the solver's operator becomes VCAMRPoissonOp2's
   alpha a(x) phi - beta div(b(x) grad phi) (VCAMRPoissonOp2Factory::define,
   VCAMRPoissonOp2.cpp:868-940, with MGnewOp's arithmetic coefficient averaging,
   :991-1070, AMRnewOp, :1074-1171, computeLambda, :283-318, and reflux with
   -b * gradient fluxes, :339-430, 787-837).  a is valid-packed over all levels;
   b_d (d = 0, 1, 2) holds, level after level and box after box, the faces normal
   to d of the box grown by one face in d, in Fortran order (the packing of a
   FluxBox component).  On a merged level, faces shared by two boxes must carry
   the same value in both.  Multigrid depths of level 0 use level 0's coefficients
   averaged arithmetically.  alpha and beta are the solver's (amrmg_solver_set_alpha_beta
   changes them and recomputes lambda).  a = NULL returns to constant
   coefficients.  With precision 1 the inner FP32 cycles use the coefficients
   rounded to FP32 (averaged in FP64 first).  The bottom solve is the
   host-driven BiCGStab. */
int amrmg_solver_set_vc_coefs(amrmg_solver* s, const double* a, const double* b0, const double* b1,
                              const double* b2);
long long amrmg_solver_level_cells(amrmg_solver* s, int lev);
int amrmg_solver_num_mg_depths(amrmg_solver* s);
long long amrmg_solver_mg_cells(amrmg_solver* s, int depth);
/* bytes of HBM the solver holds (fields + tables) */
long long amrmg_solver_device_bytes(amrmg_solver* s);

/* field 0 = phi, 1 = rhs, 2 = residual (last computeAMRResidual); valid-packed level
   after level, box after box, each box component by component */
int amrmg_solver_set_field(amrmg_solver* s, int field, const double* valid_packed);
int amrmg_solver_get_field(amrmg_solver* s, int field, double* valid_packed);

/* AMRMultiGrid::solve(phi, rhs, nlev-1, 0, zero_phi) on the device-resident
   fields.  hist[0..] = max-norm residual history (up to hist_cap entries),
   info = {iterations, exitStatus, bottom BiCGStab iterations}. */
int amrmg_solver_solve(amrmg_solver* s, int zero_phi, double* hist, int hist_cap, int* info);
/* AMRMultiGrid::solve(phi, rhs, l_max, l_base, zeroPhi) (AMRMultiGrid.H:895-1124): only
   levels l_base..l_max are solved for; the phi of level l_base-1 (as set) gives the
   coarse-fine boundary values of level l_base, and level l_base cycles with its own
   MultiGrid (MGnewOp on that AMR level, made on first use) down to the BiCGStab
   bottom.  Fields of other levels are left as they are.  FP64 with constant
   coefficients; amrmg_solver_solve is (nlev-1, 0). */
int amrmg_solver_solve_levels(amrmg_solver* s, int l_max, int l_base, int zero_phi, double* hist, int hist_cap,
                              int* info);

/* AMRMultiGrid::computeAMRResidual(resid, phi, rhs) -> field 2; norm out */
int amrmg_solver_residual(amrmg_solver* s, int homogeneous, double* norm_out);
/* AMRMultiGrid::computeAMROperator(lphi, phi) -> field 2 */
int amrmg_solver_operator(amrmg_solver* s, int homogeneous);

/* ---- level operations (host buffers in/out; one upload, one level-wide
   launch sequence, one download) ---- */
/* AMRPoissonOp::applyOp / residual / relax / restrictResidual / prolongIncrement
   on multigrid depth `depth` of AMR level 0 (depth 0 = level 0) */
int amrmg_level_apply(amrmg_solver* s, int depth, const double* phi, double* out, int homogeneous);
/* norm(BoxLayoutData, interval, p) (BoxLayoutData.cpp:23-92) through AMRPoissonOp::norm
   (AMRPoissonOp.cpp:558-562) on MG depth `depth` of level 0: p = 0 max |x|, 1 sum |x|
   (FArrayBox::norm(1) column sums), 2 sqrt(sumPow(2)); x valid-packed */
int amrmg_level_norm(amrmg_solver* s, int depth, const double* x, int p, double* out);
int amrmg_level_residual(amrmg_solver* s, int depth, const double* phi, const double* rhs, double* out,
                         int homogeneous);
int amrmg_level_relax(amrmg_solver* s, int depth, double* phi, const double* rhs, int iterations);
int amrmg_level_restrict(amrmg_solver* s, int depth, const double* phi, const double* rhs, double* res_coarse);
int amrmg_level_prolong(amrmg_solver* s, int depth, double* phi, const double* e_coarse);
/* FORT_GSRBLAPLACIAN / FORT_GSRBHELMHOLTZ on AMR level lev, ghosts supplied */
int amrmg_kernel_gsrb(amrmg_solver* s, int lev, double* phi_ghosted, const double* rhs, int color);
/* AMRPoissonOp::homogeneousCFInterp on AMR level lev (ghosted in/out) */
int amrmg_homog_cf(amrmg_solver* s, int lev, double* phi_ghosted);
/* QuadCFInterp::coarseFineInterp(phi[lev], phi[lev-1]) (ghosted in/out) */
int amrmg_quad_cf(amrmg_solver* s, int lev, double* phi_ghosted, const double* coarse_valid);
/* FORT_AVERAGE of level lev onto coarsen(grids[lev], 2) */
int amrmg_average(amrmg_solver* s, int lev, const double* fine, double* coarse);
/* MultiGrid bottom: relax(nbottom) + BiCGStabSolver::solve on the deepest MG level
   (host-driven BiCGStab); info = {exit status, iterations} */
int amrmg_bottom_solve(amrmg_solver* s, double* corr, const double* res, int nbottom, double metric, double eps,
                       int* info);
/* The same on the device bottom solver the solves use (one workgroup, no host round
   trip per iteration; the deepest multigrid level must be stored as one array over
   its whole domain, with constant coefficients); info = {iterations} */
int amrmg_bottom_solve_device(amrmg_solver* s, double* corr, const double* res, int nbottom, double metric,
                              double eps, int* info);
/* BiCGStabSolver's public parameters (BiCGStabSolver.H:105-153) for every bottom solve of
   the solver (device and host-driven, FP64 and mixed, and amrmg_bottom_solve*):
   ip = {m_imax, m_numRestarts}, rp = {m_reps, m_hang, m_small}; the defaults are the
   reference constructor's (BiCGStabSolver.H:166-183): 80, 5, 1e-12, 1e-8, 1e-30. */
int amrmg_solver_set_bottom_params(amrmg_solver* s, const int* ip, const double* rp);

/* Autotuner of the one-launch relaxation: on every level that uses it, times its
   10 tile shapes (columns 64x8, 32x8, 64x4, 32x16, 128x4, 64x16 of 256 threads;
   32x16, 64x16 of 512; 64x32, 128x8 of 1024) x z-chunk lengths {auto, 8, 16, 32}
   on the solver's own fields and keeps the fastest for later solves (results are
   bitwise the same for every choice).  Writes up to cap records of 46 doubles:
   {0 AMR level / 1 multigrid depth, index, valid cells, chosen shape, chosen
   z-chunk, its ms per iteration, then ms of each shape x chunk}; *nrec = number
   of tuned levels. */
int amrmg_solver_tune_relax(amrmg_solver* s, double* table, int cap, int* nrec);
/* Forces tile shape (0..12) and z-chunk (0 = automatic) of the one-launch
   relaxation on every level. */
int amrmg_solver_set_relax_shape(amrmg_solver* s, int shape, int kc);
/* AMRPoissonOp::s_relaxMode (AMRPoissonOp.cpp:31, 600-632) for every operator of the
   solver: 0 looseGSRB (AMRPoissonOp.cpp:1333-1413), 1 levelGSRB (default), 2
   overlapGSRB (:1416-1521, as run without MPI), 3 levelGSRBLazy (:1524-1578), 4
   levelJacobi (:1579-1596, VCAMRPoissonOp2.cpp:762-786), 5 levelMultiColor
   (:1284-1331).  With variable coefficients only 1 and 4 are accepted, as
   VCAMRPoissonOp2 aborts in the others.  With any mode but 1 the bottom solve is
   the host-driven BiCGStab. */
int amrmg_solver_set_relax_mode(amrmg_solver* s, int mode);
/* Tile shape of the one-launch variable-coefficient relaxation on every multigrid
   depth: -1 automatic, 0 64x8, 1 32x8, 2 64x4 (256 threads), 3 64x16 (512
   threads); results are bitwise the same for every choice.  No counterpart in the
   reference (its GSRBHELMHOLTZVC runs per box and colour). */
int amrmg_solver_set_vc_relax_shape(amrmg_solver* s, int shape);
/* m_coefficient_average_type of VCAMRPoissonOp2Factory (VCAMRPoissonOp2.cpp:1048-1063)
   for the multigrid coarsening (re-averaged at once when coefficients are set, else
   applied by the next amrmg_solver_set_vc_coefs): 0
   CoarseAverage::arithmetic (default), 1 harmonic (averageToCoarseHarmonic,
   AverageF.ChF AVERAGEHARMONIC and AverageFaceF.ChF AVERAGEFACEHARMONIC). */
int amrmg_solver_set_vc_average(amrmg_solver* s, int type);

/* ---- batched single-level solves: many independent problems, each V-cycle stage
   of all of them in one launch (relaxation, restriction, prolongation, residual
   with per-problem max-norm, and every problem's bottom BiCGStab as one workgroup
   each).  Per problem the arithmetic is that of amrmg_solver_solve on it alone
   (AMRMultiGrid::solve with one level).  No counterpart in the reference, which
   solves one hierarchy at a time. ---- */
/* Heat-equation steppers, d phi/dt = L phi + S with the solver's operator as L
   (AMRPoissonOp: alpha I + beta Laplacian; VCAMRPoissonOp2: alpha a - beta div(b grad)).
   scheme 0: BackwardEuler (BackwardEuler.cpp:47-254): rhs = a phi^n + dt S, then
   alpha = 1, beta = -dt and AMRMultiGrid::solve from phi^n (zero_phi = 0) or zero.
   scheme 1: AMRTGA (AMRTGA.H:204-563): the two-stage TGA step with its mu1..mu4
   (setMu), two solves from phi^n.  The step reads phi^n from the solver's field 0
   and leaves phi^{n+1} there; S is set with amrmg_heat_set_source (valid-packed,
   zero by default).  Like the reference's steppers, a step resets alpha and beta of
   every operator of the solver.  info = {iterations, exitStatus} of the first solve,
   then of the second (TGA; zero for backward Euler). */
typedef struct amrmg_heat amrmg_heat;
int amrmg_heat_create(amrmg_solver* s, int scheme, amrmg_heat** out);
int amrmg_heat_set_source(amrmg_heat* h, const double* valid_packed);
int amrmg_heat_step(amrmg_heat* h, double dt, int zero_phi, int* info);
int amrmg_heat_destroy(amrmg_heat* h);

typedef struct amrmg_batch amrmg_batch;
/* nprob problems, each a one-level (ispec, rspec) as for amrmg_solver_create,
   ispecs back to back (14 + 6*nbox ints each), rspecs 9 doubles each; each
   problem's boxes must tile its domain.  precision: 0 FP64, 1 mixed. */
int amrmg_batch_create(amrmg_ctx* ctx, int nprob, const int* ispecs, const double* rspecs, int precision,
                       amrmg_batch** out);
int amrmg_batch_destroy(amrmg_batch* b);
/* sparams / rparams as for amrmg_solver_set_params (numMG must be 1) */
int amrmg_batch_set_params(amrmg_batch* b, const int* sparams, const double* rparams);
int amrmg_batch_set_rhs(amrmg_batch* b, int problem, const double* valid_packed);
int amrmg_batch_get_phi(amrmg_batch* b, int problem, double* valid_packed);
/* solves every problem from phi = 0; per-problem iteration counts and exit status */
int amrmg_batch_solve(amrmg_batch* b, int* iterations, int* exit_status);
int amrmg_batch_history(amrmg_batch* b, int problem, double* hist, int cap, int* n);
/* autotunes the relaxation tile shape per multigrid depth; ms per depth */
int amrmg_batch_tune_relax(amrmg_batch* b, double* ms, int cap, int* ndepth);
/* No counterpart in the reference: forces the relaxation tile shape (index into the
   solver's shape table, as amrmg_solver_set_relax_shape) and z-chunk of one multigrid
   depth in every lane; results are bitwise the same for every shape that applies, and
   a shape that does not apply to every problem of the depth is rejected. */
int amrmg_batch_set_relax_shape(amrmg_batch* b, int depth, int shape, int kc);
/* cell-relaxations and bottom BiCGStab iterations of the last solve; HBM bytes held */
int amrmg_batch_stats(amrmg_batch* b, double* cell_relax, long long* bottom_iters, long long* device_bytes);
/* No counterpart in the reference.  As amrmg_batch_create, with lanes = 1, 2 or 3.
   With 2, the problems whose multigrid bottom holds at least 16^3 cells lockstep on a
   stream of their own (a lane), so the other problems' V-cycles overlap their long
   bottom solves; with 3, bottoms of at least 27^3 cells get a third lane.  Each
   problem's result is the same in every case. */
int amrmg_batch_create_lanes(amrmg_ctx* ctx, int nproblems, const int* ispecs, const double* rspecs, int precision,
                             int lanes, amrmg_batch** out);
/* number of lanes and, per problem, its lane */
int amrmg_batch_lanes(amrmg_batch* b, int* lanes, int* lane_of_problem);
/* No counterpart in the reference (its V-cycles are host loops of kernel calls).
   on = 1: every V-cycle after the first is replayed from one captured HIP graph
   (the default); 0: direct launches; -1: leave as is.  The graph is not used while
   the context's profiler is on.  launches: graph replays so far. */
int amrmg_batch_set_graph(amrmg_batch* b, int on, long long* launches);
/* Reduction order of the bottom solves' dot products and norms (no counterpart in
   the reference): 0 (default) the reference's LevelDataOps / FArrayBox::sumPow order
   (results bitwise those of solving each problem alone on the reference's path), 1 a
   fixed pairwise tree per problem (faster bottom solves; each result is still the same
   alone and in any batch, and differs from order 0 by rounding only). */
int amrmg_batch_set_reduction_order(amrmg_batch* b, int order);

/* ---- serving: a request loop with continuous admission.  nslots worker threads
   each run one request at a time on its own HIP stream, so the solves of
   different requests overlap on the device.  When a slot frees, the first request
   in arrival order whose HBM footprint fits in the budget beside the running ones is
   admitted; a head request that does not fit is passed by at most 8 later ones, then
   admission waits for it.  The footprint is the bytes the request's solver holds
   (amrmg_solver_device_bytes), computed from its layouts before it is created;
   hierarchies stay device-resident after a solve and are reused by later requests
   with the same spec (least recently used ones are evicted to make room).  Solver
   parameters are those set when a request is admitted.  No counterpart in the
   reference. ---- */
typedef struct amrmg_serve amrmg_serve;
int amrmg_serve_create(int device, int nslots, long long hbm_budget_bytes, int precision, amrmg_serve** out);
int amrmg_serve_set_params(amrmg_serve* e, const int* sparams, const double* rparams);
/* Lanes of the request loop (no counterpart in the reference): requests of at least
   large_cells AMR cells are "large", and at most max_large of them run at once, so
   the remaining slots stay free for smaller requests (default: max_large = slots,
   one lane).  Waiting requests are reclassified. */
int amrmg_serve_set_lanes(amrmg_serve* e, int max_large, long long large_cells);
/* copies the spec and the valid-packed rhs; the arrival time is the call's time */
int amrmg_serve_submit(amrmg_serve* e, const int* ispec, const double* rspec, const double* rhs, long long* id);
/* blocks until request id is solved; phi valid-packed (may be null); times =
   {arrival, admission, completion} in s since the engine started; info =
   {iterations, exit status}.  One wait per request: a second waiter on the same id
   gets an error, and the id is forgotten once the wait returns. */
int amrmg_serve_wait(amrmg_serve* e, long long id, double* phi, double* times, int* info);
/* {max HBM bytes held, hierarchies created, reused, evicted, max requests running at once} */
int amrmg_serve_stats(amrmg_serve* e, long long* stats, int n);
int amrmg_serve_destroy(amrmg_serve* e);

/* ---- operator level: device-resident levels, fields, operators and copiers ----
   The reference's operator plug-in API is AMRPoissonOp, a LinearOp / MGLevelOp /
   AMRLevelOp over LevelData<FArrayBox> (LinearSolver.H:27-149, MultiGrid.H:88-287,
   AMRMultiGrid.H:38-226) that Chombo's AMRMultiGrid, MultiGrid and BiCGStabSolver
   drive.  These entry points give a host operator class the same operations on
   device-resident data, one call per virtual:
     amrmg_level  a DisjointBoxLayout with its ProblemDomain, resident in HBM;
     amrmg_field  a LevelData<FArrayBox> on a level (one component, one ghost cell);
     amrmg_op     one AMRPoissonOp (alpha I + beta Laplacian) on a level;
     amrmg_plan   a Copier between two levels (copyTo, copyToZero) or an exchange.
   Every operation is queued on the context's stream in call order; those returning
   a host scalar (dot, norm) wait for it.  Operands are checked against the level
   each operation assumes (a field on another level is an error). */
typedef struct amrmg_level amrmg_level;
typedef struct amrmg_field amrmg_field;
typedef struct amrmg_op amrmg_op;
typedef struct amrmg_plan amrmg_plan;

/* DisjointBoxLayout::define + ProblemDomain (DisjointBoxLayout.H:30, ProblemDomain.H:141):
   boxes 6 ints each (lo0 lo1 lo2 hi0 hi1 hi2), sorted by Box::operator<, disjoint;
   domain 6 ints.  merge = 1 stores a level whose boxes tile a rectangle as one array
   (then a box's ghost cell inside another box is that box's valid cell, as after an
   exchange); 0 gives every box its own FAB. */
int amrmg_level_register(amrmg_ctx* ctx, const int* boxes, int nbox, const int* domain, int merge,
                         amrmg_level** out);
int amrmg_level_destroy(amrmg_level* L);
/* coarsen(DisjointBoxLayout, 2) (DisjointBoxLayout.H:354) of a level whose boxes are
   coarsenable by 2, with the coarsened domain */
int amrmg_level_coarsen(amrmg_level* L, amrmg_level** out);
/* valid cells (ghosted = 0) or cells of the ghosted FABs (1), summed over boxes */
long long amrmg_level_cells(amrmg_level* L, int ghosted);

/* LevelData<FArrayBox>(grids, 1, IntVect::Unit), every cell zero */
int amrmg_field_create(amrmg_level* L, amrmg_field** out);
/* LevelData<FArrayBox>(grids, ncomp, IntVect::Unit), ncomp 1..16, every cell zero.
   Transfers move box by box, each box component by component; the LevelDataOps
   calls act on every component, dot products and norms run over all of them in
   FArrayBox order; operators and copiers apply to each component (fields of one
   operation must have the same component count). */
int amrmg_field_create_ncomp(amrmg_level* L, int ncomp, amrmg_field** out);
int amrmg_field_ncomp(const amrmg_field* f);
int amrmg_field_destroy(amrmg_field* f);
/* FArrayBox-order transfers, box by box in layout order: ghosted = 0 moves each
   box's valid cells (Fortran order), 1 each box's FAB (the box grown by one cell;
   on a merged level upload skips ghost cells that are another box's valid cells) */
int amrmg_field_upload(amrmg_field* f, const double* host, int ghosted);
int amrmg_field_download(amrmg_field* f, double* host, int ghosted);

/* LevelDataOps<FArrayBox> (LevelDataOps.H:93-294) on valid cells: assignLocal (every
   cell of the FABs), incr (lhs += x*scale), axby (lhs = x*a + y*b), scale, setVal
   (setToZero is setVal 0); dotProduct (per box in Fortran order, boxes in layout
   order) and norm(BoxLayoutData, p) (BoxLayoutData.cpp:23-92; p = 0 max, 1, 2) */
int amrmg_field_assign(amrmg_field* lhs, const amrmg_field* rhs);
int amrmg_field_incr(amrmg_field* lhs, const amrmg_field* x, double scale);
int amrmg_field_axby(amrmg_field* lhs, const amrmg_field* x, const amrmg_field* y, double a, double b);
int amrmg_field_scale(amrmg_field* x, double s);
int amrmg_field_setval(amrmg_field* x, double v);
int amrmg_field_dot(const amrmg_field* a, const amrmg_field* b, double* out);
int amrmg_field_norm(const amrmg_field* x, int p, double* out);
/* amrmg_field_dot without any device allocation.  No counterpart in the reference:
   it is for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops, hosts that forbid allocation).  The per-box partial
   sums go to a fixed device array of 8192 doubles per GPU (held until the process
   ends), so levels of more than 8192 boxes are rejected with an error; the result
   is bitwise amrmg_field_dot's.  All calls on one GPU share that array: calls that
   overlap in time on one GPU race. */
int amrmg_field_dot_noalloc(const amrmg_field* a, const amrmg_field* b, double* out);

/* AMRPoissonOp::define (AMRPoissonOp.cpp:60-221) on level L: dx, dxCrse (> 0: the
   level has a coarser level, and homogeneousCFInterp fills its coarse-fine ghosts
   from its own CFRegion; <= 0: no coarse-fine faces), alpha, beta, and
   ConstDiriNeumBC-style faces bc_type[6] / bc_val[6] (lo x, y, z, then hi; 0 Neumann,
   1 Dirichlet linear, 2 Dirichlet quadratic).  coarser: the operator of the next
   coarser AMR level (ratio 2) or NULL; with it the operator gets QuadCFInterp,
   AMRRestrictS / AMRProlongS through coarsen(L, 2), and the coarser operator gets
   this level's reflux registers and zeroCovered (as AMRnewOp wires them).  An
   operator must be destroyed before its coarser one. */
int amrmg_op_create(amrmg_level* L, amrmg_op* coarser, double dx, double dxCrse, double alpha, double beta,
                    const int* bc_type, const double* bc_val, amrmg_op** out);
/* MGnewOp at the next depth (AMRPoissonOp.cpp:1923-1989): an operator on coarsen(L, 2)
   with 2 dx, the same dxCrse and coefficients, whose level is *level_out (owned by
   the new operator); the parent's restrictResidual / prolongIncrement target it */
int amrmg_op_create_coarser(amrmg_op* op, amrmg_op** out, amrmg_level** level_out);
/* the level of coarsen(L, 2) of an operator with a coarser AMR level
   (createCoarsened: resC, AMRRestrictS / AMRProlongS temporaries), owned by op */
int amrmg_op_coarsened_level(amrmg_op* op, amrmg_level** out);
int amrmg_op_destroy(amrmg_op* op);
/* AMRPoissonOp::setAlphaAndBeta (AMRPoissonOp.cpp:1103-1108) */
int amrmg_op_set_alpha_beta(amrmg_op* op, double alpha, double beta);
/* s_relaxMode of this operator (0..5, as amrmg_solver_set_relax_mode) */
int amrmg_op_set_relax_mode(amrmg_op* op, int mode);

/* LinearOp / MGLevelOp virtuals of AMRPoissonOp: residual (:224-236), applyOp
   (:332-340), preCond (:304-329), relax (:600-632), restrictResidual (:651-703),
   prolongIncrement (:706-733) */
int amrmg_op_residual(amrmg_op* op, amrmg_field* lhs, amrmg_field* phi, const amrmg_field* rhs, int homogeneous);
int amrmg_op_apply(amrmg_op* op, amrmg_field* lhs, amrmg_field* phi, int homogeneous);
int amrmg_op_precond(amrmg_op* op, amrmg_field* cor, const amrmg_field* residual);
int amrmg_op_relax(amrmg_op* op, amrmg_field* correction, const amrmg_field* residual, int iterations);
int amrmg_op_restrict_residual(amrmg_op* op, amrmg_field* res_coarse, amrmg_field* phi, const amrmg_field* rhs);
int amrmg_op_prolong_increment(amrmg_op* op, amrmg_field* phi, const amrmg_field* correct_coarse);
/* AMRLevelOp virtuals: AMRResidual / AMROperator with their NC (phi_coarse NULL) and
   NF (phi_fine NULL) forms (:736-858); AMRRestrictS (:877-919); AMRProlongS
   (:959-990); AMRUpdateResidual (:1045-1054, coarse_correction NULL on the base
   level); assignCopier (:486-492: lhs on this level <- rhs on the finer operator's
   coarsened level); zeroCovered (:495-502); localMaxNorm (:565-576) */
int amrmg_op_amr_residual(amrmg_op* op, amrmg_field* res, amrmg_field* phi_fine, amrmg_field* phi,
                          amrmg_field* phi_coarse, const amrmg_field* rhs, int homogeneous, amrmg_op* finer);
int amrmg_op_amr_operator(amrmg_op* op, amrmg_field* lphi, amrmg_field* phi_fine, amrmg_field* phi,
                          amrmg_field* phi_coarse, int homogeneous, amrmg_op* finer);
int amrmg_op_amr_restrict(amrmg_op* op, amrmg_field* res_coarse, const amrmg_field* residual,
                          amrmg_field* correction, amrmg_field* coarse_correction, amrmg_field* scratch);
int amrmg_op_amr_prolong(amrmg_op* op, amrmg_field* correction, const amrmg_field* coarse_correction,
                         amrmg_field* temp);
int amrmg_op_amr_update_residual(amrmg_op* op, amrmg_field* residual, amrmg_field* correction,
                                 amrmg_field* coarse_correction);
int amrmg_op_assign_copier(amrmg_op* op, amrmg_field* lhs, const amrmg_field* rhs);
int amrmg_op_zero_covered(amrmg_op* op, amrmg_field* lhs);
int amrmg_op_max_norm(amrmg_op* op, const amrmg_field* x, double* out);
/* the ghost-cell and box kernels the virtuals are made of: exchange (exchangeNoOverlap
   with the trimmed Copier), the BC functor on every box (homogeneous or with the face
   values), homogeneousCFInterp (INTERPHOMO / INTERPHOMOLINEAR), QuadCFInterp::
   coarseFineInterp from the coarser operator's level, one FORT_GSRBLAPLACIAN /
   FORT_GSRBHELMHOLTZ colour pass on every box with the ghosts as they are, and
   reflux (:1117-1201) of lphi with the finer level's phi */
int amrmg_op_exchange(amrmg_op* op, amrmg_field* phi);
int amrmg_op_bc(amrmg_op* op, amrmg_field* phi, int homogeneous);
/* this operator's inhomogeneous BC values from fn (as amrmg_solver_set_bc_values; a
   multigrid child made afterwards by amrmg_op_create_coarser inherits fn) */
int amrmg_op_set_bc_values(amrmg_op* op, amrmg_bc_value_fn fn, void* user);
/* components of the fields this operator's inhomogeneous boundary values serve
   (default 1): a value function fills one value per component, taken again here when
   one is set.  An inhomogeneous ghost fill of a component beyond it is an error. */
int amrmg_op_set_ncomp(amrmg_op* op, int ncomp);
int amrmg_op_cf_homog(amrmg_op* op, amrmg_field* phi);
int amrmg_op_cf_inhomog(amrmg_op* op, amrmg_field* phi, const amrmg_field* phi_coarse);
int amrmg_op_gsrb(amrmg_op* op, amrmg_field* phi, const amrmg_field* rhs, int color);
int amrmg_op_reflux(amrmg_op* op, amrmg_field* phi_fine, amrmg_field* phi, amrmg_field* lphi, amrmg_op* finer);

/* Copier (Copier.H:36-57, Copier.cpp:286-330) from level src to level dst: items are
   MotionItems of 8 ints (from box, to box, region lo0..2, hi0..2) in the order the
   copies run, or NULL to build Copier::define(src, dst) here (every intersection of a
   src box with a dst box, dst boxes in order).  amrmg_plan_exchange is the exchange
   Copier of one level (exchangeDefine + trimEdges, Copier.cpp:1464-1541) from items or
   NULL.  Regions must lie in the from box's valid cells and the to box's FAB. */
int amrmg_plan_copy(amrmg_level* src, amrmg_level* dst, const int* items, int nitems, amrmg_plan** out);
int amrmg_plan_exchange(amrmg_level* L, const int* items, int nitems, amrmg_plan** out);
/* LevelData::copyTo (op 0, dst <- src over each item's region; LevelDataI.H:221-253)
   or LevelDataOps::copyToZero (op 1, dst = 0 there, src unused; LevelDataOps.H:231-252) */
int amrmg_plan_exec(amrmg_plan* p, amrmg_field* dst, const amrmg_field* src, int op);
int amrmg_plan_destroy(amrmg_plan* p);
/* wait for every queued operation of the context */
int amrmg_ctx_sync(amrmg_ctx* ctx);

/* Plain device copy of nbytes (reads nbytes, writes nbytes), iters launches;
   average ms per launch.  Calibrates HBM byte counters and the copy bandwidth. */
int amrmg_copy_probe(amrmg_ctx* ctx, long long nbytes, int iters, double* ms_out);

/* Host-only check: builds every device table of a hierarchy (ghost rules,
   copiers, coarse-fine stencils, reflux lists) without touching a device and
   verifies each address a kernel will form is inside its field.  stats (up to 8):
   {levels, multigrid depths, ghost patches, CF coarse records, CF fine records,
   reflux cells, tiles, valid cells}. */
int amrmg_plan_check(const int* ispec, const double* rspec, long long* stats, int nstats);
/* the same with refinement ratios as for amrmg_solver_create_ratios */
int amrmg_plan_check_ratios(const int* ispec, const double* rspec, const int* ref_ratios, long long* stats,
                            int nstats);

/* ---- event profiler over kernel classes (0 GSRB, 1 residual, 2 applyOp,
   3 restrict, 4 prolong, 5 average, 6 ghost fill, 7 BLAS-1, 8 reductions,
   9 CF interpolation, 10 reflux, 11 copier).  class_mask: bit k times class k;
   -1 times all, 0 turns the profiler off.  Events go on the library's stream. ---- */
int amrmg_profile_enable(amrmg_ctx* ctx, int class_mask);
int amrmg_profile_reset(amrmg_ctx* ctx);
int amrmg_profile_read(amrmg_ctx* ctx, int kernel_class, double* ms, double* bytes, long long* launches);
/* work units of a class since the last reset: cell-relaxations (one red + one black
   update of a cell) for class 0; 0 for the others.  bytes above are algorithmic
   bytes: a relaxation that starts from a correction known to be zero reads no old
   values and counts 16 of the 24 bytes per cell-relax in FP64. */
int amrmg_profile_read_units(amrmg_ctx* ctx, int kernel_class, double* units);

#ifdef __cplusplus
}
#endif
#endif
