// Batched solver: many independent single-level problems (AMRMultiGrid with
// l_max = l_base = 0, i.e. MultiGrid V-cycles with a BiCGStab bottom) solved
// together, one launch per V-cycle stage for all of them.
//
// Per problem the arithmetic is the single-problem path's, operation for
// operation: solveNoInitResid's outer loop and convergence tests
// (AMRMultiGrid.H:935-1124), MultiGrid::cycle (MultiGrid.H:630-696) with the
// fused relaxation, RESTRICTRES, PROLONG, and BiCGStabSolver::solve
// (BiCGStabSolver.H:199-519) at the bottom with the reference's reduction orders.  In
// FP64 a problem's result is therefore bitwise the result of solving it alone;
// with precision 1 the V-cycle runs in FP32 under the FP64 outer iteration, as in
// the single-problem mixed mode.
//
// Layout: problems are sorted by their number of multigrid depths (descending),
// so the problems present at depth d are a prefix of those at depth d-1.  Each
// depth holds, for every present problem, one merged ghosted array over its
// (coarsened) domain, all in one slab per field; the box tables at every depth
// list the problems in the same order, so problem p has index p at each depth.
#pragma once
#include <memory>
#include <vector>

#include "amrpoissonop.hpp"
#include "fused.hpp"
#include "ops.hpp"
#include "plans.hpp"

namespace amrmg {

struct BatchProblemSpec {
  Box domain;                 // single level
  DBL boxes;                  // its DisjointBoxLayout (fixes the multigrid depth and the bottom boxes)
  double dx = 1, alpha = 0, beta = 1;
  BCSpec bc{};
};

// per problem and depth: the operator's coefficients and face rules
struct BatchCoef {
  double dxinv, lambda, alpha, beta, dx;
  int lap;          // 1: GSRBLAPLACIAN / OPERATORLAP forms (alpha == 0 and beta == 1)
  int kind[3][2];   // face rule: 0 Neumann, 1 Dirichlet
  double val[3][2]; // face values (outer residual only; every other use is homogeneous)
};

// the bottom of one problem: its level and boxes at its deepest depth
struct BatchBottom {
  DBox rect;        // merged array of the level
  int box0, nbox;   // its boxes in the bottom box table
  int depth;
  long long cells;
};

// The bottom solve of one hierarchy's multigrid (BiCGStabSolver::solve on the
// deepest level, homogeneous) run on the device as one workgroup: the control flow,
// level operations and reduction orders of BiCGStabSolver, with no host round
// trip per iteration.  With T = float it widens the correction and the residual
// and solves in FP64, as PromotedBottomSolver does.  Levels that are not stored
// merged over a whole domain use the host-driven BiCGStabSolver (FP64 only).
template <class T>
class DeviceBottomSolver : public LinearSolver<LevelData<T>> {
 public:
  explicit DeviceBottomSolver(Context* c) : m_ctx(c) {}
  ~DeviceBottomSolver() override;
  DeviceBottomSolver(const DeviceBottomSolver&) = delete;
  DeviceBottomSolver& operator=(const DeviceBottomSolver&) = delete;
  void define(LinearOp<LevelData<T>>* op, bool homogeneous) override;
  void solve(LevelData<T>& phi, const LevelData<T>& rhs) override;
  void setConvergenceMetrics(double metric, double tol) override;
  long long totalIterations();
  bool onDevice() const { return m_device; }
  BiCGStabSolver<LevelData<T>> m_fallback;   // parameters (m_imax, m_reps, ...) are read from here

 private:
  Context* m_ctx;
  bool m_device = false;
  const DeviceLayout* m_lay = nullptr;
  const void* m_boundOp = nullptr;
  const void* m_ap = nullptr;     // the bound AMRPoissonOp<T>
  BatchCoef m_coef{};             // coefficients uploaded for it
  long long m_vs = 0;             // elements per Krylov vector of the bound level
  double m_metric = -1.0, m_eps = 1e-6;
  void *d_bottom = nullptr, *d_bcoef = nullptr, *d_bbox = nullptr, *d_ptrs = nullptr, *d_soff = nullptr;
  void *d_active = nullptr, *d_stats = nullptr, *h_stage = nullptr;
  double *d_scratch = nullptr, *d_metric = nullptr;
  void release();
};

// valid cells of a problem's multigrid bottom (its deepest coarsening)
long long batchBottomCells(const BatchProblemSpec& P);

class BatchSolver {
 public:
  BatchSolver(Context* ctx, std::vector<BatchProblemSpec> problems, int precision);
  ~BatchSolver();
  BatchSolver(const BatchSolver&) = delete;
  BatchSolver& operator=(const BatchSolver&) = delete;

  void setSolverParameters(int pre, int post, int bottom, int numMG, int iterMax, int imin, double eps, double hang,
                           double normThresh);
  // host valid-packed data of problem q (in the caller's order)
  void setRhs(int q, const double* valid_packed);
  void getPhi(int q, double* valid_packed) const;
  // solve all problems from phi = 0; per problem iterations / exit status / residual history
  void solve();
  int iterations(int q) const { return m_iter[m_slot[q]]; }
  int exitStatus(int q) const { return m_exit[m_slot[q]]; }
  const std::vector<double>& history(int q) const { return m_hist[m_slot[q]]; }
  long long bottomIterations() const { return m_bottomIters; }
  // autotunes the batched relaxation's tile shape per depth; returns ms per depth
  std::vector<double> tuneRelax();
  // tile shape (index into ops::kFusedShapeTable) and z-chunk (0 automatic) of depth d
  void setRelaxShape(int d, int shape, int kc);
  int relaxShape(int d) const { return m_depth[d].shape; }
  int relaxChunk(int d) const { return m_depth[d].kc; }
  int numProblems() const { return m_np; }
  int numDepths() const { return (int)m_depth.size(); }
  long long deviceBytes() const { return m_bytes; }
  long long cells(int q) const;
  // profiler work units: cell-relaxations of the last solve
  double cellRelaxations() const { return m_cellRelax; }
  // V-cycles replayed from one captured HIP graph (default on; the graph is not
  // used while the context's profiler is on, so per-kernel timing stays exact)
  void setGraph(bool on);
  // reduction order of the bottom solves' dot products and norms: the reference's
  // (default) or a fixed tree per problem (same result alone and in any batch,
  // different from the reference order by rounding)
  void setTreeOrder(bool on);
  bool shapeApplies(int d, int shape) const;
  bool treeOrder() const { return m_treeOrder; }
  // run later launches on another context's stream (same device)
  void setContext(Context* c);
  Context* context() const { return m_ctx; }
  bool graph() const { return m_useGraph; }
  long long graphLaunches() const { return m_graphLaunches; }

 private:
  struct Depth {
    int np = 0;                                  // problems present (a prefix)
    std::vector<DBox> h_rect;                    // per problem: merged array of its level
    DBox* d_rect = nullptr;
    std::vector<BatchCoef> h_coef;
    BatchCoef* d_coef = nullptr;
    long long slab = 0;
    // tiles (64 x 4 x 1 cells, one 256-thread workgroup each) in problem order
    std::vector<int> tileStart;                  // per problem, prefix; [np] = total
    Tile* d_tiles = nullptr;
    int lx = 6, ly = 2;
    // column tiles (64 x 4 x kColumnPlanes cells) of the finest depth
    std::vector<int> ctileStart;
    Tile* d_ctiles = nullptr;
    // fused relaxation work items
    std::vector<ops::FusedArgs> h_items;
    ops::FusedArgs* d_items = nullptr;
    std::vector<int> wStart;                     // per problem prefix of work items
    int* d_wStart = nullptr;
    int shape = 0, kc = 0;
    long long cellsPresent = 0;                  // valid cells of the problems present
    std::vector<long long> cellStart;            // per problem prefix of valid cells
    void* corr = nullptr;                        // inner precision fields
    void* res = nullptr;
    void* tmp = nullptr;
  };
  Context* m_ctx;
  int m_prec;                                    // 0 FP64, 1 mixed
  int m_np;
  std::vector<BatchProblemSpec> m_prob;          // in slot (sorted) order
  std::vector<int> m_slot;                       // caller index -> slot
  std::vector<int> m_ndepth;                     // per slot
  std::vector<Depth> m_depth;
  double* m_phi = nullptr;                       // FP64 depth-0 fields
  double* m_rhs = nullptr;
  double* m_norm = nullptr;                      // per slot max-norm (device)
  double* h_norm = nullptr;                      // pinned mirror
  int* d_active = nullptr;
  std::vector<int> h_active;
  // bottom
  std::vector<BatchBottom> h_bottom;
  BatchBottom* d_bottom = nullptr;
  BatchCoef* m_bcoef = nullptr;                  // per slot: coefficients at its bottom depth
  void** d_corrPtr = nullptr;                    // per slot: correction / residual fields of its bottom depth
  void** d_resPtr = nullptr;
  std::vector<DBox> h_bbox;                      // bottom boxes (addresses in each bottom's merged array)
  DBox* d_bbox = nullptr;
  double* d_bscratch = nullptr;                  // 10 FP64 vectors per problem (bottom BiCGStab)
  std::vector<long long> h_bscratchOff;
  long long* d_bscratchOff = nullptr;
  double* d_metric = nullptr;                    // per slot: convergence metric and eps of the bottom solve
  int* d_bstats = nullptr;                       // per slot: BiCGStab iterations (accumulated)
  // parameters
  int m_pre = 2, m_post = 2, m_nbottom = 2, m_numMG = 1, m_iterMax = 100, m_imin = 5, m_iterMin = -1;
  double m_eps = 1e-10, m_hang = 1e-15, m_normThresh = 1e-30;
  // per slot results
  std::vector<int> m_iter, m_exit;
  std::vector<std::vector<double>> m_hist;
  long long m_bottomIters = 0;
  long long m_bytes = 0;
  double m_cellRelax = 0;

  void alloc(void** p, size_t bytes);
  template <class T> void vcycle();
  template <class T> void cycle();          // vcycle, through the captured graph when enabled
  double cycleUnits() const;                // cell-relaxations of one V-cycle of the active problems
  void dropGraph();
  bool m_useGraph = true;
  bool m_treeOrder = false;
  hipGraphExec_t m_graph = nullptr;
  long long m_graphLaunches = 0;
  template <class T> void relaxDepth(int d, int np, int iters, bool fromZero);
  template <class T> void residualNorm(bool homog);
  template <class T> void update();
  template <class T> void runSolve();
  void uploadActive();
};

}  // namespace amrmg
