// AMRPoissonOp / AMRPoissonOpFactory / AMRMultiGrid on device-resident data.
//
// Mirrors lib/src/AMRElliptic/AMRPoissonOp.{H,cpp} (operator alpha*I + beta*Laplacian
// on one AMR or multigrid level) and AMRMultiGrid.H (Martin-Cartwright AMR V-cycle).
// Every per-box FORT_* loop of the reference is one kernel launch over the level;
// Copier exchanges, BCs and coarse-fine interpolation are table-driven kernels.
#pragma once
#include <functional>
#include <memory>
#include <vector>

#include "ops.hpp"
#include "plans.hpp"
#include "solver.hpp"
#include "vc.hpp"

namespace amrmg {

template <class Real>
class AMRPoissonOp : public AMRLevelOp<LevelData<Real>> {
 public:
  using T = LevelData<Real>;
  Context* m_ctx = nullptr;
  const DeviceLayout* m_lay = nullptr;        // this level
  const DeviceLayout* m_mgCoarse = nullptr;   // coarsen(this, 2): multigrid child level
  Box m_domain;
  double m_dx = 1, m_dxCrse = -1, m_alpha = 0, m_beta = 1;
  BCSpec m_bc{};
  int m_ncomp = 1;                            // components of the fields (one BC value each)
  ops::GhostPlan m_ghost;
  // one-launch relaxation on merged levels (ops::gsrb_fused); m_relaxTmp is the
  // second buffer it relaxes into, swapped with the relaxed field afterwards
  bool m_fused = false;
  ops::FaceRules m_faces{};
  int m_relaxShape = 0, m_relaxKc = 0;        // tile shape and z-chunk (0: automatic), see tuneRelax
  LevelData<Real> m_relaxTmp;
  LevelData<Real> m_relaxTmpN;                // the same for fields of several components
  // coarser AMR level (QuadCFInterp + restriction/prolongation copiers)
  const DeviceLayout* m_coarserLay = nullptr;
  const DeviceLayout* m_cfLay = nullptr;      // coarsen(this level, refToCoarser): resC layout
  int m_refToCoarser = 2;
  CFCoarse* d_cfc = nullptr; CFGen* d_cfg = nullptr; CFFine* d_cff = nullptr;
  int n_cfc = 0, n_cfg = 0, n_cff = 0;
  Real* d_deriv = nullptr;
  ops::MotionPlan m_resCopier;                // resC (cfLay) -> coarser level
  ops::MotionPlan m_revCopier;                // coarser level -> resC (cfLay)
  // finer AMR level (reflux, zeroCovered)
  const DeviceLayout* m_finerLay = nullptr;
  int m_refToFiner = 2;
  RefluxCell* d_rcell = nullptr; RefluxCoarse* d_rc = nullptr; RefluxFine* d_rf = nullptr;
  int n_rcell = 0;
  const ops::MotionPlan* m_zeroPlan = nullptr; // finer op's m_resCopier (cfLay(finer) -> this)
  const DeviceLayout* m_finerCfLay = nullptr;
  // Variable coefficients (VCAMRPoissonOp2, one level): when set, the operator is
  // alpha a phi - beta div(b grad phi) with these coefficients and relaxes with
  // GSRBHELMHOLTZVC; the one-launch constant-coefficient relaxation is not used.
  std::shared_ptr<VCCoefs<Real>> m_vc;
  int m_vcShape = -1;                              // tile shape of the one-launch VC relaxation (-1 automatic)
  // AMRPoissonOp::s_relaxMode (AMRPoissonOp.cpp:31, 600-632): 0 looseGSRB, 1 levelGSRB,
  // 2 overlapGSRB, 3 levelGSRBLazy, 4 levelJacobi, 5 levelMultiColor
  int m_relaxMode = 1;
  LevelData<Real> m_jacobiTmp;
  void levelJacobi(T& phi, const T& rhs);
  void levelGSRBLazy(T& phi, const T& rhs);
  void levelMultiColor(T& phi, const T& rhs);
  void looseGSRB(T& phi, const T& rhs);
  void overlapGSRB(T& phi, const T& rhs);
  // tiles over each box (merged levels tile their rectangle otherwise), for looseGSRB / overlapGSRB
  Tile* d_boxTiles = nullptr;
  int n_boxTiles = 0;
  const Tile* boxTiles();
  const Real* relaxSnapshot(T& phi);
  // reflux with variable coefficients: each contribution's face in the face arrays
  std::vector<RefluxFace> m_rcFace, m_rfFace;
  VCFaceIdx* d_vcRcFace = nullptr;
  VCFaceIdx* d_vcRfFace = nullptr;
  const VCCoefs<Real>* m_vcRefluxFine = nullptr;   // the finer coefficients the tables were built for
  void setupVCReflux(const VCCoefs<Real>& fine);
  void setVC(std::shared_ptr<VCCoefs<Real>> c);   // also computes lambda (computeLambda)
  void resetLambda();                              // after alpha, beta or the coefficients change

  AMRPoissonOp() = default;
  ~AMRPoissonOp() override;

  // ---- LinearOp ----
  void residual(T& lhs, T& phi, const T& rhs, bool homogeneous = false) override;
  void preCond(T& cor, const T& residual) override;
  void applyOp(T& lhs, T& phi, bool homogeneous = false) override;
  void create(T& lhs, const T& rhs) override { lhs.define(rhs.lay, rhs.nc); }
  void assignLocal(T& lhs, const T& rhs) override;
  double dotProduct(const T& a, const T& b) override;
  void incr(T& lhs, const T& x, double scale) override;
  // lhs = lhs + x * 1.0 then x = lhs on the valid cells (incr + assignLocal fused)
  void incrAssign(T& lhs, T& x);
  void incrFromZero(T& lhs, const T& x);
  void incrAssignUpdate(T& m, T& d, T& phi);
  void incrPhiFromZero(T& phi, const T& x);
  void axby(T& lhs, const T& x, const T& y, double a, double b) override;
  void scale(T& lhs, double s) override;
  double norm(const T& x, int ord) override;
  void setToZero(T& lhs) override;
  double dx() const override { return m_dx; }
  // ---- MGLevelOp ----
  void createCoarser(T& coarse, const T& fine, bool ghosted) override;
  void relax(T& correction, const T& residual, int iterations) override;
  void relaxFromZero(T& correction, const T& residual, int iterations) override;
  void restrictResidual(T& resCoarse, T& phiFine, const T& rhsFine) override;
  void prolongIncrement(T& phiThisLevel, const T& correctCoarse) override;
  // ---- AMRLevelOp ----
  int refToCoarser() override { return m_refToCoarser; }
  void AMRResidual(T& res, T& phiFine, T& phi, T& phiCoarse, const T& rhs, bool homog, AMRLevelOp<T>* finerOp) override;
  void AMRResidualNC(T& res, T& phiFine, T& phi, const T& rhs, bool homog, AMRLevelOp<T>* finerOp) override;
  void AMRResidualNF(T& res, T& phi, T& phiCoarse, const T& rhs, bool homog) override;
  void AMROperator(T& LofPhi, T& phiFine, T& phi, T& phiCoarse, bool homog, AMRLevelOp<T>* finerOp) override;
  void AMROperatorNC(T& LofPhi, T& phiFine, T& phi, bool homog, AMRLevelOp<T>* finerOp) override;
  void AMROperatorNF(T& LofPhi, T& phi, T& phiCoarse, bool homog) override;
  void AMRRestrictS(T& resCoarse, const T& residual, T& correction, T& coarseCorrection, T& scratch) override;
  void AMRProlongS(T& correction, const T& coarseCorrection, T& temp) override;
  void AMRUpdateResidual(T& residual, T& correction, T& coarseCorrection) override;
  void createCoarsened(T& lhs, const T& rhs, int refRat) override;
  void assignCopier(T& lhs, const T& rhs) override;
  void zeroCovered(T& lhs) override;
  double localMaxNorm(const T& x) override;

  // pieces used above and by the C ABI
  void fillGhosts(T& phi, bool doCF, bool homogBC);
  void residualI(T& lhs, T& phi, const T& rhs, bool homog);
  void applyOpI(T& lhs, T& phi, bool homog);
  void levelGSRB(T& phi, const T& rhs);
  void homogeneousCFInterp(T& phi);
  void coarseFineInterp(T& phif, const T& phic);   // QuadCFInterp of this level from the coarser one
  void reflux(T& phiFine, T& phi, T& res, AMRPoissonOp* finerOp);
  void residualFused(T& res, T& phiFine, T& phi, const T& rhs, bool homog, AMRPoissonOp* finerOp);
  void localMaxNormAsync(const T& x, int slot, bool accumulate);
  // residual (homog CF fill: residual's) or AMRResidualNF with the max-norm of the
  // result folded into the context's scalar slot, as localMaxNormAsync would add it
  // (one component, constant coefficients)
  bool canFuseNorm(const T& res) const { return res.nc == 1 && !m_vc; }
  void residualNorm(T& lhs, T& phi, const T& rhs, bool homog, int slot, bool accumulate);
  void AMRResidualNFNorm(T& res, T& phi, T& phiCoarse, const T& rhs, bool homog, int slot, bool accumulate);
  // Autotuner: times the one-launch relaxation's tile shapes on (phi, rhs) of this
  // level and keeps the fastest.  No-op (returns false) when the level relaxes
  // with the four-launch sequence.  times: kFusedShapeCount x 4 ms values.
  bool tuneRelax(const T& phi, const T& rhs, double* times);
};

// AMRPoissonOp::define on one level (ghost rules from the level's own CFRegion when
// dxCrse > 0, the one-launch relaxation where it applies); mgCoarse: the multigrid
// child layout coarsen(level, 2) or null.
template <class Real>
void defineLevelOp(AMRPoissonOp<Real>* op, Context* ctx, const DeviceLayout* lay, const Box& domain, double dx,
                   double dxCrse, double alpha, double beta, const BCSpec& bc, const DeviceLayout* mgCoarse);
template <class Real>
void setLevelOpBCValues(AMRPoissonOp<Real>* op, BCValueFn fn, void* user);
// The coarse-fine tables between op's level and the coarser AMR level of coarserOp
// (ratio 2): QuadCFInterp, the restriction / prolongation copiers through cfLay =
// coarsen(op's level, 2), and coarserOp's reflux tables and zeroCovered plan.
template <class Real>
void linkCoarser(AMRPoissonOp<Real>* op, AMRPoissonOp<Real>* coarserOp, const DeviceLayout* cfLay, int ratio = 2);

// Device layouts of one AMR hierarchy: AMR levels, the multigrid levels of level
// 0, and the coarsened-fine layouts used for restriction/prolongation between
// AMR levels.  Shared by the operator factories of every precision.
struct DeviceHierarchy {
  Context* ctx = nullptr;
  std::vector<DBL> grids;
  std::vector<Box> domains;
  std::vector<double> dxs;
  std::vector<int> ratios;                            // ratios[l]: level l-1 to l (2 or 4); ratios[0] = 1
  std::vector<long long> mgBottomCells;               // per level: MultiGrid::m_bottomCells of its m_mg
  std::vector<std::unique_ptr<DeviceLayout>> amrLay;  // per AMR level
  std::vector<std::unique_ptr<DeviceLayout>> cfLay;   // coarsen(grids[l], ratios[l]); index l >= 1
  std::vector<std::unique_ptr<DeviceLayout>> mgLay;   // MG levels of AMR level 0, depth >= 1
  // MG levels of AMR level l >= 1 (depth >= 1), made by defineLevelMG for solves with
  // l_base = l: depth 1 is cfLay[l] at ratio 2, the others are owned here
  std::vector<std::vector<const DeviceLayout*>> lmgLay;
  std::vector<std::vector<std::unique_ptr<DeviceLayout>>> lmgOwned;
  bool merged = true;
  void defineLevelMG(int l);
  // merge: store each level whose boxes tile a rectangle as one array; ratio: the
  // nlev-1 refinement ratios between consecutive levels (null: 2 everywhere)
  void define(Context* c, const std::vector<DBL>& g, const Box& domain0, double dx0, bool merge = true,
              const BCSpec* bc = nullptr, const int* ratio = nullptr);
  int nlev() const { return (int)grids.size(); }
  int mgDepth() const { return 1 + (int)mgLay.size(); }
};

// AMRPoissonOpFactory::define (AMRPoissonOp.cpp:1857-1920) with AMRnewOp
// (:1992-2070) and MGnewOp (:1923-1989).
template <class Real>
class AMRPoissonOpFactory {
 public:
  DeviceHierarchy* H = nullptr;
  double alpha = 0, beta = 1;
  BCSpec bc{};
  int ncomp = 1;  // components of the fields the operators act on
  void define(DeviceHierarchy* h, double a, double b, const BCSpec& bcs, int nc = 1) {
    H = h; alpha = a; beta = b; bc = bcs; ncomp = nc;
  }
  // operator for AMR level l; the operator of level l-1 must be passed for l > 0
  std::unique_ptr<AMRPoissonOp<Real>> AMRnewOp(int l, AMRPoissonOp<Real>* coarserOp);
  std::vector<std::unique_ptr<MGLevelOp<LevelData<Real>>>> MGnewOps();
  // MGnewOp at every depth >= 1 of AMR level l >= 1 (coarser dx of the AMR level as dxCrse)
  std::vector<std::unique_ptr<MGLevelOp<LevelData<Real>>>> MGnewOpsLevel(int l);
};

// ---------------------------------------------------------------------------
// AMRMultiGrid<T> (AMRMultiGrid.H:307-1419)
// ---------------------------------------------------------------------------
template <class Real>
class AMRMultiGrid {
 public:
  using T = LevelData<Real>;
  double m_eps = 1e-6, m_hang = 1e-15, m_normThresh = 1e-30;
  int m_imin = 5, m_iterMax = 20, m_iterMin = -1, m_exitStatus = -1;
  int m_pre = 2, m_post = 2, m_bottom = 2, m_numMG = 1;
  double m_initial_rnorm = 0;
  double m_bottomSolverEpsCushion = 1.0;
  std::vector<std::unique_ptr<AMRPoissonOp<Real>>> m_op;
  std::unique_ptr<MultiGrid<T>> m_mg0;                   // multigrid of level 0
  std::vector<std::unique_ptr<MultiGrid<T>>> m_mgL;      // multigrid of level l >= 1, made on first use
  AMRPoissonOpFactory<Real>* m_factory = nullptr;
  // the MultiGrid of AMR level l (m_mg[l] at its full depth), bound to the bottom solver
  MultiGrid<T>* mg(int l);
  std::vector<std::unique_ptr<T>> m_correction, m_residual, m_resC;
  std::vector<long long> m_mgBottomCells;  // per level: its MultiGrid's m_bottomCells (ratio-4 relax)
  LinearSolver<T>* m_bottomSolver = nullptr;
  std::vector<double> m_history;
  int m_iterations = 0;
  // Mixed precision: when set, solveNoInitResid runs this in place of
  // AMRVCycle(uberCorrection, uberResidual, l_max, l_max, l_base) followed by
  // phi += uberCorrection and uberCorrection = 0: an inner correction cycle in another
  // precision, given phi and uberResidual, that adds its correction c to phi as
  // phi + (0 + c) * 1 (those two steps' arithmetic with uberCorrection zero); and
  // m_innerBottomMetrics receives the bottom solver's convergence metrics.
  std::function<void(std::vector<T*>&, std::vector<T*>&, int, int)> m_innerCycleUpdate;
  std::function<void(double, double)> m_innerBottomMetrics;
  // While solveNoInitResid runs a V-cycle with the phi update folded in (one
  // coarser cycle per level, FP64 update): the phi of every level.  The cycle then
  // adds each level's correction to phi itself and leaves uberCorrection zero.
  std::vector<T*> m_foldPhi;

  void define(AMRPoissonOpFactory<Real>& factory, LinearSolver<T>* bottom, int numLevels);
  void setSolverParameters(int pre, int post, int bottom, int numMG, int iterMax, double eps, double hang,
                           double normThresh);
  void init(std::vector<T*>& phi, std::vector<T*>& rhs, int l_max, int l_base);
  void solveNoInitResid(std::vector<T*>& phi, std::vector<T*>& uberResidual, std::vector<T*>& uberCorrection,
                        std::vector<T*>& rhs, int l_max, int l_base, bool zeroPhi);
  void AMRVCycle(std::vector<T*>& uberCorrection, std::vector<T*>& uberResidual, int ilev, int l_max, int l_base,
                 bool fromZero = true);
  double computeAMRResidual(std::vector<T*>& resid, std::vector<T*>& phi, std::vector<T*>& rhs, int l_max,
                            int l_base, bool homogeneousBC, bool computeNorm);
  void computeAMRResidualLevel(std::vector<T*>& resid, std::vector<T*>& phi, std::vector<T*>& rhs, int l_max,
                               int l_base, int ilev, bool homogeneousBC);
  void computeAMROperator(std::vector<T*>& lphi, std::vector<T*>& phi, int l_max, int l_base, bool homogeneousBC);
  void relax(T& phi, T& R, int depth, int nRelax, bool fromZero = false);
};

}  // namespace amrmg
