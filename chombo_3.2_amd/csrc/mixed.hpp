// Mixed precision: FP32 inner V-cycles under the FP64 outer iteration.
//
// The outer loop stays AMRMultiGrid<double>::solveNoInitResid (AMRMultiGrid.H:935-1124):
// FP64 composite residual, FP64 max-norm convergence tests and the FP64 update
// phi += correction.  Each outer iteration's correction cycle runs in FP32: the
// FP64 residual is rounded into FP32 fields, one AMRVCycle of AMRMultiGrid<float>
// (relax, residual, restrict, prolong, CF interpolation, reflux all in FP32)
// produces the correction, which is widened and handed back.  Because the
// residual is recomputed in FP64 every outer iteration, FP32 rounding only limits
// how well one cycle solves the correction equation, not the attainable accuracy
// (iterative refinement).  The bottom of the FP32 multigrid is solved in FP64
// (PromotedBottomSolver).  The reference has no mixed-precision mode; this is an
// addition beside its FP64 path.
#pragma once
#include "amrpoissonop.hpp"
#include "batch.hpp"

namespace amrmg {

// LinearSolver<FP32 level> that solves in FP64: widens the correction and
// residual, runs BiCGStab with the FP64 operator of the same level, rounds back.
class PromotedBottomSolver : public LinearSolver<LevelData<float>> {
 public:
  LinearOp<LevelData<double>>* m_op64 = nullptr;
  const DeviceLayout* m_lay = nullptr;
  BiCGStabSolver<LevelData<double>> m_bicg;
  LevelData<double> m_e, m_r;
  bool m_homogeneous = true;

  void bind(LinearOp<LevelData<double>>* op64, const DeviceLayout* lay) {
    m_op64 = op64;
    m_lay = lay;
    m_e.define(lay);
    m_r.define(lay);
    m_bicg.define(op64, m_homogeneous);
  }
  void define(LinearOp<LevelData<float>>* op, bool homogeneous) override {
    (void)op;
    m_homogeneous = homogeneous;
    if (m_op64) m_bicg.define(m_op64, homogeneous);
  }
  void setConvergenceMetrics(double metric, double tol) override { m_bicg.setConvergenceMetrics(metric, tol); }
  void solve(LevelData<float>& phi, const LevelData<float>& rhs) override {
    if (!m_op64) throw std::runtime_error("PromotedBottomSolver: not bound to an FP64 operator");
    if (phi.lay != m_lay || rhs.lay != m_lay) throw std::runtime_error("PromotedBottomSolver: layout mismatch");
    ops::convert_valid_f2d(*m_lay, m_e.d, phi.d);
    ops::convert_valid_f2d(*m_lay, m_r.d, rhs.d);
    m_bicg.solve(m_e, m_r);
    ops::convert_valid_d2f(*m_lay, phi.d, m_e.d);
  }
};

// The FP32 side of a mixed-precision solve over an existing FP64 hierarchy.
struct MixedInner {
  AMRPoissonOpFactory<float> fac;
  AMRMultiGrid<float> amg;
  PromotedBottomSolver bottom;                       // host-driven FP64 BiCGStab
  std::unique_ptr<DeviceBottomSolver<float>> devBottom;  // the same on the device, when the level allows
  LinearSolver<LevelData<float>>* active = nullptr;
  std::vector<std::unique_ptr<LevelData<float>>> res, corr;

  // amg64 must be defined and initialised on the same DeviceHierarchy
  void define(DeviceHierarchy* H, double alpha, double beta, const BCSpec& bc, AMRMultiGrid<double>& amg64) {
    fac.define(H, alpha, beta, bc);
    MGLevelOp<LevelData<double>>* bottomOp64 = amg64.m_mg0->m_op.back();
    bottom.bind(bottomOp64, H->mgDepth() > 1 ? H->mgLay.back().get() : H->amrLay[0].get());
    amg.define(fac, &bottom, H->nlev());
    active = &bottom;
    devBottom.reset(new DeviceBottomSolver<float>(H->ctx));
    devBottom->define(amg.m_mg0->m_op.back(), true);
    if (devBottom->onDevice()) {
      active = devBottom.get();
      amg.m_bottomSolver = active;
      amg.m_mg0->setBottomSolver(active);
    }
    for (int l = 0; l < H->nlev(); l++) {
      res.emplace_back(new LevelData<float>(H->amrLay[l].get()));
      corr.emplace_back(new LevelData<float>(H->amrLay[l].get()));
    }
    std::vector<LevelData<float>*> c, r;
    for (int l = 0; l < H->nlev(); l++) { c.push_back(corr[l].get()); r.push_back(res[l].get()); }
    amg.init(c, r, H->nlev() - 1, 0);
    amg64.m_innerCycleUpdate = [this](std::vector<LevelData<double>*>& phi, std::vector<LevelData<double>*>& ur,
                                      int l_max, int l_base) { cycle(phi, ur, l_max, l_base); };
    amg64.m_innerBottomMetrics = [this](double metric, double tol) { active->setConvergenceMetrics(metric, tol); };
  }
  // bind the bottom again after the operators' coefficients change: the device
  // solver when it applies to the FP32 bottom level, the promoted BiCGStab otherwise
  void rebindBottom() {
    devBottom->define(amg.m_mg0->m_op.back(), true);
    active = devBottom->onDevice() ? (LinearSolver<LevelData<float>>*)devBottom.get() : &bottom;
    amg.m_bottomSolver = active;
    amg.m_mg0->setBottomSolver(active);
  }
  long long bottomIterations() {
    return active == &bottom ? bottom.m_bicg.m_total_iters : devBottom->totalIterations();
  }
  void setSolverParameters(int pre, int post, int nbottom, int numMG, int iterMax, double eps, double hang,
                           double normThresh) {
    amg.setSolverParameters(pre, post, nbottom, numMG, iterMax, eps, hang, normThresh);
  }
  // phi += the FP32 correction cycle's result for uberResidual, with the arithmetic of
  // uberCorrection (zero) += widened correction, phi += uberCorrection * 1
  void cycle(std::vector<LevelData<double>*>& phi, std::vector<LevelData<double>*>& ur, int l_max, int l_base) {
    std::vector<LevelData<float>*> c, r;
    for (size_t l = 0; l < corr.size(); l++) { c.push_back(corr[l].get()); r.push_back(res[l].get()); }
    for (int l = l_base; l <= l_max; l++) {
      ops::convert_valid_d2f(*ur[l]->lay, r[l]->d, ur[l]->d);
      ops::setval<float>(*c[l]->lay, c[l]->d, 0.0);
    }
    amg.AMRVCycle(c, r, l_max, l_max, l_base);
    for (int l = l_base; l <= l_max; l++) ops::update_f2d(*phi[l]->lay, phi[l]->d, c[l]->d);
  }
};

}  // namespace amrmg
