// Operator/solver interfaces of the path, mirroring Chombo 3.2's
//   LinearOp<T>        lib/src/AMRElliptic/LinearSolver.H:27-149
//   LinearSolver<T>    LinearSolver.H:155-195
//   MGLevelOp<T>       lib/src/AMRElliptic/MultiGrid.H:88-287
//   AMRLevelOp<T>      lib/src/AMRElliptic/AMRMultiGrid.H:38-226
//   MultiGrid<T>       MultiGrid.H:330-696
//   AMRMultiGrid<T>    AMRMultiGrid.H:307-1419
//   BiCGStabSolver<T>  lib/src/AMRElliptic/BiCGStabSolver.H:199-527
// with T = a device-resident LevelData.  Names, argument meanings and the
// control flow (convergence tests, exit-status bits, restarts) are the
// reference's; the data behind T never leaves HBM.
#pragma once
#include <cmath>
#include <memory>
#include <vector>

#include "device.hpp"

namespace amrmg {

template <class T>
class LinearOp {
 public:
  virtual ~LinearOp() {}
  virtual void residual(T& lhs, T& phi, const T& rhs, bool homogeneous = false) = 0;
  virtual void preCond(T& cor, const T& residual) = 0;
  virtual void applyOp(T& lhs, T& phi, bool homogeneous = false) = 0;
  virtual void create(T& lhs, const T& rhs) = 0;
  virtual void assignLocal(T& lhs, const T& rhs) = 0;
  virtual double dotProduct(const T& a, const T& b) = 0;
  virtual void incr(T& lhs, const T& x, double scale) = 0;
  virtual void axby(T& lhs, const T& x, const T& y, double a, double b) = 0;
  virtual void scale(T& lhs, double s) = 0;
  virtual double norm(const T& x, int ord) = 0;
  virtual void setToZero(T& lhs) = 0;
  virtual double dx() const = 0;
};

template <class T>
class MGLevelOp : public LinearOp<T> {
 public:
  virtual void createCoarser(T& coarse, const T& fine, bool ghosted) = 0;
  virtual void relax(T& correction, const T& residual, int iterations) = 0;
  // setToZero(correction) followed by relax(correction, residual, iterations); an
  // operator may produce the same valid values without reading the zeros
  virtual void relaxFromZero(T& correction, const T& residual, int iterations) {
    this->setToZero(correction);
    relax(correction, residual, iterations);
  }
  virtual void restrictResidual(T& resCoarse, T& phiFine, const T& rhsFine) = 0;
  virtual void prolongIncrement(T& phiThisLevel, const T& correctCoarse) = 0;
};

template <class T>
class AMRLevelOp : public MGLevelOp<T> {
 public:
  virtual int refToCoarser() = 0;
  virtual void AMRResidual(T& residual, T& phiFine, T& phi, T& phiCoarse, const T& rhs, bool homogeneousBC,
                           AMRLevelOp<T>* finerOp) = 0;
  virtual void AMRResidualNC(T& residual, T& phiFine, T& phi, const T& rhs, bool homogeneousBC,
                             AMRLevelOp<T>* finerOp) = 0;
  virtual void AMRResidualNF(T& residual, T& phi, T& phiCoarse, const T& rhs, bool homogeneousBC) = 0;
  virtual void AMROperator(T& LofPhi, T& phiFine, T& phi, T& phiCoarse, bool homogeneousBC,
                           AMRLevelOp<T>* finerOp) = 0;
  virtual void AMROperatorNC(T& LofPhi, T& phiFine, T& phi, bool homogeneousBC, AMRLevelOp<T>* finerOp) = 0;
  virtual void AMROperatorNF(T& LofPhi, T& phi, T& phiCoarse, bool homogeneousBC) = 0;
  virtual void AMRRestrictS(T& resCoarse, const T& residual, T& correction, T& coarseCorrection, T& scratch) = 0;
  virtual void AMRProlongS(T& correction, const T& coarseCorrection, T& temp) = 0;
  virtual void AMRUpdateResidual(T& residual, T& correction, T& coarseCorrection) = 0;
  virtual void createCoarsened(T& lhs, const T& rhs, int refRat) = 0;
  virtual void assignCopier(T& lhs, const T& rhs) = 0;   // copier: coarsened-fine -> this level
  virtual void zeroCovered(T& lhs) = 0;                  // zero cells covered by the finer level
  virtual double localMaxNorm(const T& x) = 0;
};

template <class T>
class LinearSolver {
 public:
  virtual ~LinearSolver() {}
  virtual void define(LinearOp<T>* op, bool homogeneous) = 0;
  virtual void solve(T& phi, const T& rhs) = 0;
  virtual void setConvergenceMetrics(double metric, double tolerance) = 0;
};

// ---------------------------------------------------------------------------
// BiCGStabSolver<T>::solve (BiCGStabSolver.H:199-519)
// ---------------------------------------------------------------------------
template <class T>
class BiCGStabSolver : public LinearSolver<T> {
 public:
  bool m_homogeneous = false;
  LinearOp<T>* m_op = nullptr;
  int m_imax = 80, m_verbosity = 0;
  double m_eps = 1e-6, m_reps = 1e-12, m_convergenceMetric = -1.0, m_hang = 1e-8;
  int m_exitStatus = -1, m_iter = -1;
  double m_residual = -1e300, m_initial_residual = -1e300, m_small = 1e-30;
  int m_numRestarts = 5, m_normType = 2;
  long long m_total_iters = 0;

  void define(LinearOp<T>* op, bool homogeneous) override {
    m_homogeneous = homogeneous;
    m_op = op;
  }
  void setConvergenceMetrics(double metric, double tol) override {
    m_convergenceMetric = metric;
    m_eps = tol;
  }
  void solve(T& a_phi, const T& a_rhs) override {
    T r, r_tilde, e, p, p_tilde, s_tilde, t, v;
    m_op->create(r, a_rhs); m_op->create(r_tilde, a_rhs); m_op->create(e, a_phi); m_op->create(p, a_rhs);
    m_op->create(p_tilde, a_phi); m_op->create(s_tilde, a_phi); m_op->create(t, a_rhs); m_op->create(v, a_rhs);
    int recount = 0;
    m_op->setToZero(r);
    m_op->residual(r, a_phi, a_rhs, m_homogeneous);
    m_op->assignLocal(r_tilde, r);
    m_op->setToZero(e); m_op->setToZero(p_tilde); m_op->setToZero(s_tilde);
    m_iter = 0;
    double rho[4] = {0, 0, 0, 0}, norm[2];
    norm[0] = m_op->norm(r, m_normType);
    double initial_norm = norm[0];
    m_initial_residual = norm[0];
    m_residual = norm[0];
    norm[1] = norm[0];
    double alpha[2] = {0, 0}, beta[2] = {0, 0}, omega[2] = {0, 0};
    bool init = true;
    int restarts = 0;
    if (m_convergenceMetric > 0) initial_norm = m_convergenceMetric;
    while ((m_iter < m_imax && norm[0] > m_eps * norm[1]) && (norm[1] > 0)) {
      m_iter++;
      m_total_iters++;
      norm[1] = norm[0]; alpha[1] = alpha[0]; beta[1] = beta[0]; omega[1] = omega[0];
      rho[3] = rho[2]; rho[2] = rho[1];
      rho[1] = m_op->dotProduct(r_tilde, r);
      if (rho[1] == 0.0) {
        m_op->incr(a_phi, e, 1.0);
        m_residual = norm[0];
        m_exitStatus = 2;
        return;
      }
      if (init) {
        m_op->assignLocal(p, r);
        init = false;
      } else {
        beta[1] = (rho[1] / rho[2]) * (alpha[1] / omega[1]);
        m_op->scale(p, beta[1]);
        m_op->incr(p, v, -beta[1] * omega[1]);
        m_op->incr(p, r, 1.0);
      }
      m_op->preCond(p_tilde, p);
      m_op->setToZero(v);
      m_op->applyOp(v, p_tilde, true);
      double m = m_op->dotProduct(r_tilde, v);
      alpha[0] = rho[1] / m;
      if (std::fabs(m) > m_small * std::fabs(rho[1])) {
        m_op->incr(r, v, -alpha[0]);
        norm[0] = m_op->norm(r, m_normType);
        m_op->incr(e, p_tilde, alpha[0]);
      } else {
        m_op->setToZero(r);
        norm[0] = 0.0;
      }
      m_residual = norm[0];
      if (norm[0] > m_eps * initial_norm && norm[0] > m_reps * m_initial_residual) {
        m_op->preCond(s_tilde, r);
        m_op->setToZero(t);
        m_op->applyOp(t, s_tilde, true);
        omega[0] = m_op->dotProduct(t, r) / m_op->dotProduct(t, t);
        m_op->incr(e, s_tilde, omega[0]);
        m_op->incr(r, t, -omega[0]);
        norm[0] = m_op->norm(r, m_normType);
      }
      m_residual = norm[0];
      if (norm[0] <= m_eps * initial_norm || norm[0] <= m_reps * m_initial_residual) {
        m_exitStatus = 1;
        break;
      }
      if (omega[0] == 0.0 || norm[0] > (1 - m_hang) * norm[1]) {
        if (recount == 0) {
          recount = 1;
        } else {
          recount = 0;
          m_op->incr(a_phi, e, 1.0);
          if (restarts == m_numRestarts) {
            m_exitStatus = 3;
            return;
          }
          m_op->residual(r, a_phi, a_rhs, m_homogeneous);
          norm[0] = m_op->norm(r, m_normType);
          rho[1] = 0.0; rho[2] = 0.0; rho[3] = 0.0;
          alpha[0] = 0; beta[0] = 0; omega[0] = 0;
          m_op->assignLocal(r_tilde, r);
          m_op->setToZero(e);
          restarts++;
          init = true;
        }
      }
    }
    m_residual = norm[0];
    m_op->incr(a_phi, e, 1.0);
  }
};

// ---------------------------------------------------------------------------
// MultiGrid<T> (MultiGrid.H:330-696)
// ---------------------------------------------------------------------------
template <class T>
class MultiGrid {
 public:
  int m_depth = 0, m_defaultDepth = 0, m_pre = 3, m_post = 3, m_bottom = 0, m_cycle = 1, m_numMG = 1;
  long long m_bottomCells = 0;
  LinearSolver<T>* m_bottomSolver = nullptr;
  std::vector<MGLevelOp<T>*> m_op;           // m_op[0] is the AMR level operator (not owned)
  std::vector<std::unique_ptr<MGLevelOp<T>>> m_owned;
  std::vector<std::unique_ptr<T>> m_residual, m_correction;

  // ops: finest (AMR) op followed by MGnewOp(depth) results
  void define(MGLevelOp<T>* finest, std::vector<std::unique_ptr<MGLevelOp<T>>> coarser, long long domainCells,
              LinearSolver<T>* bottom) {
    m_op.clear();
    m_owned = std::move(coarser);
    m_op.push_back(finest);
    for (auto& o : m_owned) m_op.push_back(o.get());
    m_depth = (int)m_op.size();
    m_defaultDepth = m_depth;
    m_bottomCells = domainCells;
    for (int d = 1; d < m_depth; d++) m_bottomCells /= 8;
    m_residual.resize(m_depth);
    m_correction.resize(m_depth);
    setBottomSolver(bottom);
  }
  void setBottomSolver(LinearSolver<T>* s) {
    m_bottomSolver = s;
    if (s) s->define(m_op[m_depth - 1], true);
  }
  // MultiGrid::init (MultiGrid.H:534-548)
  void init(const T& e, const T& residual) {
    for (int i = 1; i < m_depth; i++) {
      m_residual[i].reset(new T());
      m_correction[i].reset(new T());
      m_op[i - 1]->createCoarser(*m_residual[i], i == 1 ? residual : *m_residual[i - 1], false);
      m_op[i - 1]->createCoarser(*m_correction[i], i == 1 ? e : *m_correction[i - 1], true);
    }
  }
  void oneCycle(T& e, const T& res) { cycle(0, e, res); }
  // oneCycle for a caller whose e is zero on entry (the AMR V-cycle's coarsest level)
  void oneCycleFromZero(T& e, const T& res) { cycle(0, e, res, true); }
  // MultiGrid::cycle (MultiGrid.H:630-696): V/W cycles (m_cycle > 0) and F cycles
  // (m_cycle < 0).  zero: correction is zero on entry (as the reference's setToZero
  // before each coarser cycle makes it), so its first relaxation need not read it.
  void cycle(int depth, T& correction, const T& residual, bool zero = false) {
    auto relax = [&](int n) {
      if (zero) m_op[depth]->relaxFromZero(correction, residual, n);
      else m_op[depth]->relax(correction, residual, n);
      zero = false;
    };
    if (depth == m_depth - 1) {
      if (m_bottomCells == 1) {
        relax(1);
      } else {
        relax(m_bottom);
        m_bottomSolver->solve(correction, residual);
      }
      return;
    }
    int cycles = m_cycle;
    if (cycles < 0) {
      // F cycle: coarse correction first, then -m_cycle V-cycles (the reference sets
      // m_cycle = 1 around them and restores it)
      cycles = -cycles;
      if (zero) {  // restrictResidual reads the correction
        m_op[depth]->setToZero(correction);
        zero = false;
      }
      m_op[depth]->restrictResidual(*m_residual[depth + 1], correction, residual);
      cycle(depth + 1, *m_correction[depth + 1], *m_residual[depth + 1], true);
      m_op[depth]->prolongIncrement(correction, *m_correction[depth + 1]);
      m_op[depth]->relax(correction, residual, m_pre);
      for (int img = 0; img < cycles; img++) {
        m_op[depth]->restrictResidual(*m_residual[depth + 1], correction, residual);
        m_cycle = 1;
        cycle(depth + 1, *m_correction[depth + 1], *m_residual[depth + 1], true);
        m_cycle = -cycles;
        m_op[depth]->prolongIncrement(correction, *m_correction[depth + 1]);
      }
      m_op[depth]->relax(correction, residual, m_post);
      return;
    }
    relax(m_pre);
    m_op[depth]->restrictResidual(*m_residual[depth + 1], correction, residual);
    // setToZero(m_correction[depth + 1]) is folded into the first coarser cycle's
    // first relaxation (relaxFromZero)
    for (int img = 0; img < cycles; img++)
      cycle(depth + 1, *m_correction[depth + 1], *m_residual[depth + 1], img == 0);
    m_op[depth]->prolongIncrement(correction, *m_correction[depth + 1]);
    m_op[depth]->relax(correction, residual, m_post);
  }
};

}  // namespace amrmg
