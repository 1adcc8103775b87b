// AMRPoissonOp, its factory and AMRMultiGrid on device-resident data.
// Method-by-method mirror of lib/src/AMRElliptic/AMRPoissonOp.cpp and
// AMRMultiGrid.H; each method cites the reference lines it follows.
#include "amrpoissonop.hpp"

#include <cmath>
#include <stdexcept>

namespace amrmg {

template <class T>
static T* upload_vec(Context* c, const std::vector<T>& v) {
  T* d = nullptr;
  dev_alloc(c, (void**)&d, sizeof(T) * v.size(), "hipMalloc table");
  dev_h2d(c, d, v.data(), sizeof(T) * v.size(), "copy table");
  return d;
}

template <class Real>
AMRPoissonOp<Real>::~AMRPoissonOp() {
  if (m_ctx && !m_ctx->dry) (void)hipStreamSynchronize(m_ctx->stream);
  m_ghost.release();
  m_resCopier.release();
  m_revCopier.release();
  if (m_ctx) {
    dev_free(m_ctx, d_cfc);
    dev_free(m_ctx, d_cfg);
    dev_free(m_ctx, d_cff);
    dev_free(m_ctx, d_deriv);
    dev_free(m_ctx, d_rcell);
    dev_free(m_ctx, d_rc);
    dev_free(m_ctx, d_rf);
    dev_free(m_ctx, d_vcRcFace);
    dev_free(m_ctx, d_vcRfFace);
    if (m_lay && d_boxTiles != m_lay->d_tiles) dev_free(m_ctx, d_boxTiles);
  }
}

// Every operand's layout is checked against the layout the operation assumes
// before anything is launched: slab sizes and box tables come from the field.
static void need(const void* have, const void* want, const char* what) {
  if (have != want || have == nullptr) throw std::runtime_error(std::string("level data layout mismatch in ") + what);
}
// Fields of several components: every kernel of the operator updates each component
// from that component alone (the reference's Fortran loops "do n = 0, ncomp-1" around
// per-cell updates, e.g. AMRPoissonOpF.ChF:158), so an operation on nc components is
// the operation on each component's one-component view in turn.  The reductions
// (dotProduct, norm, localMaxNorm) and the one-launch relaxation handle all
// components themselves.
template <class Real>
struct CompView {
  LevelData<Real> v;
  CompView(const LevelData<Real>& x, int c) { v.alias(x, c); }
};

// ---- ghost refresh: homogeneousCFInterp (AMRPoissonOp.cpp:1598), exchange
// (exchangeNoOverlap with the trimmed Copier), BC functor (BCFunc.cpp:58) ----
template <class Real>
void AMRPoissonOp<Real>::fillGhosts(T& phi, bool doCF, bool homogBC) {
  if (phi.nc > 1) {
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> phi_(phi, c);
      fillGhosts(phi_.v, doCF, homogBC);
    }
    return;
  }
  need(phi.lay, m_lay, "fillGhosts");
  if (!homogBC && m_ghost.d_bcv && phi.comp >= m_ncomp)
    throw std::runtime_error("boundary values were taken for fewer components than the field has");
  ops::fill_ghosts<Real>(*m_lay, m_ghost, phi.d, doCF ? 1 : 0, homogBC ? 1 : 0, m_dx, m_dxCrse, phi.comp);
}
template <class Real>
void AMRPoissonOp<Real>::homogeneousCFInterp(T& phi) {
  // CF patches only: exchange and BC rules also run, which only rewrite ghosts
  // that every later use refreshes again
  fillGhosts(phi, true, true);
}

// residual (AMRPoissonOp.cpp:224-236) / residualI (:260-297)
template <class Real>
void AMRPoissonOp<Real>::residual(T& lhs, T& phi, const T& rhs, bool homog) {
  if (phi.nc > 1) {
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> lhs_(lhs, c), phi_(phi, c), rhs_(rhs, c);
      residual(lhs_.v, phi_.v, rhs_.v, homog);
    }
    return;
  }
  need(lhs.lay, m_lay, "residual"); need(rhs.lay, m_lay, "residual");
  fillGhosts(phi, homog, homog);
  if (m_vc) vc::residual<Real>(*m_lay, *m_vc, lhs.d, phi.d, rhs.d, m_dx, m_alpha, m_beta);
  else ops::residual<Real>(*m_lay, lhs.d, phi.d, rhs.d, m_dx, m_alpha, m_beta);
}
template <class Real>
void AMRPoissonOp<Real>::residualI(T& lhs, T& phi, const T& rhs, bool homog) {
  if (phi.nc > 1) {
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> lhs_(lhs, c), phi_(phi, c), rhs_(rhs, c);
      residualI(lhs_.v, phi_.v, rhs_.v, homog);
    }
    return;
  }
  need(lhs.lay, m_lay, "residualI"); need(rhs.lay, m_lay, "residualI");
  fillGhosts(phi, false, homog);
  if (m_vc) vc::residual<Real>(*m_lay, *m_vc, lhs.d, phi.d, rhs.d, m_dx, m_alpha, m_beta);
  else ops::residual<Real>(*m_lay, lhs.d, phi.d, rhs.d, m_dx, m_alpha, m_beta);
}
// applyOp (:332-340) / applyOpI (:343-385)
template <class Real>
void AMRPoissonOp<Real>::applyOp(T& lhs, T& phi, bool homog) {
  if (phi.nc > 1) {
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> lhs_(lhs, c), phi_(phi, c);
      applyOp(lhs_.v, phi_.v, homog);
    }
    return;
  }
  need(lhs.lay, m_lay, "applyOp");
  fillGhosts(phi, true, homog);
  if (m_vc) vc::apply<Real>(*m_lay, *m_vc, lhs.d, phi.d, m_dx, m_alpha, m_beta);
  else ops::apply<Real>(*m_lay, lhs.d, phi.d, m_dx, m_alpha, m_beta);
}
template <class Real>
void AMRPoissonOp<Real>::applyOpI(T& lhs, T& phi, bool homog) {
  if (phi.nc > 1) {
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> lhs_(lhs, c), phi_(phi, c);
      applyOpI(lhs_.v, phi_.v, homog);
    }
    return;
  }
  need(lhs.lay, m_lay, "applyOpI");
  fillGhosts(phi, false, homog);
  if (m_vc) vc::apply<Real>(*m_lay, *m_vc, lhs.d, phi.d, m_dx, m_alpha, m_beta);
  else ops::apply<Real>(*m_lay, lhs.d, phi.d, m_dx, m_alpha, m_beta);
}
// preCond (:304-329)
template <class Real>
void AMRPoissonOp<Real>::preCond(T& phi, const T& rhs) {
  need(phi.lay, m_lay, "preCond"); need(rhs.lay, m_lay, "preCond");
  if (phi.nc > 1) {  // the scaling per component, the relaxation on the whole field
    if (m_vc) throw std::runtime_error("preCond: variable coefficients take one component");
    const double mult = 1.0 / (m_alpha - 2.0 * 3 * m_beta / (m_dx * m_dx));
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> phi_(phi, c), rhs_(rhs, c);
      ops::precond_init<Real>(*m_lay, phi_.v.d, rhs_.v.d, mult);
    }
    relax(phi, rhs, 2);
    return;
  }
  if (m_vc) {  // VCAMRPoissonOp2::preCond (VCAMRPoissonOp2.cpp:93-127): phi = rhs * lambda, then relax
    vc::precond_init<Real>(*m_lay, *m_vc, phi.d, rhs.d);
    relax(phi, rhs, 2);
    return;
  }
  double mult = 1.0 / (m_alpha - 2.0 * 3 * m_beta / (m_dx * m_dx));
  ops::precond_init<Real>(*m_lay, phi.d, rhs.d, mult);
  relax(phi, rhs, 2);
}
// LevelDataOps on the operands' own layout (LevelDataOps.H:93-294)
template <class Real>
void AMRPoissonOp<Real>::assignLocal(T& lhs, const T& rhs) {
  if (lhs.nc > 1) {
    for (int c = 0; c < lhs.nc; c++) {
      CompView<Real> lhs_(lhs, c), rhs_(rhs, c);
      assignLocal(lhs_.v, rhs_.v);
    }
    return;
  }
  need(rhs.lay, lhs.lay, "assignLocal");
  ops::copy_all<Real>(*lhs.lay, lhs.d, rhs.d);
}
template <class Real>
double AMRPoissonOp<Real>::dotProduct(const T& a, const T& b) {
  need(b.lay, a.lay, "dotProduct");
  if (b.nc != a.nc) throw std::runtime_error("dotProduct: component counts differ");
  return ops::dot_ordered<Real>(*a.lay, a.d, b.d, 1, a.nc, a.cs);
}
template <class Real>
void AMRPoissonOp<Real>::incr(T& lhs, const T& x, double s) {
  if (lhs.nc > 1) {
    for (int c = 0; c < lhs.nc; c++) {
      CompView<Real> lhs_(lhs, c), x_(x, c);
      incr(lhs_.v, x_.v, s);
    }
    return;
  }
  need(x.lay, lhs.lay, "incr");
  ops::incr<Real>(*lhs.lay, lhs.d, x.d, s);
}
template <class Real>
void AMRPoissonOp<Real>::incrAssign(T& lhs, T& x) {
  if (lhs.nc > 1) {
    for (int c = 0; c < lhs.nc; c++) {
      CompView<Real> lhs_(lhs, c), x_(x, c);
      incrAssign(lhs_.v, x_.v);
    }
    return;
  }
  need(x.lay, lhs.lay, "incrAssign");
  ops::incr_assign<Real>(*lhs.lay, lhs.d, x.d);
}
template <class Real>
void AMRPoissonOp<Real>::incrFromZero(T& lhs, const T& x) {
  if (lhs.nc > 1) {
    for (int c = 0; c < lhs.nc; c++) {
      CompView<Real> lhs_(lhs, c), x_(x, c);
      incrFromZero(lhs_.v, x_.v);
    }
    return;
  }
  need(x.lay, lhs.lay, "incrFromZero");
  ops::incr_from_zero<Real>(*lhs.lay, lhs.d, x.d);
}
template <class Real>
void AMRPoissonOp<Real>::incrAssignUpdate(T& m, T& d, T& phi) {
  if (m.nc > 1) {
    for (int c = 0; c < m.nc; c++) {
      CompView<Real> m_(m, c), d_(d, c), phi_(phi, c);
      incrAssignUpdate(m_.v, d_.v, phi_.v);
    }
    return;
  }
  need(d.lay, m.lay, "incrAssignUpdate");
  need(phi.lay, m.lay, "incrAssignUpdate");
  ops::incr_assign_update<Real>(*m.lay, m.d, d.d, phi.d);
}
template <class Real>
void AMRPoissonOp<Real>::incrPhiFromZero(T& phi, const T& x) {
  if (phi.nc > 1) {
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> phi_(phi, c), x_(x, c);
      incrPhiFromZero(phi_.v, x_.v);
    }
    return;
  }
  need(x.lay, phi.lay, "incrPhiFromZero");
  ops::incr_phi_from_zero<Real>(*phi.lay, phi.d, x.d);
}
template <class Real>
void AMRPoissonOp<Real>::axby(T& lhs, const T& x, const T& y, double a, double b) {
  if (lhs.nc > 1) {
    for (int c = 0; c < lhs.nc; c++) {
      CompView<Real> lhs_(lhs, c), x_(x, c), y_(y, c);
      axby(lhs_.v, x_.v, y_.v, a, b);
    }
    return;
  }
  need(x.lay, lhs.lay, "axby"); need(y.lay, lhs.lay, "axby");
  ops::axby<Real>(*lhs.lay, lhs.d, x.d, y.d, a, b);
}
template <class Real>
void AMRPoissonOp<Real>::scale(T& lhs, double s) {
  if (lhs.nc > 1) {
    for (int c = 0; c < lhs.nc; c++) {
      CompView<Real> lhs_(lhs, c);
      scale(lhs_.v, s);
    }
    return;
  }
  need(lhs.lay, lhs.lay, "scale");
  ops::scale<Real>(*lhs.lay, lhs.d, s);
}
// norm (:558-562 -> BoxLayoutData.cpp:23-92)
template <class Real>
double AMRPoissonOp<Real>::norm(const T& x, int ord) {
  need(x.lay, x.lay, "norm");
  if (ord == 0) return ops::maxnorm<Real>(*x.lay, x.d, 2, x.nc, x.cs);
  if (ord == 1) return ops::norm1_ordered<Real>(*x.lay, x.d, 3, x.nc, x.cs);
  if (ord == 2) return ops::norm2_ordered<Real>(*x.lay, x.d, 3, x.nc, x.cs);
  throw std::runtime_error("AMRPoissonOp::norm: order must be 0, 1 or 2");
}
template <class Real>
void AMRPoissonOp<Real>::setToZero(T& lhs) {
  if (lhs.nc > 1) {
    for (int c = 0; c < lhs.nc; c++) {
      CompView<Real> lhs_(lhs, c);
      setToZero(lhs_.v);
    }
    return;
  }
  need(lhs.lay, lhs.lay, "setToZero");
  ops::setval<Real>(*lhs.lay, lhs.d, 0.0);
}
template <class Real>
double AMRPoissonOp<Real>::localMaxNorm(const T& x) {
  need(x.lay, m_lay, "localMaxNorm");
  return ops::maxnorm<Real>(*m_lay, x.d, 4, x.nc, x.cs);
}
template <class Real>
void AMRPoissonOp<Real>::localMaxNormAsync(const T& x, int slot, bool accumulate) {
  need(x.lay, m_lay, "localMaxNorm");
  for (int c = 0; c < x.nc; c++) ops::maxnorm_async<Real>(*m_lay, x.d + c * x.cs, slot, (accumulate || c > 0) ? 1 : 0);
}

// ---- variable coefficients ----
template <class Real>
void AMRPoissonOp<Real>::setVC(std::shared_ptr<VCCoefs<Real>> c) {
  if (c && c->lay != m_lay) throw std::runtime_error("setVC: coefficients belong to another level");
  m_vc = std::move(c);
  if (m_vc) resetLambda();
}
// face indices of the reflux contributions in this level's and the finer level's
// face arrays (the arrays' layout depends only on the level layouts)
template <class Real>
void AMRPoissonOp<Real>::setupVCReflux(const VCCoefs<Real>& fine) {
  if (!m_vc) throw std::runtime_error("setupVCReflux: no coefficients on this level");
  auto idx = [](const VCCoefs<Real>& C, int box, int d, const int* iv) {
    const DBox& f = C.h_fbox[d][box];
    return f.off + (iv[0] - f.lo[0]) + (long long)(iv[1] - f.lo[1]) * f.sy + (long long)(iv[2] - f.lo[2]) * f.sz;
  };
  auto check = [](const VCCoefs<Real>& C, int box, int d, const int* iv) {
    if (box < 0 || box >= C.lay->nbox) throw std::runtime_error("reflux face: bad box");
    const DBox& f = C.h_fbox[d][box];
    for (int q = 0; q < 3; q++)
      if (iv[q] < f.lo[q] || iv[q] > f.hi[q] + (q == d ? 1 : 0))
        throw std::runtime_error("reflux face outside its box's faces");
  };
  std::vector<VCFaceIdx> cf(m_rcFace.size()), ff(m_rfFace.size());
  for (size_t r = 0; r < m_rcFace.size(); r++) {
    const RefluxFace& x = m_rcFace[r];
    check(*m_vc, x.box, x.dir, x.iv[0]);
    cf[r].dir = x.dir;
    for (int t = 0; t < 4; t++) cf[r].x[t] = idx(*m_vc, x.box, x.dir, x.iv[0]);
  }
  for (size_t r = 0; r < m_rfFace.size(); r++) {
    const RefluxFace& x = m_rfFace[r];
    ff[r].dir = x.dir;
    for (int t = 0; t < 4; t++) {
      check(fine, x.box, x.dir, x.iv[t]);
      ff[r].x[t] = idx(fine, x.box, x.dir, x.iv[t]);
    }
  }
  dev_free(m_ctx, d_vcRcFace);
  dev_free(m_ctx, d_vcRfFace);
  d_vcRcFace = cf.empty() ? nullptr : upload_vec(m_ctx, cf);
  d_vcRfFace = ff.empty() ? nullptr : upload_vec(m_ctx, ff);
  m_vcRefluxFine = &fine;
}
// computeLambda (VCAMRPoissonOp2.cpp:283-318)
template <class Real>
void AMRPoissonOp<Real>::resetLambda() {
  if (m_vc) vc::lambda<Real>(*m_vc, m_dx, m_alpha, m_beta);
}

// ---- multigrid ----
template <class Real>
void AMRPoissonOp<Real>::createCoarser(T& coarse, const T& fine, bool) {
  if (!m_mgCoarse) throw std::runtime_error("createCoarser: level not coarsenable");
  need(fine.lay, m_lay, "createCoarser");
  coarse.define(m_mgCoarse, fine.nc);
}
// relax (:600-632) with levelGSRB (:1216-1281)
template <class Real>
void AMRPoissonOp<Real>::levelGSRB(T& phi, const T& rhs) {
  need(rhs.lay, m_lay, "levelGSRB");
  if (phi.nc > 1) {
    if (rhs.nc != phi.nc) throw std::runtime_error("levelGSRB: component counts differ");
    if (m_fused && m_ctx->fused_relax && !m_vc) {
      // every component relaxed into the second buffer, then the buffers swapped
      need(phi.lay, m_lay, "levelGSRB");
      if (!m_relaxTmpN.isDefined() || m_relaxTmpN.nc != phi.nc) m_relaxTmpN.define(m_lay, phi.nc);
      if (m_relaxTmpN.cs != phi.cs || rhs.cs != phi.cs) throw std::runtime_error("levelGSRB: component strides differ");
      for (int c = 0; c < phi.nc; c++)
        ops::gsrb_fused<Real>(*m_lay, m_faces, m_relaxTmpN.d + c * phi.cs, phi.d + c * phi.cs, rhs.d + c * phi.cs, m_dx,
                              m_alpha, m_beta, m_relaxShape, m_relaxKc);
      std::swap(phi.d, m_relaxTmpN.d);
      return;
    }
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> phi_(phi, c), rhs_(rhs, c);
      levelGSRB(phi_.v, rhs_.v);
    }
    return;
  }
  if (!phi.own && m_fused && m_ctx->fused_relax)
    throw std::runtime_error("levelGSRB: the one-launch relaxation swaps buffers and takes whole fields");
  if (m_vc) {  // VCAMRPoissonOp2::levelGSRB (VCAMRPoissonOp2.cpp:547-626)
    need(phi.lay, m_lay, "levelGSRB");
    if (m_fused && m_ctx->fused_relax) {
      if (!m_relaxTmp.isDefined()) m_relaxTmp.define(m_lay);
      vc::gsrb_fused<Real>(*m_lay, m_faces, *m_vc, m_relaxTmp.d, phi.d, rhs.d, m_dx, m_alpha, m_beta, m_vcShape);
      std::swap(phi.d, m_relaxTmp.d);
      return;
    }
    for (int pass = 0; pass <= 1; pass++) {
      fillGhosts(phi, true, true);
      vc::gsrb<Real>(*m_lay, *m_vc, phi.d, rhs.d, pass, m_dx, m_alpha, m_beta);
    }
    return;
  }
  if (m_fused && m_ctx->fused_relax) {
    need(phi.lay, m_lay, "levelGSRB");
    if (!m_relaxTmp.isDefined()) m_relaxTmp.define(m_lay);
    ops::gsrb_fused<Real>(*m_lay, m_faces, m_relaxTmp.d, phi.d, rhs.d, m_dx, m_alpha, m_beta, m_relaxShape,
                          m_relaxKc);
    std::swap(phi.d, m_relaxTmp.d);
    return;
  }
  for (int pass = 0; pass <= 1; pass++) {
    fillGhosts(phi, true, true);
    ops::gsrb<Real>(*m_lay, phi.d, rhs.d, pass, m_dx, m_alpha, m_beta);
  }
}
template <class Real>
bool AMRPoissonOp<Real>::tuneRelax(const T& phi, const T& rhs, double* times) {
  if (!m_fused || m_vc) return false;
  if (phi.nc > 1) {  // timed on the first component
    CompView<Real> phi_(phi, 0), rhs_(rhs, 0);
    return tuneRelax(phi_.v, rhs_.v, times);
  }
  need(phi.lay, m_lay, "tuneRelax"); need(rhs.lay, m_lay, "tuneRelax");
  if (!m_relaxTmp.isDefined()) m_relaxTmp.define(m_lay);
  const std::pair<int, int> pick =
      ops::tune_gsrb_fused<Real>(*m_lay, m_faces, m_relaxTmp.d, phi.d, rhs.d, m_dx, m_alpha, m_beta, times);
  m_relaxShape = pick.first;
  m_relaxKc = pick.second;
  return true;
}
// setToZero + relax: on merged levels the first iteration reads no old values
// (they are all zero) and writes every valid cell, so the zeroing itself is
// dropped; ghost cells are refilled before any later read, as after any relax.
template <class Real>
void AMRPoissonOp<Real>::relaxFromZero(T& e, const T& r, int n) {
  if (!(m_fused && m_ctx->fused_relax) || m_vc || m_relaxMode != 1 || n <= 0 || e.nc > 1) {
    setToZero(e);
    relax(e, r, n);
    return;
  }
  need(e.lay, m_lay, "relaxFromZero"); need(r.lay, m_lay, "relaxFromZero");
  if (!m_relaxTmp.isDefined()) m_relaxTmp.define(m_lay);
  ops::gsrb_fused<Real>(*m_lay, m_faces, m_relaxTmp.d, e.d, r.d, m_dx, m_alpha, m_beta, m_relaxShape, m_relaxKc,
                        true);
  std::swap(e.d, m_relaxTmp.d);
  relax(e, r, n - 1);
}
template <class Real>
void AMRPoissonOp<Real>::relax(T& e, const T& r, int n) {
  if (e.nc > 1 && m_relaxMode != 1) {  // levelGSRB takes the components itself
    for (int c = 0; c < e.nc; c++) {
      CompView<Real> e_(e, c), r_(r, c);
      relax(e_.v, r_.v, n);
    }
    return;
  }
  for (int i = 0; i < n; i++) {
    if (m_relaxMode == 4) levelJacobi(e, r);
    else if (m_relaxMode == 0) looseGSRB(e, r);
    else if (m_relaxMode == 2) overlapGSRB(e, r);
    else if (m_relaxMode == 3) levelGSRBLazy(e, r);
    else if (m_relaxMode == 5) levelMultiColor(e, r);
    else levelGSRB(e, r);
  }
}
template <class Real>
const Tile* AMRPoissonOp<Real>::boxTiles() {
  if (!m_lay->merged) return m_lay->d_tiles;
  if (!d_boxTiles) {
    std::vector<Tile> tiles;
    for (int b = 0; b < m_lay->nbox; b++) {
      const DBox& x = m_lay->h_boxes[b];
      for (int k = x.lo[2]; k <= x.hi[2]; k += m_lay->TZ)
        for (int j = x.lo[1]; j <= x.hi[1]; j += m_lay->TY)
          for (int i = x.lo[0]; i <= x.hi[0]; i += m_lay->TX) tiles.push_back(Tile{b, i, j, k});
    }
    if (tiles.size() > (size_t)INT32_MAX) throw std::runtime_error("too many box tiles");
    d_boxTiles = upload_vec(m_ctx, tiles);
    n_boxTiles = (int)tiles.size();
  }
  return d_boxTiles;
}
// The neighbour values a box saw at the last ghost refresh: on per-box storage
// its own ghost cells (phi itself), on merged storage a copy of the level.
template <class Real>
const Real* AMRPoissonOp<Real>::relaxSnapshot(T& phi) {
  if (!m_lay->merged) return phi.d;
  if (!m_relaxTmp.isDefined()) m_relaxTmp.define(m_lay);
  ops::copy_all<Real>(*m_lay, m_relaxTmp.d, phi.d);
  return m_relaxTmp.d;
}
// looseGSRB (:1333-1413): one ghost refresh, then both colours box by box with
// the ghost values of that refresh
template <class Real>
void AMRPoissonOp<Real>::looseGSRB(T& phi, const T& rhs) {
  if (m_vc) throw std::runtime_error("looseGSRB is not implemented for variable coefficients");
  need(phi.lay, m_lay, "looseGSRB"); need(rhs.lay, m_lay, "looseGSRB");
  fillGhosts(phi, true, true);
  const Tile* tiles = boxTiles();
  const int nt = m_lay->merged ? n_boxTiles : m_lay->ntiles;
  const Real* snap = relaxSnapshot(phi);
  for (int pass = 0; pass <= 1; pass++)
    ops::gsrb_region<Real>(*m_lay, tiles, nt, phi.d, snap, rhs.d, pass, 0, 0, 0, false, m_dx, m_alpha, m_beta);
}
// overlapGSRB (:1416-1521) without MPI (exchangeBegin copies every ghost at once,
// LevelDataI.H:533-538): looseGSRB unless alpha == 1, beta == 0 and several boxes;
// a ghost refresh; GSRBLAPLACIAN on every box's interior, then on its six face
// layers in the order x lo, x hi, y lo, ..., both colours each.  The boxes are
// independent (each reads the others only through the refresh's copies), so
// one launch covers a layer of every box.
template <class Real>
void AMRPoissonOp<Real>::overlapGSRB(T& phi, const T& rhs) {
  if (m_vc) throw std::runtime_error("overlapGSRB is not implemented for variable coefficients");
  need(phi.lay, m_lay, "overlapGSRB"); need(rhs.lay, m_lay, "overlapGSRB");
  if (m_lay->nbox == 1 || m_alpha != 1 || m_beta != 0) looseGSRB(phi, rhs);
  fillGhosts(phi, true, true);
  const Tile* tiles = boxTiles();
  const int nt = m_lay->merged ? n_boxTiles : m_lay->ntiles;
  const Real* snap = relaxSnapshot(phi);
  for (int pass = 0; pass <= 1; pass++)
    ops::gsrb_region<Real>(*m_lay, tiles, nt, phi.d, snap, rhs.d, pass, 1, 0, 0, true, m_dx, m_alpha, m_beta);
  for (int dir = 0; dir < 3; dir++)
    for (int side = 0; side < 2; side++)
      for (int pass = 0; pass <= 1; pass++)
        ops::gsrb_region<Real>(*m_lay, tiles, nt, phi.d, snap, rhs.d, pass, 2, dir, side, true, m_dx, m_alpha,
                               m_beta);
}
// levelGSRBLazy (:1524-1578): the eight parity colours in nextColorLoc's order
// (x fastest); for each, the homogeneous operator on the whole level, then
// GSRBLAZY on the cells of that colour
template <class Real>
void AMRPoissonOp<Real>::levelGSRBLazy(T& phi, const T& rhs) {
  if (m_vc) throw std::runtime_error("levelGSRBLazy is not implemented for variable coefficients");
  need(phi.lay, m_lay, "levelGSRBLazy"); need(rhs.lay, m_lay, "levelGSRBLazy");
  if (!m_jacobiTmp.isDefined()) m_jacobiTmp.define(m_lay);
  for (int c = 0; c < 8; c++) {
    const int color[3] = {c & 1, (c >> 1) & 1, (c >> 2) & 1};
    applyOp(m_jacobiTmp, phi, true);
    ops::gsrb_lazy<Real>(*m_lay, phi.d, m_jacobiTmp.d, rhs.d, color, m_dx, m_alpha, m_beta);
  }
}
// levelMultiColor (:1284-1331) over amrpgetMultiColors' 3D order (:36-56): for each
// colour a homogeneous ghost refresh, then GSMCAMRPOP on the cells of that colour
template <class Real>
void AMRPoissonOp<Real>::levelMultiColor(T& phi, const T& rhs) {
  if (m_vc) throw std::runtime_error("levelMultiColor is not implemented for variable coefficients");
  need(phi.lay, m_lay, "levelMultiColor"); need(rhs.lay, m_lay, "levelMultiColor");
  static const int colors[8][3] = {{0, 0, 0}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1},
                                   {0, 1, 0}, {1, 0, 0}, {0, 0, 1}, {1, 1, 1}};
  for (int c = 0; c < 8; c++) {
    fillGhosts(phi, true, true);
    ops::gsmc<Real>(*m_lay, phi.d, rhs.d, colors[c], m_dx, m_alpha, m_beta);
  }
}
// levelJacobi (AMRPoissonOp.cpp:1579-1596; VCAMRPoissonOp2.cpp:762-786): the homogeneous
// residual, times 0.666 / (alpha - 2 D beta / dx^2), or times lambda and 0.5 with
// variable coefficients, added to phi
template <class Real>
void AMRPoissonOp<Real>::levelJacobi(T& phi, const T& rhs) {
  if (!m_jacobiTmp.isDefined()) m_jacobiTmp.define(m_lay);
  residual(m_jacobiTmp, phi, rhs, true);
  if (m_vc) {
    vc::scale_by_lambda<Real>(*m_lay, *m_vc, m_jacobiTmp.d);
    incr(phi, m_jacobiTmp, 0.5);
  } else {
    const double weight = m_alpha - 2.0 * 3 * m_beta / (m_dx * m_dx);
    incr(phi, m_jacobiTmp, 0.666 / weight);
  }
}
// restrictResidual (:651-703)
template <class Real>
void AMRPoissonOp<Real>::restrictResidual(T& resC, T& phi, const T& rhs) {
  if (phi.nc > 1) {
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> resC_(resC, c), phi_(phi, c), rhs_(rhs, c);
      restrictResidual(resC_.v, phi_.v, rhs_.v);
    }
    return;
  }
  need(rhs.lay, m_lay, "restrictResidual"); need(resC.lay, m_mgCoarse, "restrictResidual");
  fillGhosts(phi, true, true);
  if (m_vc)  // VCAMRPoissonOp2::restrictResidual (VCAMRPoissonOp2.cpp:195-256)
    vc::restrict_res<Real>(*m_lay, *m_mgCoarse, *m_vc, resC.d, phi.d, rhs.d, m_dx, m_alpha, m_beta);
  else
    ops::restrict_res<Real>(*m_lay, *m_mgCoarse, resC.d, phi.d, rhs.d, m_dx, m_alpha, m_beta);
}
// prolongIncrement (:706-733)
template <class Real>
void AMRPoissonOp<Real>::prolongIncrement(T& phi, const T& c) {
  if (phi.nc > 1) {
    for (int q = 0; q < phi.nc; q++) {
      CompView<Real> phi_(phi, q), c_(c, q);
      prolongIncrement(phi_.v, c_.v);
    }
    return;
  }
  need(phi.lay, m_lay, "prolongIncrement"); need(c.lay, m_mgCoarse, "prolongIncrement");
  ops::prolong<Real>(*m_lay, *m_mgCoarse, phi.d, c.d);
}

// ---- AMR ----
template <class Real>
void AMRPoissonOp<Real>::coarseFineInterp(T& phif, const T& phic) {
  if (phif.nc > 1) {
    for (int c = 0; c < phif.nc; c++) {
      CompView<Real> phif_(phif, c), phic_(phic, c);
      coarseFineInterp(phif_.v, phic_.v);
    }
    return;
  }
  if (!m_coarserLay) return;
  need(phif.lay, m_lay, "coarseFineInterp"); need(phic.lay, m_coarserLay, "coarseFineInterp");
  ops::cf_inhomog<Real>(m_ctx, d_cfc, n_cfc, d_cfg, d_cff, n_cff, d_deriv, phif.d, phic.d, m_dx,
                        m_refToCoarser * m_dx, m_refToCoarser);
}
// AMRResidual (:736-750)
template <class Real>
void AMRPoissonOp<Real>::AMRResidual(T& res, T& phiFine, T& phi, T& phiCoarse, const T& rhs, bool homog,
                                     AMRLevelOp<T>* finerOp) {
  if (phi.nc > 1) {
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> res_(res, c), phiFine_(phiFine, c), phi_(phi, c), phiCoarse_(phiCoarse, c), rhs_(rhs, c);
      AMRResidual(res_.v, phiFine_.v, phi_.v, phiCoarse_.v, rhs_.v, homog, finerOp);
    }
    return;
  }
  if (!m_vc && phiFine.isDefined()) {
    if (phiCoarse.isDefined()) coarseFineInterp(phi, phiCoarse);
    residualFused(res, phiFine, phi, rhs, homog, static_cast<AMRPoissonOp*>(finerOp));
    return;
  }
  AMROperator(res, phiFine, phi, phiCoarse, homog, finerOp);
  axby(res, res, rhs, -1.0, 1.0);
}
template <class Real>
void AMRPoissonOp<Real>::AMRResidualNC(T& res, T& phiFine, T& phi, const T& rhs, bool homog,
                                       AMRLevelOp<T>* finerOp) {
  if (phi.nc > 1) {
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> res_(res, c), phiFine_(phiFine, c), phi_(phi, c), rhs_(rhs, c);
      AMRResidualNC(res_.v, phiFine_.v, phi_.v, rhs_.v, homog, finerOp);
    }
    return;
  }
  if (!m_vc && phiFine.isDefined()) {
    residualFused(res, phiFine, phi, rhs, homog, static_cast<AMRPoissonOp*>(finerOp));
    return;
  }
  AMROperatorNC(res, phiFine, phi, homog, finerOp);
  axby(res, res, rhs, -1.0, 1.0);
}
// applyOpI + reflux + axby(res, res, rhs, -1, 1) of AMRResidual / AMRResidualNC in two
// launches (valid cells, then the refluxed cells), bitwise the same values
template <class Real>
void AMRPoissonOp<Real>::residualFused(T& res, T& phiFine, T& phi, const T& rhs, bool homog, AMRPoissonOp* finerOp) {
  if (phi.nc > 1) {
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> res_(res, c), phiFine_(phiFine, c), phi_(phi, c), rhs_(rhs, c);
      residualFused(res_.v, phiFine_.v, phi_.v, rhs_.v, homog, finerOp);
    }
    return;
  }
  need(res.lay, m_lay, "AMRResidual"); need(rhs.lay, m_lay, "AMRResidual");
  need(phiFine.lay, m_finerLay, "AMRResidual"); need(phi.lay, m_lay, "AMRResidual");
  fillGhosts(phi, false, homog);
  ops::apply_residual<Real>(*m_lay, res.d, phi.d, rhs.d, m_dx, m_alpha, m_beta);
  finerOp->coarseFineInterp(phiFine, phi);
  const double a_scale = 1.0 / m_dx;
  ops::reflux_residual<Real>(m_ctx, d_rcell, n_rcell, d_rc, d_rf, res.d, phi.d, phiFine.d, rhs.d, m_beta * 1 / m_dx,
                             m_beta * m_refToFiner / m_dx, -a_scale, m_dx, m_alpha, m_beta);
}
// AMRResidualNF (:770-787)
template <class Real>
void AMRPoissonOp<Real>::AMRResidualNF(T& res, T& phi, T& phiCoarse, const T& rhs, bool homog) {
  if (phi.nc > 1) {
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> res_(res, c), phi_(phi, c), phiCoarse_(phiCoarse, c), rhs_(rhs, c);
      AMRResidualNF(res_.v, phi_.v, phiCoarse_.v, rhs_.v, homog);
    }
    return;
  }
  if (phiCoarse.isDefined()) coarseFineInterp(phi, phiCoarse);
  residualI(res, phi, rhs, homog);
}
template <class Real>
void AMRPoissonOp<Real>::residualNorm(T& lhs, T& phi, const T& rhs, bool homog, int slot, bool accumulate) {
  if (!canFuseNorm(lhs)) throw std::runtime_error("residualNorm: one component, constant coefficients");
  need(lhs.lay, m_lay, "residual"); need(rhs.lay, m_lay, "residual");
  fillGhosts(phi, homog, homog);
  ops::residual_norm<Real>(*m_lay, lhs.d, phi.d, rhs.d, m_dx, m_alpha, m_beta, slot, accumulate ? 1 : 0);
}
template <class Real>
void AMRPoissonOp<Real>::AMRResidualNFNorm(T& res, T& phi, T& phiCoarse, const T& rhs, bool homog, int slot,
                                          bool accumulate) {
  if (!canFuseNorm(res)) throw std::runtime_error("AMRResidualNFNorm: one component, constant coefficients");
  if (phiCoarse.isDefined()) coarseFineInterp(phi, phiCoarse);
  need(res.lay, m_lay, "residualI"); need(rhs.lay, m_lay, "residualI");
  fillGhosts(phi, false, homog);
  ops::residual_norm<Real>(*m_lay, res.d, phi.d, rhs.d, m_dx, m_alpha, m_beta, slot, accumulate ? 1 : 0);
}
// AMROperator (:790-815)
template <class Real>
void AMRPoissonOp<Real>::AMROperator(T& LofPhi, T& phiFine, T& phi, T& phiCoarse, bool homog,
                                     AMRLevelOp<T>* finerOp) {
  if (phi.nc > 1) {
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> LofPhi_(LofPhi, c), phiFine_(phiFine, c), phi_(phi, c), phiCoarse_(phiCoarse, c);
      AMROperator(LofPhi_.v, phiFine_.v, phi_.v, phiCoarse_.v, homog, finerOp);
    }
    return;
  }
  if (phiCoarse.isDefined()) coarseFineInterp(phi, phiCoarse);
  applyOpI(LofPhi, phi, homog);
  if (phiFine.isDefined()) reflux(phiFine, phi, LofPhi, static_cast<AMRPoissonOp*>(finerOp));
}
template <class Real>
void AMRPoissonOp<Real>::AMROperatorNC(T& LofPhi, T& phiFine, T& phi, bool homog, AMRLevelOp<T>* finerOp) {
  if (phi.nc > 1) {
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> LofPhi_(LofPhi, c), phiFine_(phiFine, c), phi_(phi, c);
      AMROperatorNC(LofPhi_.v, phiFine_.v, phi_.v, homog, finerOp);
    }
    return;
  }
  applyOpI(LofPhi, phi, homog);
  if (phiFine.isDefined()) reflux(phiFine, phi, LofPhi, static_cast<AMRPoissonOp*>(finerOp));
}
template <class Real>
void AMRPoissonOp<Real>::AMROperatorNF(T& LofPhi, T& phi, T& phiCoarse, bool homog) {
  if (phi.nc > 1) {
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> LofPhi_(LofPhi, c), phi_(phi, c), phiCoarse_(phiCoarse, c);
      AMROperatorNF(LofPhi_.v, phi_.v, phiCoarse_.v, homog);
    }
    return;
  }
  if (phiCoarse.isDefined()) coarseFineInterp(phi, phiCoarse);
  applyOpI(LofPhi, phi, homog);
}
// AMRRestrictS (:877-919)
template <class Real>
void AMRPoissonOp<Real>::AMRRestrictS(T& resCoarse, const T& residual, T& correction, T& coarseCorrection,
                                      T& scratch) {
  if (correction.nc > 1) {
    for (int c = 0; c < correction.nc; c++) {
      CompView<Real> resCoarse_(resCoarse, c), residual_(residual, c), correction_(correction, c), coarseCorrection_(coarseCorrection, c), scratch_(scratch, c);
      AMRRestrictS(resCoarse_.v, residual_.v, correction_.v, coarseCorrection_.v, scratch_.v);
    }
    return;
  }
  need(resCoarse.lay, m_cfLay, "AMRRestrictS");
  if (!m_vc) {
    // AMRResidualNF (coarse-fine interpolation, homogeneous ghost refresh,
    // OPERATORLAPRES) and the average in one pass: scratch is only scratch space
    // here and is not written
    need(correction.lay, m_lay, "AMRRestrictS"); need(residual.lay, m_lay, "AMRRestrictS");
    if (coarseCorrection.isDefined()) coarseFineInterp(correction, coarseCorrection);
    fillGhosts(correction, false, true);
    ops::residual_average<Real>(*m_lay, *m_cfLay, resCoarse.d, correction.d, residual.d, m_dx, m_alpha, m_beta,
                                m_refToCoarser);
    return;
  }
  AMRResidualNF(scratch, correction, coarseCorrection, residual, true);
  ops::average<Real>(*m_lay, *m_cfLay, resCoarse.d, scratch.d, m_refToCoarser);
}
// AMRProlongS (:959-990)
template <class Real>
void AMRPoissonOp<Real>::AMRProlongS(T& correction, const T& coarseCorrection, T& temp) {
  if (correction.nc > 1) {
    for (int c = 0; c < correction.nc; c++) {
      CompView<Real> correction_(correction, c), coarseCorrection_(coarseCorrection, c), temp_(temp, c);
      AMRProlongS(correction_.v, coarseCorrection_.v, temp_.v);
    }
    return;
  }
  need(temp.lay, m_cfLay, "AMRProlongS"); need(coarseCorrection.lay, m_coarserLay, "AMRProlongS");
  need(correction.lay, m_lay, "AMRProlongS");
  ops::motion<Real>(m_revCopier, *m_coarserLay, *m_cfLay, temp.d, coarseCorrection.d, 0);
  ops::prolong<Real>(*m_lay, *m_cfLay, correction.d, temp.d, m_refToCoarser);
}
// AMRUpdateResidual (:1045-1054)
template <class Real>
void AMRPoissonOp<Real>::AMRUpdateResidual(T& residual, T& correction, T& coarseCorrection) {
  AMRResidualNF(residual, correction, coarseCorrection, residual, true);
}
template <class Real>
void AMRPoissonOp<Real>::createCoarsened(T& lhs, const T& rhs, int) {
  lhs.define(m_cfLay, rhs.nc);
}
// assignCopier (:486-492): lhs (this level) <- rhs (coarsened finer level)
template <class Real>
void AMRPoissonOp<Real>::assignCopier(T& lhs, const T& rhs) {
  if (lhs.nc > 1) {
    for (int c = 0; c < lhs.nc; c++) {
      CompView<Real> lhs_(lhs, c), rhs_(rhs, c);
      assignCopier(lhs_.v, rhs_.v);
    }
    return;
  }
  need(lhs.lay, m_lay, "assignCopier"); need(rhs.lay, m_finerCfLay, "assignCopier");
  ops::motion<Real>(*m_zeroPlan, *m_finerCfLay, *m_lay, lhs.d, rhs.d, 0);
}
// zeroCovered (:495-502) -> LevelDataOps::copyToZero
template <class Real>
void AMRPoissonOp<Real>::zeroCovered(T& lhs) {
  if (lhs.nc > 1) {
    for (int c = 0; c < lhs.nc; c++) {
      CompView<Real> lhs_(lhs, c);
      zeroCovered(lhs_.v);
    }
    return;
  }
  need(lhs.lay, m_lay, "zeroCovered");
  if (!m_zeroPlan) throw std::runtime_error("zeroCovered: no finer level");
  ops::motion<Real>(*m_zeroPlan, *m_finerCfLay, *m_lay, lhs.d, (const Real*)nullptr, 1);
}
// reflux (:1117-1201)
template <class Real>
void AMRPoissonOp<Real>::reflux(T& phiFine, T& phi, T& res, AMRPoissonOp* finerOp) {
  if (phi.nc > 1) {
    for (int c = 0; c < phi.nc; c++) {
      CompView<Real> phiFine_(phiFine, c), phi_(phi, c), res_(res, c);
      reflux(phiFine_.v, phi_.v, res_.v, finerOp);
    }
    return;
  }
  need(phiFine.lay, m_finerLay, "reflux"); need(phi.lay, m_lay, "reflux"); need(res.lay, m_lay, "reflux");
  if (!m_vc != !finerOp->m_vc) throw std::runtime_error("reflux: variable coefficients on one level of two");
  finerOp->coarseFineInterp(phiFine, phi);
  double a_scale = 1.0 / m_dx;
  if (m_vc) {  // VCAMRPoissonOp2::reflux (VCAMRPoissonOp2.cpp:339-430) with getFlux (:787-837)
    if (!d_vcRcFace || m_vcRefluxFine != finerOp->m_vc.get()) setupVCReflux(*finerOp->m_vc);
    vc::reflux<Real>(m_ctx, d_rcell, n_rcell, d_rc, d_rf, d_vcRcFace, d_vcRfFace, res.d, phi.d, phiFine.d, *m_vc,
                     *finerOp->m_vc, m_beta * 1 / m_dx, m_beta * m_refToFiner / m_dx, -a_scale);
    return;
  }
  ops::reflux<Real>(m_ctx, d_rcell, n_rcell, d_rc, d_rf, res.d, phi.d, phiFine.d, m_beta * 1 / m_dx,
                    m_beta * m_refToFiner / m_dx, -a_scale);
}

// ---------------------------------------------------------------------------
// DeviceHierarchy / factory
// ---------------------------------------------------------------------------
// MultiGrid::define on AMR level l via MGnewOp (AMRPoissonOp.cpp:1955-1980): the level's
// boxes coarsened by 2, 4, ... while coarsenable by coarsening * s_maxCoarse (2)
void DeviceHierarchy::defineLevelMG(int l) {
  if (l < 1 || l >= nlev()) throw std::runtime_error("defineLevelMG: level out of range");
  if (!lmgLay[l].empty() || !grids[l].coarsenable(4)) return;
  for (int depth = 1;; depth++) {
    const int coarsening = 1 << depth;
    if (!grids[l].coarsenable(coarsening * 2)) break;
    if (depth == 1 && ratios[l] == 2) {
      lmgLay[l].push_back(cfLay[l].get());  // coarsen(level, 2) over coarsen(domain, 2): the same layout
      continue;
    }
    lmgOwned[l].emplace_back(
        new DeviceLayout(ctx, grids[l].coarsened(coarsening), coarsen(domains[l], coarsening), merged));
    lmgLay[l].push_back(lmgOwned[l].back().get());
  }
}

void DeviceHierarchy::define(Context* c, const std::vector<DBL>& g, const Box& domain0, double dx0, bool merge,
                             const BCSpec* bc, const int* ratio) {
  ctx = c;
  grids = g;
  merged = merge;
  lmgLay.assign(g.size(), {});
  lmgOwned.clear();
  lmgOwned.resize(g.size());
  domains.assign(g.size(), domain0);
  dxs.assign(g.size(), dx0);
  ratios.assign(g.size(), 1);
  for (size_t l = 1; l < g.size(); l++) {
    const int r = ratio ? ratio[l - 1] : 2;
    if (r != 2 && r != 4) throw std::runtime_error("refinement ratio must be 2 or 4");
    ratios[l] = r;
    domains[l] = refine(domains[l - 1], r);
    dxs[l] = dxs[l - 1] / r;
  }
  // periodic directions: a layout with a box on a periodic face is stored per box
  // (its ghosts there come from periodic images); the levels above the base stay 2
  // coarse cells away from periodic faces, so coarse-fine stencils and flux
  // registers never reach across them
  bool per[3] = {false, false, false};
  if (bc)
    for (int d = 0; d < 3; d++) {
      if ((bc->type[0][d] == 3) != (bc->type[1][d] == 3))
        throw std::runtime_error("a periodic direction needs BC type 3 on both faces");
      per[d] = bc->type[0][d] == 3;
    }
  for (size_t l = 1; l < g.size(); l++)
    for (int d = 0; d < 3; d++)
      if (per[d])
        for (const Box& b : g[l].boxes)
          if (b.lo[d] < domains[l].lo[d] + 2 * ratios[l] || b.hi[d] > domains[l].hi[d] - 2 * ratios[l])
            throw std::runtime_error("a refined level comes within 2 coarse cells of a periodic face");
  auto touches = [&](const DBL& d, const Box& dom) {
    for (const Box& b : d.boxes)
      for (int q = 0; q < 3; q++)
        if (per[q] && (b.lo[q] == dom.lo[q] || b.hi[q] == dom.hi[q])) return true;
    return false;
  };
  amrLay.clear();
  cfLay.clear();
  mgLay.clear();
  for (size_t l = 0; l < g.size(); l++) {
    if (!g[l].sorted()) throw std::runtime_error("boxes of a level must be sorted (Box::operator<)");
    amrLay.emplace_back(new DeviceLayout(c, g[l], domains[l], merge && !touches(g[l], domains[l])));
    if (l == 0) cfLay.emplace_back(nullptr);
    else {
      if (!g[l].coarsenable(ratios[l]))
        throw std::runtime_error("fine level boxes must be coarsenable by the refinement ratio");
      cfLay.emplace_back(new DeviceLayout(c, g[l].coarsened(ratios[l]), domains[l - 1], merge));
    }
  }
  // MultiGrid::define's m_bottomCells for each level's m_mg (MultiGrid.H:477-508): the
  // level's domain divided by 8 for every multigrid level MGnewOp creates
  mgBottomCells.assign(g.size(), 0);
  for (size_t l = 0; l < g.size(); l++) {
    long long cells = domains[l].npts();
    for (int cz = 2; cz < (1 << 24) && g[l].coarsenable(cz * 2); cz *= 2) cells /= 8;
    mgBottomCells[l] = cells;
  }
  // MultiGrid::define via MGnewOp: stop when not coarsenable by coarsening*2
  // (AMRPoissonOp.cpp:1955-1959, s_maxCoarse = 2)
  for (int depth = 1;; depth++) {
    int coarsening = 1 << depth;
    if (!g[0].coarsenable(coarsening * 2)) break;
    const DBL gc = g[0].coarsened(coarsening);
    const Box dc = coarsen(domain0, coarsening);
    mgLay.emplace_back(new DeviceLayout(c, gc, dc, merge && !touches(gc, dc)));
  }
}

// ---------------------------------------------------------------------------
// host validation of device tables: every address a kernel forms from a table
// entry must lie inside the field slab it indexes.  Violations throw before any
// launch.
// ---------------------------------------------------------------------------
static void require(bool ok, const char* what) {
  if (!ok) throw std::runtime_error(std::string("device table check failed: ") + what);
}
static bool in_grown(const DBL& d, int b, const int lo[3], const int hi[3], int g) {
  if (b < 0 || b >= d.size()) return false;
  Box gb = grow(d[b], g);
  return gb.contains(Box(IV{lo[0], lo[1], lo[2]}, IV{hi[0], hi[1], hi[2]}));
}
static void validate_ghost_plan(const ops::GhostPlan& P, const DBL& d) {
  for (const Patch& p : P.h_patches) {
    require(in_grown(d, p.dst, p.lo, p.hi, 1), "ghost patch outside its box");
    if (p.kind == PK_EXCHANGE) {
      const int slo[3] = {p.lo[0] + p.sh[0], p.lo[1] + p.sh[1], p.lo[2] + p.sh[2]};
      const int shi[3] = {p.hi[0] + p.sh[0], p.hi[1] + p.sh[1], p.hi[2] + p.sh[2]};
      require(in_grown(d, p.src, slo, shi, 0), "exchange source outside its box");
    }
    else {
      // the cells read (one and two inward) must stay inside the grown box
      int lo[3] = {p.lo[0], p.lo[1], p.lo[2]}, hi[3] = {p.hi[0], p.hi[1], p.hi[2]};
      int s = p.side ? -2 : 2;
      lo[p.dir] += s;
      hi[p.dir] += s;
      require(in_grown(d, p.dst, lo, hi, 1), "ghost rule reads outside its box");
      // quadratic Dirichlet reads two valid cells: the box must be two cells wide there
      if (p.kind == PK_DIRICHLET2) require(in_grown(d, p.dst, lo, hi, 0), "quadratic Dirichlet face of a box one cell wide");
    }
  }
}
static void validate_motion(const ops::MotionPlan& P, const DBL& src, const DBL& dst) {
  for (const Motion& m : P.h_items) {
    require(in_grown(src, m.from, m.lo, m.hi, 1), "copier source region outside its box");
    require(in_grown(dst, m.to, m.lo, m.hi, 1), "copier destination region outside its box");
  }
}
static void validate_pair(const DeviceLayout& f, const DeviceLayout& c, int ratio = 2) {
  require(f.dbl.size() == c.dbl.size(), "coarsened layout must match the fine layout box by box");
  for (int b = 0; b < f.dbl.size(); b++) require(coarsen(f.dbl[b], ratio) == c.dbl[b], "coarsened box mismatch");
}

// The one-launch relaxation applies when the level is stored merged and every
// ghost rule is the rule of one whole face of its rectangle: homogeneous
// Dirichlet (linear) or Neumann, or quadratic coarse-fine (a coarse-fine face of a
// box one cell wide uses the linear rule and keeps the two-launch path; so does a
// quadratic Dirichlet face, whose rule the one-launch kernel does not form).
template <class Real>
static void setup_fused(AMRPoissonOp<Real>* op) {
  const DeviceLayout& L = *op->m_lay;
  op->m_fused = false;
  if (!L.merged) return;
  int kind[3][2] = {{-1, -1}, {-1, -1}, {-1, -1}};
  for (const Patch& p : op->m_ghost.h_patches) {
    int k;
    if (p.kind == PK_NEUMANN) k = 0;
    else if (p.kind == PK_DIRICHLET) k = 1;
    else if (p.kind == PK_CFHOMO) k = 2;
    else return;
    const int face = p.side ? L.rect.hi[p.dir] + 1 : L.rect.lo[p.dir] - 1;
    if (p.lo[p.dir] != face || p.hi[p.dir] != face) return;
    if (kind[p.dir][p.side] == -1) kind[p.dir][p.side] = k;
    else if (kind[p.dir][p.side] != k) return;
  }
  for (int d = 0; d < 3; d++)
    for (int s = 0; s < 2; s++) {
      if (kind[d][s] < 0) return;
      op->m_faces.kind[d][s] = kind[d][s];
    }
  // the coefficients k_fill_ghosts forms for PK_CFHOMO (AMRPoissonOp.cpp:1774-1775)
  const double dx = op->m_dx, dxCrse = op->m_dxCrse;
  op->m_faces.c1 = 2.0 * (dxCrse - dx) / (dxCrse + dx);
  op->m_faces.c2 = -((dxCrse - dx) / (dxCrse + 3.0 * dx));
  op->m_fused = true;
}

// AMRPoissonOp::define (AMRPoissonOp.cpp:60-221) on one level: exchange, BC and (when
// the level has a coarser one) homogeneous coarse-fine ghost rules from the level's
// own CFRegion, and the one-launch relaxation where it applies.
template <class Real>
void defineLevelOp(AMRPoissonOp<Real>* op, Context* ctx, const DeviceLayout* lay, const Box& domain, double dx,
                   double dxCrse, double alpha, double beta, const BCSpec& bc, const DeviceLayout* mgCoarse) {
  op->m_ctx = ctx;
  op->m_lay = lay;
  op->m_domain = domain;
  op->m_dx = dx;
  op->m_dxCrse = dxCrse;
  op->m_alpha = alpha;
  op->m_beta = beta;
  op->m_bc = bc;
  op->m_mgCoarse = mgCoarse;
  op->m_ghost = build_ghost_plan(lay->dbl, domain, bc, dxCrse > 0, lay->merged);
  validate_ghost_plan(op->m_ghost, lay->dbl);
  if (bc.fn) attach_bc_values(op->m_ghost, dx, bc, op->m_ncomp);
  op->m_ghost.upload(ctx);
  setup_fused(op);
  if (op->m_mgCoarse) validate_pair(*op->m_lay, *op->m_mgCoarse);
}

// A BCValueHolder for the operator's domain faces (fn null: the face constants again);
// every later inhomogeneous ghost refresh uses the values fn gives here.
template <class Real>
void setLevelOpBCValues(AMRPoissonOp<Real>* op, BCValueFn fn, void* user) {
  op->m_bc.fn = fn;
  op->m_bc.user = user;
  attach_bc_values(op->m_ghost, op->m_dx, op->m_bc, op->m_ncomp);
  op->m_ghost.upload(op->m_ctx);
}

// The coarse-fine part of AMRnewOp for a level with a coarser AMR level (refinement
// ratio 2 or 4): QuadCFInterp tables, the copiers between coarsen(fine, ratio) (cfLay)
// and the coarser level, and the coarser operator's reflux tables and zeroCovered plan.
template <class Real>
void linkCoarser(AMRPoissonOp<Real>* op, AMRPoissonOp<Real>* coarserOp, const DeviceLayout* cfLay, int ratio) {
  Context* ctx = op->m_ctx;
  const DeviceLayout& F = *op->m_lay;
  const DeviceLayout& Cl = *coarserOp->m_lay;
  if (ratio != 2 && ratio != 4) throw std::runtime_error("refinement ratio must be 2 or 4");
  if (!(refine(coarserOp->m_domain, ratio) == op->m_domain))
    throw std::runtime_error("coarser level: domain must be coarsen(domain, ratio)");
  op->m_refToCoarser = ratio;
  op->m_coarserLay = &Cl;
  op->m_cfLay = cfLay;
  CFPlanHost cf = build_cf_plan(F.dbl, F, Cl.dbl, Cl, op->m_domain, op->m_dx, ratio);
  {
    const long long fs = F.slab, cs = Cl.slab;
    for (const CFCoarse& c : cf.coarse) {
      require(c.pc >= 0 && c.pc < cs, "CF coarse cell");
      if (c.std)
        for (int q = 0; q < 8; q++) require(c.nbr[q] >= 0 && c.nbr[q] < cs, "CF coarse neighbour");
      else require(c.gen >= 0 && c.gen < (int)cf.gens.size(), "CF stencil index");
    }
    for (const CFGen& g : cf.gens)
      for (int q = 0; q < 5; q++) {
        require(g.n[q] >= 0 && g.n[q] <= 9, "CF stencil length");
        for (int t = 0; t < g.n[q]; t++) require(g.off[q][t] >= 0 && g.off[q][t] < cs, "CF stencil cell");
      }
    for (const CFFine& f : cf.fine) {
      require(f.g >= 0 && f.g < fs && f.a1 >= 0 && f.a1 < fs && f.a2 >= 0 && f.a2 < fs, "CF fine cell");
      require(f.coarse >= 0 && f.coarse < (int)cf.coarse.size(), "CF fine record");
    }
  }
  op->n_cfc = (int)cf.coarse.size();
  op->n_cfg = (int)cf.gens.size();
  op->n_cff = (int)cf.fine.size();
  op->d_cfc = upload_vec(ctx, cf.coarse);
  op->d_cfg = upload_vec(ctx, cf.gens);
  op->d_cff = upload_vec(ctx, cf.fine);
  dev_alloc(ctx, (void**)&op->d_deriv, sizeof(Real) * 6 * (size_t)op->n_cfc, "hipMalloc deriv");
  const DBL fc = F.dbl.coarsened(ratio);
  op->m_resCopier = build_copier(fc, Cl.dbl);
  validate_motion(op->m_resCopier, fc, Cl.dbl);
  op->m_resCopier.upload(ctx);
  op->m_revCopier = build_copier(Cl.dbl, fc);
  validate_motion(op->m_revCopier, Cl.dbl, fc);
  op->m_revCopier.upload(ctx);
  validate_pair(*op->m_lay, *op->m_cfLay, ratio);
  // the coarser operator refluxes against this level and zeroes what it covers
  coarserOp->m_refToFiner = ratio;
  coarserOp->m_finerLay = op->m_lay;
  coarserOp->m_zeroPlan = &op->m_resCopier;
  coarserOp->m_finerCfLay = op->m_cfLay;
  RefluxPlanHost rp = build_reflux_plan(Cl.dbl, Cl, F.dbl, F, coarserOp->m_domain, ratio);
  {
    const long long fs = F.slab, cs = Cl.slab;
    for (const RefluxCell& c : rp.cells) {
      require(c.u >= 0 && c.u < cs, "reflux cell");
      require(c.c0 >= 0 && c.nc >= 0 && c.c0 + c.nc <= (int)rp.coarse.size(), "reflux coarse range");
      require(c.f0 >= 0 && c.nf >= 0 && c.f0 + c.nf <= (int)rp.fine.size(), "reflux fine range");
    }
    for (const RefluxCoarse& x : rp.coarse) require(x.a >= 0 && x.a < cs && x.b >= 0 && x.b < cs, "reflux coarse face");
    for (const RefluxFine& x : rp.fine)
      for (int t = 0; t < 4; t++) require(x.a[t] >= 0 && x.a[t] < fs && x.b[t] >= 0 && x.b[t] < fs, "reflux fine face");
  }
  dev_free(ctx, coarserOp->d_rcell);
  dev_free(ctx, coarserOp->d_rc);
  dev_free(ctx, coarserOp->d_rf);
  coarserOp->n_rcell = (int)rp.cells.size();
  coarserOp->m_rcFace = rp.coarseFace;
  coarserOp->m_rfFace = rp.fineFace;
  coarserOp->d_rcell = upload_vec(ctx, rp.cells);
  coarserOp->d_rc = upload_vec(ctx, rp.coarse);
  coarserOp->d_rf = upload_vec(ctx, rp.fine);
}

template <class Real>
std::unique_ptr<AMRPoissonOp<Real>> AMRPoissonOpFactory<Real>::AMRnewOp(int l, AMRPoissonOp<Real>* coarserOp) {
  std::unique_ptr<AMRPoissonOp<Real>> op(new AMRPoissonOp<Real>());
  op->m_ncomp = ncomp;
  // the multigrid child: level 0's first multigrid depth; for a level above the base
  // with ratio 2, coarsen(level, 2) is its coarsened-fine layout (ratio 4 has none)
  const DeviceLayout* mgc = (l == 0 && !H->mgLay.empty()) ? H->mgLay[0].get()
                            : (l > 0 && H->ratios[l] == 2 ? H->cfLay[l].get() : nullptr);
  defineLevelOp(op.get(), H->ctx, H->amrLay[l].get(), H->domains[l], H->dxs[l], (l > 0) ? H->dxs[l - 1] : -1.0,
                alpha, beta, bc, mgc);
  if (l > 0) linkCoarser(op.get(), coarserOp, H->cfLay[l].get(), H->ratios[l]);
  return op;
}

template <class Real>
std::vector<std::unique_ptr<MGLevelOp<LevelData<Real>>>> AMRPoissonOpFactory<Real>::MGnewOpsLevel(int l) {
  std::vector<std::unique_ptr<MGLevelOp<LevelData<Real>>>> out;
  H->defineLevelMG(l);
  const auto& lay = H->lmgLay[l];
  for (size_t d = 0; d < lay.size(); d++) {
    const int coarsening = 1 << (d + 1);
    AMRPoissonOp<Real>* op = new AMRPoissonOp<Real>();
    op->m_ncomp = ncomp;
    // MGnewOp keeps the AMR level's dxCrse: the coarse-fine faces of every depth get
    // the homogeneous interpolation with the coarser AMR level's spacing
    defineLevelOp(op, H->ctx, lay[d], coarsen(H->domains[l], coarsening), H->dxs[l] * coarsening, H->dxs[l - 1],
                  alpha, beta, bc, (d + 1 < lay.size()) ? lay[d + 1] : nullptr);
    out.emplace_back(op);
  }
  return out;
}
template <class Real>
std::vector<std::unique_ptr<MGLevelOp<LevelData<Real>>>> AMRPoissonOpFactory<Real>::MGnewOps() {
  std::vector<std::unique_ptr<MGLevelOp<LevelData<Real>>>> out;
  for (size_t d = 0; d < H->mgLay.size(); d++) {
    int coarsening = 1 << (d + 1);
    AMRPoissonOp<Real>* op = new AMRPoissonOp<Real>();
    op->m_ncomp = ncomp;
    defineLevelOp(op, H->ctx, H->mgLay[d].get(), coarsen(H->domains[0], coarsening), H->dxs[0] * coarsening, -1.0,
                  alpha, beta, bc, (d + 1 < H->mgLay.size()) ? H->mgLay[d + 1].get() : nullptr);
    out.emplace_back(op);
  }
  return out;
}

// ---------------------------------------------------------------------------
// AMRMultiGrid
// ---------------------------------------------------------------------------
template <class Real>
void AMRMultiGrid<Real>::define(AMRPoissonOpFactory<Real>& factory, LinearSolver<T>* bottom, int numLevels) {
  m_op.clear();
  for (int l = 0; l < numLevels; l++) m_op.push_back(factory.AMRnewOp(l, l > 0 ? m_op[l - 1].get() : nullptr));
  m_bottomSolver = bottom;
  m_factory = &factory;
  m_mgL.clear();
  m_mgL.resize(numLevels);
  m_mgBottomCells = factory.H->mgBottomCells;
  m_mg0.reset(new MultiGrid<T>());
  m_mg0->define(m_op[0].get(), factory.MGnewOps(), factory.H->domains[0].npts(), bottom);
  m_mg0->m_pre = m_pre;
  m_mg0->m_post = m_post;
  m_mg0->m_bottom = m_bottom;
  m_correction.resize(numLevels);
  m_residual.resize(numLevels);
  m_resC.resize(numLevels);
}
// setSolverParameters (AMRMultiGrid.H:528-553)
template <class Real>
void AMRMultiGrid<Real>::setSolverParameters(int pre, int post, int bottom, int numMG, int iterMax, double eps,
                                             double hang, double normThresh) {
  m_pre = pre; m_post = post; m_bottom = bottom; m_eps = eps; m_hang = hang; m_normThresh = normThresh;
  m_iterMax = iterMax; m_numMG = numMG;
  m_mg0->m_pre = pre; m_mg0->m_post = post; m_mg0->m_bottom = bottom;
  m_mg0->m_numMG = numMG; m_mg0->m_cycle = numMG;
  for (auto& m : m_mgL)
    if (m) {
      m->m_pre = pre; m->m_post = post; m->m_bottom = bottom;
      m->m_numMG = numMG; m->m_cycle = numMG;
    }
  m_bottomSolverEpsCushion = 1.0;
}
// init (AMRMultiGrid.H:1212-1249)
template <class Real>
void AMRMultiGrid<Real>::init(std::vector<T*>& phi, std::vector<T*>& rhs, int l_max, int l_base) {
  for (int i = l_base; i <= l_max; i++) {
    AMRPoissonOp<Real>& op = *m_op[i];
    m_correction[i].reset(new T());
    m_residual[i].reset(new T());
    op.create(*m_correction[i], *phi[i]);
    op.create(*m_residual[i], *rhs[i]);
    if (i == 0) m_mg0->init(*phi[i], *rhs[i]);
    if (i != l_base) {
      m_resC[i].reset(new T());
      op.createCoarsened(*m_resC[i], *rhs[i], 2);
    }
  }
}
template <class Real>
MultiGrid<LevelData<Real>>* AMRMultiGrid<Real>::mg(int l) {
  if (l == 0) return m_mg0.get();
  if (l < 0 || l >= (int)m_op.size()) throw std::runtime_error("l_base out of range");
  if (!m_mgL[l]) {
    if (m_op[l]->m_vc) throw std::runtime_error("l_base > 0 is not supported with variable coefficients");
    if (!m_correction[l] || !m_residual[l]) throw std::runtime_error("AMRMultiGrid: init first");
    AMRPoissonOp<Real>* top = m_op[l].get();
    auto ops = m_factory->MGnewOpsLevel(l);
    // the AMR level's multigrid child: coarsen(level, 2) (its coarsened-fine layout
    // serves at ratio 2; at ratio 4 the level has none of its own)
    if (!ops.empty() && !top->m_mgCoarse) top->m_mgCoarse = static_cast<AMRPoissonOp<Real>*>(ops[0].get())->m_lay;
    std::unique_ptr<MultiGrid<T>> m(new MultiGrid<T>());
    m->define(top, std::move(ops), m_factory->H->domains[l].npts(), nullptr);
    m->m_pre = m_pre; m->m_post = m_post; m->m_bottom = m_bottom;
    m->m_numMG = m_numMG; m->m_cycle = m_numMG;
    m->init(*m_correction[l], *m_residual[l]);
    m_mgL[l] = std::move(m);
  }
  return m_mgL[l].get();
}
template <class Real>
void AMRMultiGrid<Real>::relax(T& phi, T& R, int depth, int nRelax, bool fromZero) {
  // ratio 4 to the coarser level: the level's MultiGrid cycles with m_depth 1, its
  // bottom branch (AMRMultiGrid.H:638-657 with init's intermediateDepth,
  // MultiGrid.H:634-646): one relaxation when its m_bottomCells is 1, else m_bottom
  // (the bottom solver there is the no-op one)
  if (m_op[depth]->m_refToCoarser > 2) nRelax = (m_mgBottomCells[depth] == 1) ? 1 : m_bottom;
  if (fromZero) m_op[depth]->relaxFromZero(phi, R, nRelax);
  else m_op[depth]->relax(phi, R, nRelax);
}
// computeAMRResidualLevel (AMRMultiGrid.H:837-892)
template <class Real>
void AMRMultiGrid<Real>::computeAMRResidualLevel(std::vector<T*>& resid, std::vector<T*>& phi, std::vector<T*>& rhs,
                                                 int l_max, int l_base, int ilev, bool homog) {
  if (l_max != l_base) {
    if (ilev == l_max) {
      m_op[l_max]->AMRResidualNF(*resid[l_max], *phi[l_max], *phi[l_max - 1], *rhs[l_max], homog);
    } else if (ilev == l_base) {
      if (l_base == 0)
        m_op[l_base]->AMRResidualNC(*resid[l_base], *phi[l_base + 1], *phi[l_base], *rhs[l_base], homog,
                                    m_op[l_base + 1].get());
      else
        m_op[l_base]->AMRResidual(*resid[l_base], *phi[l_base + 1], *phi[l_base], *phi[l_base - 1], *rhs[l_base],
                                  homog, m_op[l_base + 1].get());
    } else {
      m_op[ilev]->AMRResidual(*resid[ilev], *phi[ilev + 1], *phi[ilev], *phi[ilev - 1], *rhs[ilev], homog,
                              m_op[ilev + 1].get());
    }
  } else {
    if (l_base == 0) m_op[l_max]->residual(*resid[l_max], *phi[l_max], *rhs[l_max], homog);
    else m_op[l_max]->AMRResidualNF(*resid[l_max], *phi[l_max], *phi[l_max - 1], *rhs[l_max], homog);
  }
}
// computeAMRResidual (AMRMultiGrid.H:683-736); the level max-norms accumulate on
// the device and are read once.
template <class Real>
double AMRMultiGrid<Real>::computeAMRResidual(std::vector<T*>& resid, std::vector<T*>& phi, std::vector<T*>& rhs,
                                              int l_max, int l_base, bool homog, bool computeNorm) {
  for (int ilev = l_base; ilev <= l_max; ilev++) {
    // the finest level's residual and its max-norm in one pass (the same values)
    if (computeNorm && ilev == l_max && m_op[ilev]->canFuseNorm(*resid[ilev])) {
      if (l_max == l_base && l_base == 0)
        m_op[ilev]->residualNorm(*resid[ilev], *phi[ilev], *rhs[ilev], homog, 0, ilev != l_base);
      else
        m_op[ilev]->AMRResidualNFNorm(*resid[ilev], *phi[ilev], *phi[ilev - 1], *rhs[ilev], homog, 0,
                                      ilev != l_base);
      continue;
    }
    computeAMRResidualLevel(resid, phi, rhs, l_max, l_base, ilev, homog);
    if (computeNorm) {
      if (ilev != l_max) m_op[ilev]->zeroCovered(*resid[ilev]);
      m_op[ilev]->localMaxNormAsync(*resid[ilev], 0, ilev != l_base);
    }
  }
  if (!computeNorm) return 0.0;
  Context* ctx = m_op[0]->m_ctx;
  hip_check(hipMemcpyAsync(ctx->h_scalars, ctx->d_scalars, sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
            "copy rnorm");
  ctx->sync();
  return ctx->h_scalars[0];
}
// computeAMROperator (AMRMultiGrid.H:790-833)
template <class Real>
void AMRMultiGrid<Real>::computeAMROperator(std::vector<T*>& lphi, std::vector<T*>& phi, int l_max, int l_base,
                                            bool homog) {
  for (int ilev = l_base; ilev <= l_max; ilev++) {
    if (l_max != l_base) {
      if (ilev == l_max) m_op[l_max]->AMROperatorNF(*lphi[l_max], *phi[l_max], *phi[l_max - 1], homog);
      else if (ilev == l_base && l_base == 0)
        m_op[l_base]->AMROperatorNC(*lphi[l_base], *phi[l_base + 1], *phi[l_base], homog, m_op[l_base + 1].get());
      else
        m_op[ilev]->AMROperator(*lphi[ilev], *phi[ilev + 1], *phi[ilev], *phi[ilev - 1], homog,
                                m_op[ilev + 1].get());
    } else {
      m_op[l_max]->applyOp(*lphi[l_max], *phi[l_max], homog);
    }
  }
}
// AMRVCycle (AMRMultiGrid.H:1342-1419)
template <class Real>
void AMRMultiGrid<Real>::AMRVCycle(std::vector<T*>& uberCorrection, std::vector<T*>& uberResidual, int ilev,
                                   int l_max, int l_base, bool fromZero) {
  if (ilev == l_max && l_max > l_base) {
    // (with one level the cycle works on uberCorrection / uberResidual directly and
    // the reference's copies below are never read)
    for (int level = l_base; level <= l_max; level++) {
      if (level == l_max) {
        // uberResidual[l_max] is read only by this copy until the next
        // computeAMRResidual rewrites it, so the buffers trade places instead;
        // m_correction[l_max] is first written by its relaxation from zero
        need(m_residual[level]->lay, uberResidual[level]->lay, "AMRVCycle");
        std::swap(m_residual[level]->d, uberResidual[level]->d);
        continue;
      }
      m_op[level]->assignLocal(*m_residual[level], *uberResidual[level]);
      m_op[level]->setToZero(*m_correction[level]);
    }
  }
  // fromZero: the correction this call relaxes first is zero on entry (set at the
  // top of the cycle or by the finer level just before recursing; only the first
  // of numMG recursive calls sees it zero)
  if (l_max == l_base) {
    // uberCorrection is zero on entry: zeroed before the first iteration and after each update
    if (fromZero) mg(l_base)->oneCycleFromZero(*uberCorrection[ilev], *uberResidual[ilev]);
    else mg(l_base)->oneCycle(*uberCorrection[ilev], *uberResidual[ilev]);
  } else if (ilev == l_base) {
    if (fromZero) mg(l_base)->oneCycleFromZero(*m_correction[ilev], *m_residual[ilev]);
    else mg(l_base)->oneCycle(*m_correction[ilev], *m_residual[ilev]);
    // uberCorrection[l_base] is zero here and read next only by the update of phi
    if (!m_foldPhi.empty()) m_op[ilev]->incrPhiFromZero(*m_foldPhi[ilev], *m_correction[ilev]);
    else m_op[ilev]->incr(*uberCorrection[ilev], *m_correction[ilev], 1.0);
  } else {
    relax(*m_correction[ilev], *m_residual[ilev], ilev, m_pre, fromZero);
    // on the first visit of a cycle uberCorrection[ilev] is zero (zeroed before the
    // cycle and not written since), so the incr need not read it
    if (fromZero) m_op[ilev]->incrFromZero(*uberCorrection[ilev], *m_correction[ilev]);
    else m_op[ilev]->incr(*uberCorrection[ilev], *m_correction[ilev], 1.0);
    m_op[ilev - 1]->setToZero(*m_correction[ilev - 1]);
    std::vector<T*> res(m_residual.size());
    for (size_t q = 0; q < res.size(); q++) res[q] = m_residual[q].get();
    computeAMRResidualLevel(res, uberCorrection, uberResidual, l_max, l_base, ilev - 1, true);
    m_op[ilev]->AMRRestrictS(*m_resC[ilev], *m_residual[ilev], *m_correction[ilev], *m_correction[ilev - 1],
                             *uberCorrection[ilev]);
    m_op[ilev - 1]->assignCopier(*m_residual[ilev - 1], *m_resC[ilev]);
    for (int img = 0; img < m_numMG; img++)
      AMRVCycle(uberCorrection, uberResidual, ilev - 1, l_max, l_base, img == 0);
    m_op[ilev]->AMRProlongS(*m_correction[ilev], *m_correction[ilev - 1], *m_resC[ilev]);
    m_op[ilev]->AMRUpdateResidual(*m_residual[ilev], *m_correction[ilev], *m_correction[ilev - 1]);
    T& dCorr = *uberCorrection[ilev];
    relax(dCorr, *m_residual[ilev], ilev, m_post, true);  // setToZero(dCorr) + relax
    // incr(m_correction, dCorr, 1.0) then assignLocal(uberCorrection, m_correction),
    // in one pass over the valid cells (every later read of a ghost refills it);
    // with the phi update folded in, uberCorrection[ilev] is read next only by
    // solveNoInitResid's incr(phi) and setToZero, which this pass performs too
    if (!m_foldPhi.empty()) m_op[ilev]->incrAssignUpdate(*m_correction[ilev], dCorr, *m_foldPhi[ilev]);
    else m_op[ilev]->incrAssign(*m_correction[ilev], dCorr);
  }
}
// solveNoInitResid (AMRMultiGrid.H:935-1124)
template <class Real>
void AMRMultiGrid<Real>::solveNoInitResid(std::vector<T*>& phi, std::vector<T*>& uberResidual,
                                          std::vector<T*>& uberCorrection, std::vector<T*>& rhs, int l_max,
                                          int l_base, bool zeroPhi) {
  if (l_base < 0 || l_base > l_max || l_max >= (int)m_op.size()) throw std::runtime_error("bad l_base / l_max");
  mg(l_base)->setBottomSolver(m_bottomSolver);
  // with l_base > 0 the residual of level l_base takes coarse-fine data from the
  // uberCorrection one level lower, which stays zero (AMRMultiGrid.H:952-967)
  if (l_base > 0) m_op[l_base - 1]->setToZero(*uberCorrection[l_base - 1]);
  for (int ilev = l_base; ilev <= l_max; ilev++) {
    m_op[ilev]->setToZero(*uberResidual[ilev]);
    m_op[ilev]->setToZero(*uberCorrection[ilev]);
  }
  if (zeroPhi)
    for (int ilev = l_base; ilev <= l_max; ilev++) m_op[ilev]->setToZero(*phi[ilev]);
  m_initial_rnorm = computeAMRResidual(uberResidual, phi, rhs, l_max, l_base, false, true);
  double rnorm = m_initial_rnorm, norm_last = 2 * m_initial_rnorm;
  m_bottomSolver->setConvergenceMetrics(m_initial_rnorm, m_bottomSolverEpsCushion * m_eps);
  if (m_innerBottomMetrics) m_innerBottomMetrics(m_initial_rnorm, m_bottomSolverEpsCushion * m_eps);
  int iter = 0;
  m_history.assign(1, rnorm);
  bool goNorm = rnorm > m_normThresh, goRedu = rnorm > m_eps * m_initial_rnorm, goIter = iter < m_iterMax;
  bool goHang = iter < m_imin || rnorm < (1 - m_hang) * norm_last, goMin = iter < m_iterMin;
  while (goMin || (goIter && goRedu && goHang && goNorm)) {
    norm_last = rnorm;
    if (m_innerCycleUpdate) {
      m_innerCycleUpdate(phi, uberResidual, l_max, l_base);  // uberCorrection stays zero
    } else if (m_numMG == 1 && l_max > l_base) {
      // every level is visited once per cycle: its correction goes into phi at the
      // end of its upsweep, with the arithmetic of the update below
      m_foldPhi.assign(phi.begin(), phi.end());
      try {
        AMRVCycle(uberCorrection, uberResidual, l_max, l_max, l_base);
      } catch (...) {
        m_foldPhi.clear();
        throw;
      }
      m_foldPhi.clear();
    } else {
      AMRVCycle(uberCorrection, uberResidual, l_max, l_max, l_base);
      for (int ilev = l_base; ilev <= l_max; ilev++) {
        m_op[ilev]->incr(*phi[ilev], *uberCorrection[ilev], 1.0);
        m_op[ilev]->setToZero(*uberCorrection[ilev]);
      }
    }
    rnorm = computeAMRResidual(uberResidual, phi, rhs, l_max, l_base, false, true);
    iter++;
    m_history.push_back(rnorm);
    goNorm = rnorm > m_normThresh;
    goRedu = rnorm > m_eps * m_initial_rnorm;
    goIter = iter < m_iterMax;
    goHang = iter < m_imin || rnorm < (1 - m_hang) * norm_last;
    goMin = iter < m_iterMin;
  }
  m_iterations = iter;
  m_exitStatus = int(!goRedu) + int(!goIter) * 2 + int(!goHang) * 4 + int(!goNorm) * 8;
}

template void setLevelOpBCValues<double>(AMRPoissonOp<double>*, BCValueFn, void*);
template void setLevelOpBCValues<float>(AMRPoissonOp<float>*, BCValueFn, void*);
template void defineLevelOp<double>(AMRPoissonOp<double>*, Context*, const DeviceLayout*, const Box&, double, double,
                                    double, double, const BCSpec&, const DeviceLayout*);
template void defineLevelOp<float>(AMRPoissonOp<float>*, Context*, const DeviceLayout*, const Box&, double, double,
                                   double, double, const BCSpec&, const DeviceLayout*);
template void linkCoarser<double>(AMRPoissonOp<double>*, AMRPoissonOp<double>*, const DeviceLayout*, int);
template void linkCoarser<float>(AMRPoissonOp<float>*, AMRPoissonOp<float>*, const DeviceLayout*, int);
template class AMRPoissonOp<double>;
template class AMRPoissonOp<float>;
template class AMRPoissonOpFactory<double>;
template class AMRPoissonOpFactory<float>;
template class AMRMultiGrid<double>;
template class AMRMultiGrid<float>;

}  // namespace amrmg
