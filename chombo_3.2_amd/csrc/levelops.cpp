// Operator-level C ABI: device-resident levels (DisjointBoxLayout + ProblemDomain),
// fields (LevelData<FArrayBox>), operators (AMRPoissonOp) and copiers, one entry
// point per LinearOp / MGLevelOp / AMRLevelOp virtual (declarations and conventions
// in include/amrmg.h).  Each call maps onto the same AMRPoissonOp<double> methods the
// whole-solve entry points run, so a host operator class built on these calls gets
// bitwise the same arithmetic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "../../include/amrmg.h"
#include "handles.hpp"

using namespace amrmg;

namespace {

Box box6(const int* p) { return Box(IV{p[0], p[1], p[2]}, IV{p[3], p[4], p[5]}); }

DBL parse_boxes(const int* boxes, int nbox) {
  if (nbox <= 0) throw std::runtime_error("a level needs at least one box");
  DBL g;
  for (int q = 0; q < nbox; q++) {
    g.boxes.push_back(box6(boxes + 6 * q));
    if (g.boxes.back().empty()) throw std::runtime_error("empty box");
  }
  if (!g.sorted()) throw std::runtime_error("boxes of a level must be sorted (Box::operator<)");
  for (int a = 0; a < nbox; a++)
    for (int b = a + 1; b < nbox; b++)
      if (intersects(g[a], g[b])) throw std::runtime_error("boxes of a level must be disjoint");
  return g;
}

amrmg_level* new_level(amrmg_ctx* ctx, const DBL& g, const Box& domain, bool merge) {
  for (const Box& b : g.boxes)
    if (!domain.contains(b)) throw std::runtime_error("a box lies outside the level's domain");
  std::unique_ptr<amrmg_level> L(new amrmg_level());
  L->ctx = ctx;
  L->lay.reset(new DeviceLayout(ctx->c.get(), g, domain, merge));
  return L.release();
}

const DeviceLayout& lay(const amrmg_field* f) {
  if (!f) throw std::runtime_error("null field");
  return *f->lev->lay;
}
LevelData<double>& ld(amrmg_field* f) {
  if (!f) throw std::runtime_error("null field");
  return f->ld;
}
const LevelData<double>& ld(const amrmg_field* f) {
  if (!f) throw std::runtime_error("null field");
  return f->ld;
}
void same_level(const amrmg_field* a, const amrmg_field* b, const char* what) {
  if (!a || !b) throw std::runtime_error(std::string("null field in ") + what);
  if (a->lev != b->lev) throw std::runtime_error(std::string("fields on different levels in ") + what);
  if (a->ld.nc != b->ld.nc) throw std::runtime_error(std::string("fields of different component counts in ") + what);
}
// component c of a field (its one-component storage)
double* comp(amrmg_field* f, int c) { return f->ld.d + (long long)c * f->ld.cs; }
const double* comp(const amrmg_field* f, int c) { return f->ld.d + (long long)c * f->ld.cs; }
constexpr int kMaxFieldComps = 16;
// an optional operand: NULL is the reference's undefined LevelData
LevelData<double>& opt(amrmg_field* f, LevelData<double>& none) { return f ? f->ld : none; }
AMRPoissonOp<double>& aop(amrmg_op* op) {
  if (!op || !op->op) throw std::runtime_error("null operator");
  return *op->op;
}

// FArrayBox-order transfers between host buffers (box by box, each box's valid box or
// FAB in Fortran order) and a level's slab.
void pack_transfer(const DeviceLayout& L, double* dev, double* host, bool ghosted, bool up, int nc = 1,
                   long long cs = 0) {
  // the whole field goes through the host: an upload keeps the cells it does not write.
  // Box by box, each box component by component (an FArrayBox's data order).
  if (nc == 1) cs = L.slab;
  const size_t n = (size_t)(cs * (nc - 1) + L.slab);
  std::vector<double> slab(n, 0.0);
  L.ctx->sync();
  hip_check(hipMemcpy(slab.data(), dev, sizeof(double) * n, hipMemcpyDeviceToHost), "read field");
  long long o = 0;
  for (int b = 0; b < L.dbl.size(); b++)
  for (int m = 0; m < nc; m++) {
    const Box& v = L.dbl[b];
    const Box r = ghosted ? grow(v, 1) : v;
    for (int k = r.lo[2]; k <= r.hi[2]; k++)
      for (int j = r.lo[1]; j <= r.hi[1]; j++)
        for (int i = r.lo[0]; i <= r.hi[0]; i++, o++) {
          const long long c = L.index(b, i, j, k) + m * cs;
          if (!up) {
            host[o] = slab[(size_t)c];
            continue;
          }
          if (L.merged && !v.contains(i, j, k)) {
            // ghost cells of a merged level: only face ghosts outside the rectangle are
            // this box's own cells (the others are valid cells of other boxes, or edge
            // and corner cells shared with other boxes that the stencils never read)
            const int ox = (i < v.lo[0] || i > v.hi[0]) ? 1 : 0;
            const int oy = (j < v.lo[1] || j > v.hi[1]) ? 1 : 0;
            const int oz = (k < v.lo[2] || k > v.hi[2]) ? 1 : 0;
            if (ox + oy + oz != 1 || L.rect.contains(i, j, k)) continue;
          }
          slab[(size_t)c] = host[o];
        }
  }
  if (up) {
    hip_check(hipMemcpyAsync(dev, slab.data(), sizeof(double) * n, hipMemcpyHostToDevice, L.ctx->stream),
              "write field");
    L.ctx->sync();
  }
}

ops::MotionPlan motion_items(const int* items, int n, const DBL& src, const DBL& dst, bool exchange) {
  ops::MotionPlan P;
  for (int q = 0; q < n; q++) {
    const int* m = items + 8 * q;
    Motion x;
    x.from = m[0];
    x.to = m[1];
    for (int d = 0; d < 3; d++) {
      x.lo[d] = m[2 + d];
      x.hi[d] = m[5 + d];
    }
    if (x.from < 0 || x.from >= src.size() || x.to < 0 || x.to >= dst.size())
      throw std::runtime_error("motion item: box index out of range");
    const Box r(IV{x.lo[0], x.lo[1], x.lo[2]}, IV{x.hi[0], x.hi[1], x.hi[2]});
    if (r.empty()) throw std::runtime_error("motion item: empty region");
    if (!src[x.from].contains(r)) throw std::runtime_error("motion item: region outside the from box's valid cells");
    if (!grow(dst[x.to], 1).contains(r)) throw std::runtime_error("motion item: region outside the to box's FAB");
    if (exchange && dst[x.to].contains(r)) throw std::runtime_error("exchange item writes valid cells of its box");
    P.h_items.push_back(x);
  }
  return P;
}

}  // namespace

extern "C" {

int amrmg_ctx_sync(amrmg_ctx* ctx) {
  API_BEGIN
  ctx->c->sync();
  API_END
}

// ---------------- levels ----------------
int amrmg_level_register(amrmg_ctx* ctx, const int* boxes, int nbox, const int* domain, int merge,
                         amrmg_level** out) {
  API_BEGIN
  DBL g = parse_boxes(boxes, nbox);
  *out = new_level(ctx, g, box6(domain), merge != 0);
  API_END
}
int amrmg_level_destroy(amrmg_level* L) {
  API_BEGIN
  if (!L) return 0;
  if (L->ownedByOp) throw std::runtime_error("this level belongs to an operator (destroyed with it)");
  if (L->refs) throw std::runtime_error("fields or operators are still defined on this level");
  delete L;
  API_END
}
int amrmg_level_coarsen(amrmg_level* L, amrmg_level** out) {
  API_BEGIN
  const DeviceLayout& D = *L->lay;
  if (!D.dbl.coarsenable(2)) throw std::runtime_error("level not coarsenable by 2");
  *out = new_level(L->ctx, D.dbl.coarsened(2), coarsen(D.domain, 2), D.merged);
  API_END
}
long long amrmg_level_cells(amrmg_level* L, int ghosted) {
  if (!L) return -1;
  if (!ghosted) return L->lay->valid_cells;
  long long n = 0;
  for (const Box& b : L->lay->dbl.boxes) n += grow(b, 1).npts();
  return n;
}

// ---------------- fields ----------------
int amrmg_field_create(amrmg_level* L, amrmg_field** out) { return amrmg_field_create_ncomp(L, 1, out); }
int amrmg_field_create_ncomp(amrmg_level* L, int ncomp, amrmg_field** out) {
  API_BEGIN
  if (ncomp < 1 || ncomp > kMaxFieldComps) throw std::runtime_error("ncomp must be 1..16");
  std::unique_ptr<amrmg_field> f(new amrmg_field());
  f->lev = L;
  f->ld.define(L->lay.get(), ncomp);
  hip_check(hipMemsetAsync(f->ld.d, 0, sizeof(double) * f->ld.cs * ncomp, L->ctx->c->stream), "zero field");
  L->refs++;
  *out = f.release();
  API_END
}
int amrmg_field_ncomp(const amrmg_field* f) { return f ? f->ld.nc : -1; }
int amrmg_field_destroy(amrmg_field* f) {
  API_BEGIN
  if (!f) return 0;
  f->lev->ctx->c->sync();
  f->lev->refs--;
  delete f;
  API_END
}
int amrmg_field_upload(amrmg_field* f, const double* host, int ghosted) {
  API_BEGIN
  pack_transfer(lay(f), f->ld.d, const_cast<double*>(host), ghosted != 0, true, f->ld.nc, f->ld.cs);
  API_END
}
int amrmg_field_download(amrmg_field* f, double* host, int ghosted) {
  API_BEGIN
  pack_transfer(lay(f), f->ld.d, host, ghosted != 0, false, f->ld.nc, f->ld.cs);
  API_END
}
int amrmg_field_assign(amrmg_field* lhs, const amrmg_field* rhs) {
  API_BEGIN
  same_level(lhs, rhs, "assign");
  for (int c = 0; c < lhs->ld.nc; c++) ops::copy_all<double>(lay(lhs), comp(lhs, c), comp(rhs, c));
  API_END
}
int amrmg_field_incr(amrmg_field* lhs, const amrmg_field* x, double scale) {
  API_BEGIN
  same_level(lhs, x, "incr");
  for (int c = 0; c < lhs->ld.nc; c++) ops::incr<double>(lay(lhs), comp(lhs, c), comp(x, c), scale);
  API_END
}
int amrmg_field_axby(amrmg_field* lhs, const amrmg_field* x, const amrmg_field* y, double a, double b) {
  API_BEGIN
  same_level(lhs, x, "axby");
  same_level(lhs, y, "axby");
  for (int c = 0; c < lhs->ld.nc; c++) ops::axby<double>(lay(lhs), comp(lhs, c), comp(x, c), comp(y, c), a, b);
  API_END
}
int amrmg_field_scale(amrmg_field* x, double s) {
  API_BEGIN
  for (int c = 0; c < ld(x).nc; c++) ops::scale<double>(lay(x), comp(x, c), s);
  API_END
}
int amrmg_field_setval(amrmg_field* x, double v) {
  API_BEGIN
  for (int c = 0; c < ld(x).nc; c++) ops::setval<double>(lay(x), comp(x, c), v);
  API_END
}
int amrmg_field_dot(const amrmg_field* a, const amrmg_field* b, double* out) {
  API_BEGIN
  same_level(a, b, "dot");
  *out = ops::dot_ordered<double>(lay(a), a->ld.d, b->ld.d, 1, a->ld.nc, a->ld.cs);
  API_END
}
int amrmg_field_dot_noalloc(const amrmg_field* a, const amrmg_field* b, double* out) {
  API_BEGIN
  same_level(a, b, "dot");
  *out = ops::dot_ordered_noalloc<double>(lay(a), a->ld.d, b->ld.d, 1, a->ld.nc, a->ld.cs);
  API_END
}
int amrmg_field_norm(const amrmg_field* x, int p, double* out) {
  API_BEGIN
  const DeviceLayout& L = lay(x);
  if (p == 0) *out = ops::maxnorm<double>(L, x->ld.d, 2, x->ld.nc, x->ld.cs);
  else if (p == 1) *out = ops::norm1_ordered<double>(L, x->ld.d, 3, x->ld.nc, x->ld.cs);
  else if (p == 2) *out = ops::norm2_ordered<double>(L, x->ld.d, 3, x->ld.nc, x->ld.cs);
  else throw std::runtime_error("norm: p must be 0, 1 or 2");
  API_END
}

// ---------------- operators ----------------
int amrmg_op_create(amrmg_level* L, amrmg_op* coarser, double dx, double dxCrse, double alpha, double beta,
                    const int* bc_type, const double* bc_val, amrmg_op** out) {
  API_BEGIN
  if (!(dx > 0)) throw std::runtime_error("dx must be positive");
  BCSpec bc{};
  for (int s = 0; s < 2; s++)
    for (int d = 0; d < 3; d++) {
      bc.type[s][d] = bc_type[3 * s + d];
      bc.val[s][d] = bc_val[3 * s + d];
      if (bc.type[s][d] < 0 || bc.type[s][d] > 2) throw std::runtime_error("unknown BC type");
    }
  std::unique_ptr<amrmg_op> h(new amrmg_op());
  h->ctx = L->ctx;
  h->lev = L;
  h->op.reset(new AMRPoissonOp<double>());
  Context* c = L->ctx->c.get();
  const DeviceLayout* mgc = nullptr;
  int ratio = 2;
  if (coarser) {
    if (coarser->finer) throw std::runtime_error("the coarser operator already has a finer level");
    if (coarser->ctx != L->ctx) throw std::runtime_error("operators of different contexts");
    if (!(dxCrse > 0)) throw std::runtime_error("a level with a coarser AMR level needs dxCrse > 0");
    // the refinement ratio to the coarser level, dxCrse / dx: 2 or 4
    ratio = (int)std::lround(dxCrse / dx);
    if ((ratio != 2 && ratio != 4) || std::fabs(dxCrse - ratio * dx) > 1e-12 * dxCrse)
      throw std::runtime_error("dxCrse / dx must be the refinement ratio, 2 or 4");
    if (!L->lay->dbl.coarsenable(ratio))
      throw std::runtime_error("a level with a coarser AMR level must be coarsenable by the refinement ratio");
    h->cfLevel.reset(new amrmg_level());
    h->cfLevel->ctx = L->ctx;
    h->cfLevel->ownedByOp = true;
    h->cfLevel->lay.reset(
        new DeviceLayout(c, L->lay->dbl.coarsened(ratio), coarsen(L->lay->domain, ratio), L->lay->merged));
    if (ratio == 2) mgc = h->cfLevel->lay.get();  // coarsen(level, 2) is also its multigrid child
  }
  defineLevelOp<double>(h->op.get(), c, L->lay.get(), L->lay->domain, dx, dxCrse, alpha, beta, bc, mgc);
  if (coarser) {
    linkCoarser<double>(h->op.get(), coarser->op.get(), h->cfLevel->lay.get(), ratio);
    coarser->finer = h.get();
    h->coarser = coarser;
  }
  L->refs++;
  *out = h.release();
  API_END
}
int amrmg_op_create_coarser(amrmg_op* op, amrmg_op** out, amrmg_level** level_out) {
  API_BEGIN
  AMRPoissonOp<double>& p = aop(op);
  if (op->mgChild) throw std::runtime_error("this operator already has a multigrid child");
  const DeviceLayout& L = *op->lev->lay;
  if (!L.dbl.coarsenable(2)) throw std::runtime_error("level not coarsenable by 2");
  std::unique_ptr<amrmg_op> h(new amrmg_op());
  h->ctx = op->ctx;
  if (op->cfLevel && p.m_refToCoarser == 2) {
    // an AMR level with a coarser one at ratio 2 already holds coarsen(level, 2),
    // which its restrictResidual targets
    h->lev = op->cfLevel.get();
    h->lev->refs++;
  } else {
    h->ownLevel.reset(new amrmg_level());
    h->ownLevel->ctx = op->ctx;
    h->ownLevel->ownedByOp = true;
    h->ownLevel->lay.reset(new DeviceLayout(op->ctx->c.get(), L.dbl.coarsened(2), coarsen(L.domain, 2), L.merged));
    h->lev = h->ownLevel.get();
  }
  h->op.reset(new AMRPoissonOp<double>());
  h->op->m_ncomp = p.m_ncomp;
  // MGnewOp: 2 dx, the AMR level's dxCrse, the same coefficients and BC
  defineLevelOp<double>(h->op.get(), op->ctx->c.get(), h->lev->lay.get(), h->lev->lay->domain, 2 * p.m_dx, p.m_dxCrse,
                        p.m_alpha, p.m_beta, p.m_bc, nullptr);
  h->op->m_relaxMode = p.m_relaxMode;
  p.m_mgCoarse = h->lev->lay.get();
  h->mgParent = op;
  op->mgChild = h.get();
  if (level_out) *level_out = h->lev;
  *out = h.release();
  API_END
}
int amrmg_op_coarsened_level(amrmg_op* op, amrmg_level** out) {
  API_BEGIN
  if (!op->cfLevel) throw std::runtime_error("operator has no coarser AMR level");
  *out = op->cfLevel.get();
  API_END
}
int amrmg_op_destroy(amrmg_op* op) {
  API_BEGIN
  if (!op) return 0;
  if (op->finer) throw std::runtime_error("destroy the finer AMR level's operator first");
  if (op->mgChild) throw std::runtime_error("destroy the multigrid child operator first");
  if (op->ownLevel && op->ownLevel->refs) throw std::runtime_error("fields are still defined on this operator's level");
  if (op->cfLevel && op->cfLevel->refs) throw std::runtime_error("fields are still defined on this operator's coarsened level");
  op->ctx->c->sync();
  if (op->coarser) {  // the coarser operator loses its reflux / zeroCovered partner
    AMRPoissonOp<double>& c = *op->coarser->op;
    c.m_finerLay = nullptr;
    c.m_zeroPlan = nullptr;
    c.m_finerCfLay = nullptr;
    c.n_rcell = 0;
    op->coarser->finer = nullptr;
  }
  if (op->mgParent) {
    amrmg_op* par = op->mgParent;
    par->op->m_mgCoarse = par->cfLevel ? par->cfLevel->lay.get() : nullptr;
    par->mgChild = nullptr;
  }
  if (!op->ownLevel) op->lev->refs--;
  op->op.reset();
  delete op;
  API_END
}
int amrmg_op_set_alpha_beta(amrmg_op* op, double alpha, double beta) {
  API_BEGIN
  aop(op).m_alpha = alpha;
  aop(op).m_beta = beta;
  aop(op).resetLambda();
  API_END
}
int amrmg_op_set_relax_mode(amrmg_op* op, int mode) {
  API_BEGIN
  if (mode < 0 || mode > 5) throw std::runtime_error("relax mode must be 0..5");
  aop(op).m_relaxMode = mode;
  API_END
}

// ---- LinearOp / MGLevelOp ----
int amrmg_op_residual(amrmg_op* op, amrmg_field* lhs, amrmg_field* phi, const amrmg_field* rhs, int homogeneous) {
  API_BEGIN
  aop(op).residual(ld(lhs), ld(phi), ld(rhs), homogeneous != 0);
  API_END
}
int amrmg_op_apply(amrmg_op* op, amrmg_field* lhs, amrmg_field* phi, int homogeneous) {
  API_BEGIN
  aop(op).applyOp(ld(lhs), ld(phi), homogeneous != 0);
  API_END
}
int amrmg_op_precond(amrmg_op* op, amrmg_field* cor, const amrmg_field* residual) {
  API_BEGIN
  aop(op).preCond(ld(cor), ld(residual));
  API_END
}
int amrmg_op_relax(amrmg_op* op, amrmg_field* correction, const amrmg_field* residual, int iterations) {
  API_BEGIN
  if (iterations < 0) throw std::runtime_error("relax: negative iteration count");
  aop(op).relax(ld(correction), ld(residual), iterations);
  API_END
}
int amrmg_op_restrict_residual(amrmg_op* op, amrmg_field* res_coarse, amrmg_field* phi, const amrmg_field* rhs) {
  API_BEGIN
  if (!aop(op).m_mgCoarse) throw std::runtime_error("restrictResidual: operator has no coarser multigrid level");
  aop(op).restrictResidual(ld(res_coarse), ld(phi), ld(rhs));
  API_END
}
int amrmg_op_prolong_increment(amrmg_op* op, amrmg_field* phi, const amrmg_field* correct_coarse) {
  API_BEGIN
  if (!aop(op).m_mgCoarse) throw std::runtime_error("prolongIncrement: operator has no coarser multigrid level");
  aop(op).prolongIncrement(ld(phi), ld(correct_coarse));
  API_END
}

// ---- AMRLevelOp ----
static void need_finer(amrmg_op* op, amrmg_op* finer, amrmg_field* phiFine) {
  if (!phiFine) return;
  if (!finer || op->finer != finer) throw std::runtime_error("finer operator is not this operator's finer AMR level");
}
static void need_coarser(amrmg_op* op, amrmg_field* phiCoarse) {
  if (phiCoarse && !op->coarser) throw std::runtime_error("operator has no coarser AMR level");
}
int amrmg_op_amr_residual(amrmg_op* op, amrmg_field* res, amrmg_field* phi_fine, amrmg_field* phi,
                          amrmg_field* phi_coarse, const amrmg_field* rhs, int homogeneous, amrmg_op* finer) {
  API_BEGIN
  need_finer(op, finer, phi_fine);
  need_coarser(op, phi_coarse);
  LevelData<double> none1, none2;
  AMRPoissonOp<double>& p = aop(op);
  LevelData<double>& pf = opt(phi_fine, none1);
  LevelData<double>& pc = opt(phi_coarse, none2);
  const bool h = homogeneous != 0;
  if (phi_fine && phi_coarse) p.AMRResidual(ld(res), pf, ld(phi), pc, ld(rhs), h, finer->op.get());
  else if (phi_fine) p.AMRResidualNC(ld(res), pf, ld(phi), ld(rhs), h, finer->op.get());
  else if (phi_coarse) p.AMRResidualNF(ld(res), ld(phi), pc, ld(rhs), h);
  else p.residual(ld(res), ld(phi), ld(rhs), h);
  API_END
}
int amrmg_op_amr_operator(amrmg_op* op, amrmg_field* lphi, amrmg_field* phi_fine, amrmg_field* phi,
                          amrmg_field* phi_coarse, int homogeneous, amrmg_op* finer) {
  API_BEGIN
  need_finer(op, finer, phi_fine);
  need_coarser(op, phi_coarse);
  LevelData<double> none1, none2;
  AMRPoissonOp<double>& p = aop(op);
  LevelData<double>& pf = opt(phi_fine, none1);
  LevelData<double>& pc = opt(phi_coarse, none2);
  const bool h = homogeneous != 0;
  if (phi_fine && phi_coarse) p.AMROperator(ld(lphi), pf, ld(phi), pc, h, finer->op.get());
  else if (phi_fine) p.AMROperatorNC(ld(lphi), pf, ld(phi), h, finer->op.get());
  else if (phi_coarse) p.AMROperatorNF(ld(lphi), ld(phi), pc, h);
  else p.applyOp(ld(lphi), ld(phi), h);
  API_END
}
int amrmg_op_amr_restrict(amrmg_op* op, amrmg_field* res_coarse, const amrmg_field* residual,
                          amrmg_field* correction, amrmg_field* coarse_correction, amrmg_field* scratch) {
  API_BEGIN
  if (!op->coarser) throw std::runtime_error("AMRRestrictS: operator has no coarser AMR level");
  LevelData<double> none;
  aop(op).AMRRestrictS(ld(res_coarse), ld(residual), ld(correction), opt(coarse_correction, none), ld(scratch));
  API_END
}
int amrmg_op_amr_prolong(amrmg_op* op, amrmg_field* correction, const amrmg_field* coarse_correction,
                         amrmg_field* temp) {
  API_BEGIN
  if (!op->coarser) throw std::runtime_error("AMRProlongS: operator has no coarser AMR level");
  aop(op).AMRProlongS(ld(correction), ld(coarse_correction), ld(temp));
  API_END
}
int amrmg_op_amr_update_residual(amrmg_op* op, amrmg_field* residual, amrmg_field* correction,
                                 amrmg_field* coarse_correction) {
  API_BEGIN
  need_coarser(op, coarse_correction);
  LevelData<double> none;
  aop(op).AMRUpdateResidual(ld(residual), ld(correction), opt(coarse_correction, none));
  API_END
}
int amrmg_op_assign_copier(amrmg_op* op, amrmg_field* lhs, const amrmg_field* rhs) {
  API_BEGIN
  if (!op->finer) throw std::runtime_error("assignCopier: operator has no finer AMR level");
  aop(op).assignCopier(ld(lhs), ld(rhs));
  API_END
}
int amrmg_op_zero_covered(amrmg_op* op, amrmg_field* lhs) {
  API_BEGIN
  aop(op).zeroCovered(ld(lhs));
  API_END
}
int amrmg_op_max_norm(amrmg_op* op, const amrmg_field* x, double* out) {
  API_BEGIN
  *out = aop(op).localMaxNorm(ld(x));
  API_END
}
int amrmg_op_exchange(amrmg_op* op, amrmg_field* phi) {
  API_BEGIN
  AMRPoissonOp<double>& p = aop(op);
  if (phi->ld.lay != p.m_lay) throw std::runtime_error("exchange: field on another level");
  for (int c = 0; c < phi->ld.nc; c++) ops::fill_ghosts<double>(*p.m_lay, p.m_ghost, comp(phi, c), 3, 1, p.m_dx, p.m_dxCrse, c);
  API_END
}
int amrmg_op_bc(amrmg_op* op, amrmg_field* phi, int homogeneous) {
  API_BEGIN
  AMRPoissonOp<double>& p = aop(op);
  if (phi->ld.lay != p.m_lay) throw std::runtime_error("bc: field on another level");
  if (phi->ld.nc > p.m_ncomp && !homogeneous && p.m_ghost.d_bcv)
    throw std::runtime_error("bc: boundary values were taken for fewer components (amrmg_op_set_ncomp)");
  for (int c = 0; c < phi->ld.nc; c++)
    ops::fill_ghosts<double>(*p.m_lay, p.m_ghost, comp(phi, c), 4, homogeneous != 0, p.m_dx, p.m_dxCrse, c);
  API_END
}
int amrmg_op_set_bc_values(amrmg_op* op, amrmg_bc_value_fn fn, void* user) {
  API_BEGIN
  setLevelOpBCValues<double>(op->op.get(), fn, user);
  API_END
}
int amrmg_op_set_ncomp(amrmg_op* op, int ncomp) {
  API_BEGIN
  if (ncomp < 1 || ncomp > kMaxFieldComps) throw std::runtime_error("ncomp must be 1..16");
  AMRPoissonOp<double>& p = aop(op);
  p.m_ncomp = ncomp;
  if (p.m_bc.fn) setLevelOpBCValues<double>(&p, p.m_bc.fn, p.m_bc.user);  // one value block per component
  API_END
}
int amrmg_op_cf_homog(amrmg_op* op, amrmg_field* phi) {
  API_BEGIN
  AMRPoissonOp<double>& p = aop(op);
  if (phi->ld.lay != p.m_lay) throw std::runtime_error("homogeneousCFInterp: field on another level");
  for (int c = 0; c < phi->ld.nc; c++) ops::fill_ghosts<double>(*p.m_lay, p.m_ghost, comp(phi, c), 2, 1, p.m_dx, p.m_dxCrse, c);
  API_END
}
int amrmg_op_cf_inhomog(amrmg_op* op, amrmg_field* phi, const amrmg_field* phi_coarse) {
  API_BEGIN
  if (!op->coarser) throw std::runtime_error("coarseFineInterp: operator has no coarser AMR level");
  aop(op).coarseFineInterp(ld(phi), ld(phi_coarse));
  API_END
}
int amrmg_op_gsrb(amrmg_op* op, amrmg_field* phi, const amrmg_field* rhs, int color) {
  API_BEGIN
  AMRPoissonOp<double>& p = aop(op);
  if (phi->ld.lay != p.m_lay || rhs->ld.lay != p.m_lay) throw std::runtime_error("gsrb: field on another level");
  if (color != 0 && color != 1) throw std::runtime_error("gsrb: colour must be 0 or 1");
  if (phi->ld.nc != rhs->ld.nc) throw std::runtime_error("gsrb: fields of different component counts");
  for (int c = 0; c < phi->ld.nc; c++)
    ops::gsrb<double>(*p.m_lay, comp(phi, c), comp(rhs, c), color, p.m_dx, p.m_alpha, p.m_beta);
  API_END
}
int amrmg_op_reflux(amrmg_op* op, amrmg_field* phi_fine, amrmg_field* phi, amrmg_field* lphi, amrmg_op* finer) {
  API_BEGIN
  if (!finer || op->finer != finer) throw std::runtime_error("reflux: finer operator is not this operator's finer level");
  aop(op).reflux(ld(phi_fine), ld(phi), ld(lphi), finer->op.get());
  API_END
}

// ---------------- copiers ----------------
int amrmg_plan_copy(amrmg_level* src, amrmg_level* dst, const int* items, int nitems, amrmg_plan** out) {
  API_BEGIN
  if (src->ctx != dst->ctx) throw std::runtime_error("levels of different contexts");
  std::unique_ptr<amrmg_plan> p(new amrmg_plan());
  p->src = src;
  p->dst = dst;
  if (items) p->P = motion_items(items, nitems, src->lay->dbl, dst->lay->dbl, false);
  else p->P = build_copier(src->lay->dbl, dst->lay->dbl);
  p->P.upload(src->ctx->c.get());
  src->refs++;
  dst->refs++;
  *out = p.release();
  API_END
}
int amrmg_plan_exchange(amrmg_level* L, const int* items, int nitems, amrmg_plan** out) {
  API_BEGIN
  std::unique_ptr<amrmg_plan> p(new amrmg_plan());
  p->src = L;
  p->dst = L;
  if (items) {
    p->P = motion_items(items, nitems, L->lay->dbl, L->lay->dbl, true);
  } else {  // exchangeDefine + trimEdges: the face ghosts of each box that lie in another box
    BCSpec bc{};
    for (int s = 0; s < 2; s++)
      for (int d = 0; d < 3; d++) bc.type[s][d] = 1;
    const ops::GhostPlan G = build_ghost_plan(L->lay->dbl, L->lay->domain, bc, false, false);
    for (const Patch& q : G.h_patches) {
      if (q.kind != PK_EXCHANGE) continue;
      Motion m;
      m.from = q.src;
      m.to = q.dst;
      for (int d = 0; d < 3; d++) {
        m.lo[d] = q.lo[d];
        m.hi[d] = q.hi[d];
      }
      p->P.h_items.push_back(m);
    }
  }
  p->P.upload(L->ctx->c.get());
  L->refs += 2;
  *out = p.release();
  API_END
}
int amrmg_plan_exec(amrmg_plan* p, amrmg_field* dst, const amrmg_field* src, int op) {
  API_BEGIN
  if (!dst || dst->lev != p->dst) throw std::runtime_error("copier: destination field on another level");
  if (op == 0) {
    if (!src || src->lev != p->src) throw std::runtime_error("copier: source field on another level");
    if (src->ld.nc != dst->ld.nc) throw std::runtime_error("copier: fields of different component counts");
    for (int c = 0; c < dst->ld.nc && p->P.nitems; c++)
      ops::motion<double>(p->P, *p->src->lay, *p->dst->lay, comp(dst, c), comp(src, c), 0);
  } else if (op == 1) {
    for (int c = 0; c < dst->ld.nc && p->P.nitems; c++)
      ops::motion<double>(p->P, *p->src->lay, *p->dst->lay, comp(dst, c), (const double*)nullptr, 1);
  } else {
    throw std::runtime_error("copier op must be 0 (copyTo) or 1 (copyToZero)");
  }
  API_END
}
int amrmg_plan_destroy(amrmg_plan* p) {
  API_BEGIN
  if (!p) return 0;
  p->src->ctx->c->sync();
  p->P.release();
  p->src->refs--;
  p->dst->refs--;
  delete p;
  API_END
}

}  // extern "C"
