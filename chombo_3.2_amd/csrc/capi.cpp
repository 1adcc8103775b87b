// extern "C" boundary of libamrmg.so (declarations and conventions in include/amrmg.h).
#include "../../include/amrmg.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <stdexcept>
#include <string>
#include <vector>

#include "amrpoissonop.hpp"
#include "fused.hpp"
#include "batch.hpp"
#include "handles.hpp"
#include "mixed.hpp"

using namespace amrmg;

// components a solver's fields may have (amrmg_solver_create_ncomp)
constexpr int kMaxComps = 16;

thread_local std::string amrmg_g_err;

extern "C" const char* amrmg_last_error(void) { return amrmg_g_err.c_str(); }

// A batch runs in one or more lanes: each lane is a BatchSolver over a subset of the
// problems, and lanes after the first run on streams of their own, so their lockstep
// iterations overlap on the device (problem results do not depend on the lanes).
struct amrmg_batch {
  amrmg_ctx* ctx = nullptr;
  struct Lane {
    std::unique_ptr<Context> own;  // the stream of a lane after the first
    std::unique_ptr<BatchSolver> solver;
    std::vector<int> probs;        // caller indices of the lane's problems
  };
  std::vector<Lane> lanes;
  std::vector<std::pair<int, int>> at;  // caller index -> (lane, index in the lane)
  int nprob = 0;
  BatchSolver& solver(int q, int& local) {
    if (q < 0 || q >= nprob) throw std::runtime_error("batch: bad problem index");
    local = at[q].second;
    return *lanes[at[q].first].solver;
  }
};

struct amrmg_solver {
  amrmg_ctx* ctx = nullptr;
  int precision = 0;
  int ncomp = 1;  // components of phi, rhs and the residual
  DeviceHierarchy H;
  BCSpec bc{};
  double alpha = 0, beta = 1;
  AMRPoissonOpFactory<double> fac;
  AMRMultiGrid<double> amg;
  std::unique_ptr<DeviceBottomSolver<double>> bottom;
  std::vector<std::unique_ptr<LevelData<double>>> phi, rhs, res, uberRes, uberCorr;
  int sp[6] = {2, 2, 2, 1, 100, 5};
  double rp[3] = {1e-10, 1e-15, 1e-30};
  long long table_bytes = 0;
  std::unique_ptr<MixedInner> mixed;  // precision 1: FP32 inner cycles
  // variable coefficients per multigrid depth (empty: constant coefficients)
  std::vector<std::shared_ptr<VCCoefs<double>>> vc;
  std::vector<std::shared_ptr<VCCoefs<float>>> vc32;
  int vcHarmonic = 0;  // VCAMRPoissonOp2Factory's coefficient average type: 0 arithmetic, 1 harmonic
  // Per-box storage for the FAB-level entry points (amrmg_kernel_gsrb,
  // amrmg_homog_cf, amrmg_quad_cf), whose ghost cells are caller inputs like the
  // ghost cells of the reference's FArrayBoxes; built on first use.
  std::unique_ptr<DeviceHierarchy> fabH;
  std::unique_ptr<AMRPoissonOpFactory<double>> fabFac;
  std::vector<std::unique_ptr<AMRPoissonOp<double>>> fabOps;

  int nlev() const { return H.nlev(); }
  AMRPoissonOp<double>* fabOp(int lev) {
    if (!fabH) {
      fabH.reset(new DeviceHierarchy());
      fabH->define(ctx->c.get(), H.grids, H.domains[0], H.dxs[0], false, nullptr, H.ratios.data() + 1);
      fabFac.reset(new AMRPoissonOpFactory<double>());
      fabFac->define(fabH.get(), alpha, beta, bc);
    }
    while ((int)fabOps.size() <= lev) {
      const int l = (int)fabOps.size();
      fabOps.emplace_back(fabFac->AMRnewOp(l, l > 0 ? fabOps[l - 1].get() : nullptr));
    }
    return fabOps[lev].get();
  }
  std::vector<LevelData<double>*> ptrs(std::vector<std::unique_ptr<LevelData<double>>>& v) {
    std::vector<LevelData<double>*> p;
    for (auto& x : v) p.push_back(x.get());
    return p;
  }
  AMRPoissonOp<double>* mgop(int depth) {
    if (depth < 0 || depth >= H.mgDepth()) throw std::runtime_error("bad multigrid depth");
    if (depth == 0) return amg.m_op[0].get();
    return static_cast<AMRPoissonOp<double>*>(amg.m_mg0->m_op[depth]);
  }
  const DeviceLayout* mglay(int depth) { return depth == 0 ? H.amrLay[0].get() : H.mgLay[depth - 1].get(); }
  AMRPoissonOp<float>* mixedop(int depth) {
    if (depth == 0) return mixed->amg.m_op[0].get();
    return static_cast<AMRPoissonOp<float>*>(mixed->amg.m_mg0->m_op[depth]);
  }
};

// ---- host <-> device transfers of valid-packed / ghosted-packed data ----
// Valid-packed host order: box by box, in each box component by component (an
// FArrayBox's data order), cells x fastest.  On the device component c is the field
// at c * cs.
static void upload_valid(const DeviceLayout& L, double* d, const double* h, int nc = 1, long long cs = 0) {
  if (nc == 1) cs = L.slab;
  const size_t n = (size_t)(cs * (nc - 1) + L.slab);
  std::vector<double> slab(n, 0.0);
  long long o = 0;
  for (int b = 0; b < L.dbl.size(); b++) {
    const Box& bx = L.dbl[b];
    for (int m = 0; m < nc; m++)
      for (int k = bx.lo[2]; k <= bx.hi[2]; k++)
        for (int j = bx.lo[1]; j <= bx.hi[1]; j++) {
          long long c = L.index(b, bx.lo[0], j, k) + m * cs;
          std::memcpy(&slab[(size_t)c], h + o, sizeof(double) * bx.len(0));
          o += bx.len(0);
        }
  }
  hip_check(hipMemcpyAsync(d, slab.data(), sizeof(double) * n, hipMemcpyHostToDevice, L.ctx->stream), "upload");
  L.ctx->sync();
}
static void download_valid(const DeviceLayout& L, const double* d, double* h, int nc = 1, long long cs = 0) {
  if (nc == 1) cs = L.slab;
  const size_t n = (size_t)(cs * (nc - 1) + L.slab);
  std::vector<double> slab(n);
  L.ctx->sync();
  hip_check(hipMemcpy(slab.data(), d, sizeof(double) * n, hipMemcpyDeviceToHost), "download");
  long long o = 0;
  for (int b = 0; b < L.dbl.size(); b++) {
    const Box& bx = L.dbl[b];
    for (int m = 0; m < nc; m++)
      for (int k = bx.lo[2]; k <= bx.hi[2]; k++)
        for (int j = bx.lo[1]; j <= bx.hi[1]; j++) {
          long long c = L.index(b, bx.lo[0], j, k) + m * cs;
          std::memcpy(h + o, &slab[(size_t)c], sizeof(double) * bx.len(0));
          o += bx.len(0);
        }
  }
}
static void upload_ghosted(const DeviceLayout& L, double* d, const double* h) {
  hip_check(hipMemcpyAsync(d, h, sizeof(double) * L.slab, hipMemcpyHostToDevice, L.ctx->stream), "upload");
  L.ctx->sync();
}
static void download_ghosted(const DeviceLayout& L, const double* d, double* h) {
  L.ctx->sync();
  hip_check(hipMemcpy(h, d, sizeof(double) * L.slab, hipMemcpyDeviceToHost), "download");
}

static std::vector<DBL> parse_spec(const int* is, const double* rs, Box& dom, BCSpec& bc, double& dx0, double& a,
                                   double& b) {
  int nlev = is[0];
  if (nlev < 1 || nlev > 16) throw std::runtime_error("nlev out of range");
  dom = Box(IV{is[1], is[2], is[3]}, IV{is[4], is[5], is[6]});
  for (int d = 0; d < 3; d++) {
    bc.type[0][d] = is[7 + d];
    bc.type[1][d] = is[10 + d];
    bc.val[0][d] = rs[3 + d];
    bc.val[1][d] = rs[6 + d];
    for (int s = 0; s < 2; s++)
      if (bc.type[s][d] < 0 || bc.type[s][d] > 3) throw std::runtime_error("unknown BC type");
  }
  dx0 = rs[0];
  a = rs[1];
  b = rs[2];
  std::vector<DBL> g(nlev);
  const int* bp = is + 13 + nlev;
  for (int l = 0; l < nlev; l++) {
    int n = is[13 + l];
    if (n <= 0) throw std::runtime_error("a level has no boxes");
    for (int q = 0; q < n; q++) {
      g[l].boxes.push_back(Box(IV{bp[0], bp[1], bp[2]}, IV{bp[3], bp[4], bp[5]}));
      if (g[l].boxes.back().empty()) throw std::runtime_error("empty box");
      bp += 6;
    }
  }
  return g;
}

// Bytes of the fields a solver holds on a hierarchy.  FP64 per AMR level: phi, rhs,
// res, uber residual / correction, V-cycle correction, residual, resC and the
// relaxation buffer; per multigrid level: correction, residual, relaxation buffer.
// Mixed precision adds the FP32 copies of the V-cycle fields.
static long long hierarchy_field_bytes(const DeviceHierarchy& H, int precision) {
  long long n = 0, mg = 0;
  for (auto& l : H.amrLay) n += l->slab;
  for (auto& m : H.mgLay) mg += m->slab;
  long long total = n * 9 * 8 + mg * 3 * 8;
  if (precision == 1) total += n * 6 * 4 + mg * 3 * 4;
  return total;
}

extern "C" {

int amrmg_ctx_create(int device, amrmg_ctx** out) {
  API_BEGIN
  amrmg_ctx* c = new amrmg_ctx();
  c->c.reset(new Context(device));
  *out = c;
  API_END
}
int amrmg_ctx_set_fused_relax(amrmg_ctx* ctx, int on) {
  API_BEGIN
  ctx->c->fused_relax = (on != 0);
  API_END
}
int amrmg_ctx_destroy(amrmg_ctx* ctx) {
  API_BEGIN
  delete ctx;
  API_END
}

int amrmg_solver_create(amrmg_ctx* ctx, const int* ispec, const double* rspec, int precision, amrmg_solver** out) {
  return amrmg_solver_create_ratios(ctx, ispec, rspec, nullptr, precision, out);
}
int amrmg_solver_create_ratios(amrmg_ctx* ctx, const int* ispec, const double* rspec, const int* ref_ratios,
                               int precision, amrmg_solver** out) {
  return amrmg_solver_create_ncomp(ctx, ispec, rspec, ref_ratios, 1, precision, out);
}
int amrmg_solver_create_ncomp(amrmg_ctx* ctx, const int* ispec, const double* rspec, const int* ref_ratios,
                              int ncomp, int precision, amrmg_solver** out) {
  API_BEGIN
  if (precision != 0 && precision != 1) throw std::runtime_error("precision must be 0 (FP64) or 1 (mixed)");
  if (ncomp < 1 || ncomp > kMaxComps) throw std::runtime_error("ncomp must be 1.." + std::to_string(kMaxComps));
  if (ncomp > 1 && precision != 0) throw std::runtime_error("ncomp > 1 is not supported in mixed precision");
  std::unique_ptr<amrmg_solver> s(new amrmg_solver());
  s->ctx = ctx;
  s->precision = precision;
  s->ncomp = ncomp;
  Box dom;
  double dx0;
  std::vector<DBL> g = parse_spec(ispec, rspec, dom, s->bc, dx0, s->alpha, s->beta);
  s->H.define(ctx->c.get(), g, dom, dx0, true, &s->bc, ref_ratios);
  s->fac.define(&s->H, s->alpha, s->beta, s->bc, ncomp);
  s->bottom.reset(new DeviceBottomSolver<double>(ctx->c.get()));
  s->amg.define(s->fac, s->bottom.get(), s->nlev());
  s->amg.setSolverParameters(s->sp[0], s->sp[1], s->sp[2], s->sp[3], s->sp[4], s->rp[0], s->rp[1], s->rp[2]);
  s->amg.m_imin = s->sp[5];
  for (int l = 0; l < s->nlev(); l++) {
    const DeviceLayout* L = s->H.amrLay[l].get();
    s->phi.emplace_back(new LevelData<double>(L, ncomp));
    s->rhs.emplace_back(new LevelData<double>(L, ncomp));
    s->res.emplace_back(new LevelData<double>(L, ncomp));
    s->uberRes.emplace_back(new LevelData<double>(L, ncomp));
    s->uberCorr.emplace_back(new LevelData<double>(L, ncomp));
  }
  auto ph = s->ptrs(s->phi), rh = s->ptrs(s->rhs);
  s->amg.init(ph, rh, s->nlev() - 1, 0);
  if (precision == 1) {
    s->mixed.reset(new MixedInner());
    s->mixed->define(&s->H, s->alpha, s->beta, s->bc, s->amg);
    s->mixed->setSolverParameters(s->sp[0], s->sp[1], s->sp[2], s->sp[3], s->sp[4], s->rp[0], s->rp[1], s->rp[2]);
  }
  ctx->c->sync();
  *out = s.release();
  API_END
}
int amrmg_solver_destroy(amrmg_solver* s) {
  API_BEGIN
  delete s;
  API_END
}
int amrmg_solver_set_params(amrmg_solver* s, const int* sp, const double* rp) {
  API_BEGIN
  std::memcpy(s->sp, sp, sizeof s->sp);
  std::memcpy(s->rp, rp, sizeof s->rp);
  s->amg.setSolverParameters(sp[0], sp[1], sp[2], sp[3], sp[4], rp[0], rp[1], rp[2]);
  s->amg.m_imin = sp[5];
  if (s->mixed) s->mixed->setSolverParameters(sp[0], sp[1], sp[2], sp[3], sp[4], rp[0], rp[1], rp[2]);
  API_END
}
int amrmg_solver_set_alpha_beta(amrmg_solver* s, double alpha, double beta) {
  API_BEGIN
  s->alpha = alpha;
  s->beta = beta;
  auto set = [&](auto* op) {
    op->m_alpha = alpha;
    op->m_beta = beta;
  };
  for (auto& op : s->amg.m_op) set(op.get());
  for (int d = 1; d < s->H.mgDepth(); d++) set(s->mgop(d));
  if (s->mixed) {
    for (auto& op : s->mixed->amg.m_op) set(op.get());
    for (int d = 1; d < s->H.mgDepth(); d++) set(static_cast<AMRPoissonOp<float>*>(s->mixed->amg.m_mg0->m_op[d]));
  }
  s->fabOps.clear();
  s->fabFac.reset();
  s->fabH.reset();
  if (!s->vc.empty()) {
    for (auto& op : s->amg.m_op) op->resetLambda();
    for (int d = 1; d < s->H.mgDepth(); d++) s->mgop(d)->resetLambda();
    if (s->mixed) {
      for (auto& op : s->mixed->amg.m_op) op->resetLambda();
      for (int d = 1; d < s->H.mgDepth(); d++) s->mixedop(d)->resetLambda();
    }
  }
  API_END
}
// DiriBC / NeumBC with a BCValueHolder (BCFunc.cpp:217-258, 291-354) on every
// operator of the solver: the AMR levels and multigrid depths (FP64 and the FP32
// inner ones); the values are taken once, here, at every domain face cell.
int amrmg_solver_set_bc_values(amrmg_solver* s, amrmg_bc_value_fn fn, void* user) {
  API_BEGIN
  s->bc.fn = fn;
  s->bc.user = user;
  s->fac.bc.fn = fn;
  s->fac.bc.user = user;
  for (auto& op : s->amg.m_op) setLevelOpBCValues(op.get(), fn, user);
  for (int d = 1; d < s->H.mgDepth(); d++) setLevelOpBCValues(s->mgop(d), fn, user);
  if (s->mixed) {
    s->mixed->fac.bc.fn = fn;
    s->mixed->fac.bc.user = user;
    for (auto& op : s->mixed->amg.m_op) setLevelOpBCValues(op.get(), fn, user);
    for (int d = 1; d < s->H.mgDepth(); d++) setLevelOpBCValues(s->mixedop(d), fn, user);
  }
  s->fabOps.clear();
  s->fabFac.reset();
  s->fabH.reset();
  API_END
}
int amrmg_solver_set_vc_coefs(amrmg_solver* s, const double* a, const double* b0, const double* b1,
                              const double* b2) {
  API_BEGIN
  if (s->ncomp > 1) throw std::runtime_error("variable coefficients take one component (ncomp > 1 solver)");
  // the operators that carry coefficients: AMR levels 0..nlev-1, then the multigrid
  // depths 1.. of level 0 (depth 0 is AMR level 0)
  const int nd = s->H.mgDepth(), nl = s->nlev();
  std::vector<AMRPoissonOp<double>*> ops;
  std::vector<AMRPoissonOp<float>*> ops32;
  std::vector<const DeviceLayout*> lays;
  for (int l = 0; l < nl; l++) {
    ops.push_back(s->amg.m_op[l].get());
    if (s->mixed) ops32.push_back(s->mixed->amg.m_op[l].get());
    lays.push_back(s->H.amrLay[l].get());
  }
  for (int d = 1; d < nd; d++) {
    ops.push_back(s->mgop(d));
    if (s->mixed) ops32.push_back(s->mixedop(d));
    lays.push_back(s->mglay(d));
  }
  Context* c = s->ctx->c.get();
  if (a && ops[0]->m_relaxMode != 1 && ops[0]->m_relaxMode != 4)
    throw std::runtime_error("only relax modes 1 and 4 are implemented for variable coefficients");
  if (!a) {  // back to constant coefficients
    for (auto* op : ops) op->setVC(nullptr);
    for (auto* op : ops32) op->setVC(nullptr);
    s->vc.clear();
    s->vc32.clear();
  } else {
    if (!b0 || !b1 || !b2) throw std::runtime_error("variable coefficients: bCoef arrays missing");
    std::vector<std::shared_ptr<VCCoefs<double>>> v;
    for (const DeviceLayout* L : lays) v.emplace_back(new VCCoefs<double>(L));
    // every AMR level its own coefficients (VCAMRPoissonOp2Factory::define, AMRnewOp)
    const double* bs[3] = {b0, b1, b2};
    long long ao = 0, bo[3] = {0, 0, 0};
    for (int l = 0; l < nl; l++) {
      v[l]->setCells(a + ao);
      ao += lays[l]->valid_cells;
      for (int q = 0; q < 3; q++) {
        v[l]->setFaces(q, bs[q] + bo[q]);
        bo[q] += v[l]->packedFaces(q);
      }
    }
    // multigrid depths: level 0's coefficients averaged with ratio 2^d (MGnewOp)
    for (int d = 1; d < nd; d++) vc::average_coefs<double>(*v[0], *v[nl + d - 1], 1 << d, s->vcHarmonic);
    for (size_t q = 0; q < ops.size(); q++) ops[q]->setVC(v[q]);
    s->vc = v;
    s->vc32.clear();
    for (size_t q = 0; q < ops32.size(); q++) {
      std::shared_ptr<VCCoefs<float>> f(new VCCoefs<float>(lays[q]));
      vc::narrow(*v[q], *f);
      ops32[q]->setVC(f);
      s->vc32.push_back(f);
    }
  }
  // the bottom of the multigrid: the device solver for constant coefficients, the
  // host-driven BiCGStab otherwise (both bind again here)
  s->amg.m_mg0->setBottomSolver(s->bottom.get());
  if (s->mixed) s->mixed->rebindBottom();
  c->sync();
  API_END
}
long long amrmg_solver_level_cells(amrmg_solver* s, int lev) {
  if (lev < 0 || lev >= s->nlev()) return -1;
  return s->H.amrLay[lev]->valid_cells;
}
int amrmg_solver_num_mg_depths(amrmg_solver* s) { return s->H.mgDepth(); }
long long amrmg_solver_mg_cells(amrmg_solver* s, int depth) {
  if (depth < 0 || depth >= s->H.mgDepth()) return -1;
  return s->mglay(depth)->valid_cells;
}
long long amrmg_solver_device_bytes(amrmg_solver* s) {
  long long total = hierarchy_field_bytes(s->H, s->mixed ? 1 : 0) * s->ncomp;
  for (auto& v : s->vc) total += v->bytes();
  for (auto& v : s->vc32) total += v->bytes();
  return total;
}

static std::vector<std::unique_ptr<LevelData<double>>>& field_of(amrmg_solver* s, int field) {
  if (field == 0) return s->phi;
  if (field == 1) return s->rhs;
  if (field == 2) return s->res;
  throw std::runtime_error("field must be 0 (phi), 1 (rhs) or 2 (residual)");
}
int amrmg_solver_set_field(amrmg_solver* s, int field, const double* h) {
  API_BEGIN
  auto& f = field_of(s, field);
  long long o = 0;
  for (int l = 0; l < s->nlev(); l++) {
    upload_valid(*s->H.amrLay[l], f[l]->d, h + o, f[l]->nc, f[l]->cs);
    o += s->H.amrLay[l]->valid_cells * f[l]->nc;
  }
  API_END
}
int amrmg_solver_get_field(amrmg_solver* s, int field, double* h) {
  API_BEGIN
  auto& f = field_of(s, field);
  long long o = 0;
  for (int l = 0; l < s->nlev(); l++) {
    download_valid(*s->H.amrLay[l], f[l]->d, h + o, f[l]->nc, f[l]->cs);
    o += s->H.amrLay[l]->valid_cells * f[l]->nc;
  }
  API_END
}

int amrmg_solver_solve(amrmg_solver* s, int zero_phi, double* hist, int hist_cap, int* info) {
  return amrmg_solver_solve_levels(s, s->nlev() - 1, 0, zero_phi, hist, hist_cap, info);
}
int amrmg_solver_solve_levels(amrmg_solver* s, int l_max, int l_base, int zero_phi, double* hist, int hist_cap,
                              int* info) {
  API_BEGIN
  if (l_base < 0 || l_base > l_max || l_max >= s->nlev()) throw std::runtime_error("need 0 <= l_base <= l_max < nlev");
  if (l_base > 0 && s->mixed) throw std::runtime_error("l_base > 0 is not supported in mixed precision");
  if (l_base > 0 && !s->vc.empty()) throw std::runtime_error("l_base > 0 is not supported with variable coefficients");
  auto ph = s->ptrs(s->phi), rh = s->ptrs(s->rhs), ur = s->ptrs(s->uberRes), uc = s->ptrs(s->uberCorr);
  long long b0 = s->mixed ? s->mixed->bottomIterations() : s->bottom->totalIterations();
  s->amg.solveNoInitResid(ph, ur, uc, rh, l_max, l_base, zero_phi != 0);
  s->ctx->c->sync();
  if (hist)
    for (int q = 0; q < hist_cap && q < (int)s->amg.m_history.size(); q++) hist[q] = s->amg.m_history[q];
  if (info) {
    info[0] = s->amg.m_iterations;
    info[1] = s->amg.m_exitStatus;
    info[2] = (int)((s->mixed ? s->mixed->bottomIterations() : s->bottom->totalIterations()) - b0);
  }
  API_END
}
int amrmg_solver_residual(amrmg_solver* s, int homog, double* norm_out) {
  API_BEGIN
  auto ph = s->ptrs(s->phi), rh = s->ptrs(s->rhs), rs = s->ptrs(s->res);
  double n = s->amg.computeAMRResidual(rs, ph, rh, s->nlev() - 1, 0, homog != 0, true);
  if (norm_out) *norm_out = n;
  API_END
}
int amrmg_solver_operator(amrmg_solver* s, int homog) {
  API_BEGIN
  auto ph = s->ptrs(s->phi), rs = s->ptrs(s->res);
  s->amg.computeAMROperator(rs, ph, s->nlev() - 1, 0, homog != 0);
  s->ctx->c->sync();
  API_END
}

int amrmg_level_apply(amrmg_solver* s, int depth, const double* phi, double* out, int homog) {
  API_BEGIN
  const DeviceLayout* L = s->mglay(depth);
  LevelData<double> p(L), o(L);
  upload_valid(*L, p.d, phi);
  s->mgop(depth)->applyOp(o, p, homog != 0);
  download_valid(*L, o.d, out);
  API_END
}
int amrmg_level_norm(amrmg_solver* s, int depth, const double* x, int p, double* out) {
  API_BEGIN
  const DeviceLayout* L = s->mglay(depth);
  LevelData<double> v(L);
  upload_valid(*L, v.d, x);
  *out = s->mgop(depth)->norm(v, p);
  API_END
}
int amrmg_level_residual(amrmg_solver* s, int depth, const double* phi, const double* rhs, double* out,
                         int homog) {
  API_BEGIN
  const DeviceLayout* L = s->mglay(depth);
  LevelData<double> p(L), r(L), o(L);
  upload_valid(*L, p.d, phi);
  upload_valid(*L, r.d, rhs);
  s->mgop(depth)->residual(o, p, r, homog != 0);
  download_valid(*L, o.d, out);
  API_END
}
int amrmg_level_relax(amrmg_solver* s, int depth, double* phi, const double* rhs, int iters) {
  API_BEGIN
  const DeviceLayout* L = s->mglay(depth);
  LevelData<double> p(L), r(L);
  upload_valid(*L, p.d, phi);
  upload_valid(*L, r.d, rhs);
  s->mgop(depth)->relax(p, r, iters);
  download_valid(*L, p.d, phi);
  API_END
}
int amrmg_level_restrict(amrmg_solver* s, int depth, const double* phi, const double* rhs, double* resc) {
  API_BEGIN
  if (depth + 1 >= s->H.mgDepth()) throw std::runtime_error("no coarser multigrid level");
  const DeviceLayout* L = s->mglay(depth);
  const DeviceLayout* Lc = s->mglay(depth + 1);
  LevelData<double> p(L), r(L), c(Lc);
  upload_valid(*L, p.d, phi);
  upload_valid(*L, r.d, rhs);
  s->mgop(depth)->restrictResidual(c, p, r);
  download_valid(*Lc, c.d, resc);
  API_END
}
int amrmg_level_prolong(amrmg_solver* s, int depth, double* phi, const double* ec) {
  API_BEGIN
  if (depth + 1 >= s->H.mgDepth()) throw std::runtime_error("no coarser multigrid level");
  const DeviceLayout* L = s->mglay(depth);
  const DeviceLayout* Lc = s->mglay(depth + 1);
  LevelData<double> p(L), c(Lc);
  upload_valid(*L, p.d, phi);
  upload_valid(*Lc, c.d, ec);
  s->mgop(depth)->prolongIncrement(p, c);
  download_valid(*L, p.d, phi);
  API_END
}
int amrmg_kernel_gsrb(amrmg_solver* s, int lev, double* phi_g, const double* rhs, int color) {
  API_BEGIN
  if (lev < 0 || lev >= s->nlev()) throw std::runtime_error("bad level");
  const DeviceLayout* L = s->fabOp(lev)->m_lay;
  LevelData<double> p(L), r(L);
  upload_ghosted(*L, p.d, phi_g);
  upload_valid(*L, r.d, rhs);
  ops::gsrb<double>(*L, p.d, r.d, color, s->H.dxs[lev], s->alpha, s->beta);
  download_ghosted(*L, p.d, phi_g);
  API_END
}
int amrmg_homog_cf(amrmg_solver* s, int lev, double* phi_g) {
  API_BEGIN
  if (lev < 0 || lev >= s->nlev()) throw std::runtime_error("bad level");
  AMRPoissonOp<double>* op = s->fabOp(lev);
  const DeviceLayout* L = op->m_lay;
  LevelData<double> p(L);
  upload_ghosted(*L, p.d, phi_g);
  ops::fill_ghosts<double>(*L, op->m_ghost, p.d, 2, 1, op->m_dx, op->m_dxCrse);
  download_ghosted(*L, p.d, phi_g);
  API_END
}
int amrmg_quad_cf(amrmg_solver* s, int lev, double* phi_g, const double* coarse) {
  API_BEGIN
  if (lev < 1 || lev >= s->nlev()) throw std::runtime_error("bad level");
  AMRPoissonOp<double>* op = s->fabOp(lev);
  const DeviceLayout* L = op->m_lay;
  const DeviceLayout* Lc = op->m_coarserLay;
  LevelData<double> p(L), c(Lc);
  upload_ghosted(*L, p.d, phi_g);
  upload_valid(*Lc, c.d, coarse);
  op->coarseFineInterp(p, c);
  download_ghosted(*L, p.d, phi_g);
  API_END
}
int amrmg_average(amrmg_solver* s, int lev, const double* fine, double* coarse) {
  API_BEGIN
  if (lev < 1 || lev >= s->nlev()) throw std::runtime_error("bad level");
  const DeviceLayout* L = s->H.amrLay[lev].get();
  const DeviceLayout* Lc = s->H.cfLay[lev].get();
  LevelData<double> f(L), c(Lc);
  upload_valid(*L, f.d, fine);
  ops::average<double>(*L, *Lc, c.d, f.d, s->H.ratios[lev]);
  download_valid(*Lc, c.d, coarse);
  API_END
}
// BiCGStabSolver's public parameters of every bottom solve of the solver
static void copy_bottom_params(BiCGStabSolver<LevelData<double>>& to, const BiCGStabSolver<LevelData<double>>& from) {
  to.m_imax = from.m_imax;
  to.m_numRestarts = from.m_numRestarts;
  to.m_reps = from.m_reps;
  to.m_hang = from.m_hang;
  to.m_small = from.m_small;
}
int amrmg_solver_set_bottom_params(amrmg_solver* s, const int* ip, const double* rp) {
  API_BEGIN
  if (ip[0] < 0 || ip[1] < 0) throw std::runtime_error("BiCGStab: imax and numRestarts must be >= 0");
  BiCGStabSolver<LevelData<double>>& f = s->bottom->m_fallback;
  f.m_imax = ip[0];
  f.m_numRestarts = ip[1];
  f.m_reps = rp[0];
  f.m_hang = rp[1];
  f.m_small = rp[2];
  if (s->mixed) {
    copy_bottom_params(s->mixed->bottom.m_bicg, f);
    if (s->mixed->devBottom) {
      auto& g = s->mixed->devBottom->m_fallback;
      g.m_imax = f.m_imax; g.m_numRestarts = f.m_numRestarts; g.m_reps = f.m_reps; g.m_hang = f.m_hang;
      g.m_small = f.m_small;
    }
  }
  API_END
}
int amrmg_bottom_solve(amrmg_solver* s, double* corr, const double* res, int nbottom, double metric, double eps,
                       int* info) {
  API_BEGIN
  int d = s->H.mgDepth() - 1;
  const DeviceLayout* L = s->mglay(d);
  LevelData<double> c(L), r(L);
  upload_valid(*L, c.d, corr);
  upload_valid(*L, r.d, res);
  AMRPoissonOp<double>* op = s->mgop(d);
  BiCGStabSolver<LevelData<double>> b;
  copy_bottom_params(b, s->bottom->m_fallback);
  b.define(op, true);
  b.setConvergenceMetrics(metric, eps);
  op->relax(c, r, nbottom);
  b.solve(c, r);
  download_valid(*L, c.d, corr);
  if (info) {
    info[0] = b.m_exitStatus;
    info[1] = b.m_iter;
  }
  API_END
}
int amrmg_bottom_solve_device(amrmg_solver* s, double* corr, const double* res, int nbottom, double metric,
                              double eps, int* info) {
  API_BEGIN
  int d = s->H.mgDepth() - 1;
  const DeviceLayout* L = s->mglay(d);
  AMRPoissonOp<double>* op = s->mgop(d);
  DeviceBottomSolver<double> b(s->ctx->c.get());
  copy_bottom_params(b.m_fallback, s->bottom->m_fallback);
  b.define(op, true);
  if (!b.onDevice()) throw std::runtime_error("the deepest multigrid level does not allow the device bottom solver");
  LevelData<double> c(L), r(L);
  upload_valid(*L, c.d, corr);
  upload_valid(*L, r.d, res);
  b.setConvergenceMetrics(metric, eps);
  op->relax(c, r, nbottom);
  b.solve(c, r);
  const long long it = b.totalIterations();
  download_valid(*L, c.d, corr);
  if (info) info[0] = (int)it;
  API_END
}

int amrmg_solver_tune_relax(amrmg_solver* s, double* table, int cap, int* nrec) {
  API_BEGIN
  constexpr int W = 6 + ops::kFusedShapeCount * 4;
  int n = 0;
  auto rec = [&](int kind, int idx, auto* op, auto& phi, auto& rhs) {
    double times[ops::kFusedShapeCount * 4];
    if (!op->tuneRelax(phi, rhs, times)) return;
    if (table && n < cap) {
      double* r = table + (size_t)n * W;
      r[0] = kind;
      r[1] = idx;
      r[2] = (double)op->m_lay->valid_cells;
      r[3] = op->m_relaxShape;
      r[4] = op->m_relaxKc;
      r[5] = times[op->m_relaxShape * 4 + (op->m_relaxKc == 0 ? 0 : op->m_relaxKc == 8 ? 1 : op->m_relaxKc == 16 ? 2 : 3)];
      for (int q = 0; q < ops::kFusedShapeCount * 4; q++) r[6 + q] = times[q];
    }
    n++;
  };
  if (s->mixed) {  // the relaxation of a mixed solve runs in FP32
    auto& m = *s->mixed;
    for (int l = 0; l < s->nlev(); l++) rec(0, l, m.amg.m_op[l].get(), *m.corr[l], *m.res[l]);
    for (int d = 1; d < s->H.mgDepth(); d++)
      rec(1, d, static_cast<AMRPoissonOp<float>*>(m.amg.m_mg0->m_op[d]), *m.amg.m_mg0->m_correction[d],
          *m.amg.m_mg0->m_residual[d]);
  } else {
    for (int l = 0; l < s->nlev(); l++) rec(0, l, s->amg.m_op[l].get(), *s->phi[l], *s->rhs[l]);
    for (int d = 1; d < s->H.mgDepth(); d++)
      rec(1, d, s->mgop(d), *s->amg.m_mg0->m_correction[d], *s->amg.m_mg0->m_residual[d]);
  }
  s->ctx->c->sync();
  if (nrec) *nrec = n;
  API_END
}
int amrmg_solver_set_relax_shape(amrmg_solver* s, int shape, int kc) {
  API_BEGIN
  if (shape < 0 || shape >= ops::kFusedShapeCount || kc < 0 || kc > 1024)
    throw std::runtime_error("relax shape out of range");
  // a shape that does not apply to a level (x quads on FP64 or on an x extent that
  // is not a multiple of 4) leaves that level's shape as it is
  auto set = [&](auto* op, int bytes) {
    if (!ops::fused_shape_ok(shape, bytes, op->m_lay->rect.len(0))) return;
    op->m_relaxShape = shape;
    op->m_relaxKc = kc;
  };
  for (int l = 0; l < s->nlev(); l++) set(s->amg.m_op[l].get(), 8);
  for (int d = 1; d < s->H.mgDepth(); d++) set(s->mgop(d), 8);
  if (s->mixed) {
    auto& m = s->mixed->amg;
    for (int l = 0; l < s->nlev(); l++) set(m.m_op[l].get(), 4);
    for (int d = 1; d < s->H.mgDepth(); d++) set(static_cast<AMRPoissonOp<float>*>(m.m_mg0->m_op[d]), 4);
  }
  API_END
}
int amrmg_solver_set_relax_mode(amrmg_solver* s, int mode) {
  API_BEGIN
  if (mode < 0 || mode > 5)
    throw std::runtime_error(
        "relax mode must be 0 (looseGSRB), 1 (levelGSRB), 2 (overlapGSRB), 3 (levelGSRBLazy), 4 (levelJacobi) or "
        "5 (levelMultiColor)");
  // VCAMRPoissonOp2 aborts in looseGSRB, overlapGSRB, levelGSRBLazy and levelMultiColor
  // (VCAMRPoissonOp2.cpp:628-640, 748-760)
  if (mode != 1 && mode != 4 && !s->vc.empty())
    throw std::runtime_error("only relax modes 1 and 4 are implemented for variable coefficients");
  for (auto& op : s->amg.m_op) op->m_relaxMode = mode;
  for (int d = 1; d < s->H.mgDepth(); d++) s->mgop(d)->m_relaxMode = mode;
  if (s->mixed) {
    for (auto& op : s->mixed->amg.m_op) op->m_relaxMode = mode;
    for (int d = 1; d < s->H.mgDepth(); d++) s->mixedop(d)->m_relaxMode = mode;
  }
  s->amg.m_mg0->setBottomSolver(s->bottom.get());
  if (s->mixed) s->mixed->rebindBottom();
  API_END
}
int amrmg_solver_set_vc_average(amrmg_solver* s, int type) {
  API_BEGIN
  if (type != 0 && type != 1) throw std::runtime_error("coefficient average type must be 0 (arithmetic) or 1 (harmonic)");
  s->vcHarmonic = type;
  if (!s->vc.empty()) {
    // coefficients are already set: the multigrid depths take level 0's coefficients
    // averaged the new way now (MGnewOp), with lambda and the FP32 copies redone
    const int nd = s->H.mgDepth(), nl = s->nlev();
    for (int d = 1; d < nd; d++) {
      auto& v = s->vc[nl + d - 1];
      vc::average_coefs<double>(*s->vc[0], *v, 1 << d, s->vcHarmonic);
      s->mgop(d)->setVC(v);
      if (s->mixed) {
        auto& f = s->vc32[nl + d - 1];
        vc::narrow(*v, *f);
        s->mixedop(d)->setVC(f);
      }
    }
    s->ctx->c->sync();
  }
  API_END
}
int amrmg_solver_set_vc_relax_shape(amrmg_solver* s, int shape) {
  API_BEGIN
  if (shape < -1 || shape > 3) throw std::runtime_error("variable-coefficient relax shape out of range");
  for (int d = 0; d < s->H.mgDepth(); d++) s->mgop(d)->m_vcShape = shape;
  if (s->mixed)
    for (int d = 0; d < s->H.mgDepth(); d++) s->mixedop(d)->m_vcShape = shape;
  API_END
}
// ---- heat-equation steppers (BackwardEuler, AMRTGA) over a solver ----
// d phi/dt = L phi + S with the solver's operator as L: every step resets alpha and
// beta of all the solver's operators, as the reference's steppers do.
struct amrmg_heat {
  amrmg_solver* s = nullptr;
  int scheme = 0;                                   // 0 BackwardEuler, 1 AMRTGA
  double mu1 = 0, mu2 = 0, mu3 = 0, mu4 = 0;
  std::vector<std::unique_ptr<LevelData<double>>> src, old, rhst, srct;
  AMRPoissonOp<double>* op(int l) { return s->amg.m_op[l].get(); }
  // VCAMRPoissonOp2::diagonalScale / divideByIdentityCoef; no-ops for AMRPoissonOp
  void aScale(int l, LevelData<double>& x, bool divide) {
    if (op(l)->m_vc) vc::scale_by_a<double>(*x.lay, *op(l)->m_vc, x.d, divide);
  }
  void setAlphaBeta(double alpha, double beta) {
    if (amrmg_solver_set_alpha_beta(s, alpha, beta) != 0) throw std::runtime_error(amrmg_g_err);
  }
  // AMRMultiGrid::solveNoInit (zeroPhi as given) on the solver's phi with rhs rhst
  void solve(bool zeroPhi, int* info) {
    auto ph = s->ptrs(s->phi), rh = s->ptrs(rhst), ur = s->ptrs(s->uberRes), uc = s->ptrs(s->uberCorr);
    s->amg.solveNoInitResid(ph, ur, uc, rh, s->nlev() - 1, 0, zeroPhi);
    info[0] = s->amg.m_iterations;
    info[1] = s->amg.m_exitStatus;
  }
};

int amrmg_heat_create(amrmg_solver* s, int scheme, amrmg_heat** out) {
  API_BEGIN
  if (scheme != 0 && scheme != 1) throw std::runtime_error("heat scheme must be 0 (backward Euler) or 1 (TGA)");
  if (s->ncomp > 1) throw std::runtime_error("the heat steppers take one component (ncomp > 1 solver)");
  std::unique_ptr<amrmg_heat> h(new amrmg_heat());
  h->s = s;
  h->scheme = scheme;
  // AMRTGA::setMu (AMRTGA.H:368-384)
  const double tgaEpsilon = 1.e-12;
  const double a = 2.0 - std::sqrt(2.0) - tgaEpsilon;
  h->mu1 = (a - std::sqrt(a * a - 4.0 * a + 2.0)) / 2.0;
  h->mu2 = (a + std::sqrt(a * a - 4.0 * a + 2.0)) / 2.0;
  h->mu3 = (1.0 - a);
  h->mu4 = 0.5 - a;
  for (int l = 0; l < s->nlev(); l++) {
    const DeviceLayout* L = s->H.amrLay[l].get();
    h->src.emplace_back(new LevelData<double>(L));
    h->old.emplace_back(new LevelData<double>(L));
    h->rhst.emplace_back(new LevelData<double>(L));
    h->srct.emplace_back(new LevelData<double>(L));
    ops::setval<double>(*L, h->src[l]->d, 0.0);
  }
  s->ctx->c->sync();
  *out = h.release();
  API_END
}
int amrmg_heat_destroy(amrmg_heat* h) {
  API_BEGIN
  delete h;
  API_END
}
int amrmg_heat_set_source(amrmg_heat* h, const double* valid_packed) {
  API_BEGIN
  long long o = 0;
  for (int l = 0; l < h->s->nlev(); l++) {
    const DeviceLayout& L = *h->s->H.amrLay[l];
    upload_valid(L, h->src[l]->d, valid_packed + o);
    o += L.valid_cells;
  }
  API_END
}
int amrmg_heat_step(amrmg_heat* h, double dt, int zero_phi, int* info) {
  API_BEGIN
  amrmg_solver* s = h->s;
  const int nl = s->nlev();
  int i1[2] = {0, 0}, i2[2] = {0, 0};
  for (int l = 0; l < nl; l++) h->op(l)->assignLocal(*h->old[l], *s->phi[l]);
  if (h->scheme == 0) {
    // BackwardEuler::createEulerRHS (BackwardEuler.cpp:196-233): rhs = a phi^n + dt S
    for (int l = 0; l < nl; l++) {
      AMRPoissonOp<double>* op = h->op(l);
      op->setToZero(*h->rhst[l]);
      op->incr(*h->rhst[l], *h->old[l], 1.0);
      h->aScale(l, *h->rhst[l], false);
      op->setToZero(*h->srct[l]);
      op->incr(*h->srct[l], *h->src[l], 1.0);
      op->scale(*h->srct[l], dt);
      op->incr(*h->rhst[l], *h->srct[l], 1.0);
    }
    // oneStep: phi^{n+1} starts from phi^n; solveHelm (:236-254): alpha 1, beta -dt
    h->setAlphaBeta(1.0, -dt);
    h->solve(zero_phi != 0, i1);
  } else {
    // AMRTGA::oneStep (AMRTGA.H:420-510)
    for (int l = 0; l < nl; l++) {
      h->op(l)->setToZero(*h->srct[l]);
      h->op(l)->incr(*h->srct[l], *h->src[l], 1.0);
      h->aScale(l, *h->srct[l], true);
    }
    auto srct = s->ptrs(h->srct), rhst = s->ptrs(h->rhst), ph = s->ptrs(s->phi), old = s->ptrs(h->old);
    h->setAlphaBeta(1.0, h->mu4 * dt);                     // applyHelm (:514-537)
    s->amg.computeAMROperator(rhst, srct, nl - 1, 0, true);
    for (int l = 0; l < nl; l++) h->op(l)->scale(*h->rhst[l], dt);
    h->setAlphaBeta(1.0, h->mu3 * dt);
    s->amg.computeAMROperator(ph, old, nl - 1, 0, false);
    for (int l = 0; l < nl; l++) {
      h->op(l)->incr(*h->rhst[l], *s->phi[l], 1.0);
      h->op(l)->assignLocal(*s->phi[l], *h->old[l]);
    }
    h->setAlphaBeta(1.0, -h->mu2 * dt);                    // solveHelm (:540-560)
    h->solve(false, i1);
    for (int l = 0; l < nl; l++) {
      h->op(l)->assignLocal(*h->rhst[l], *s->phi[l]);
      h->aScale(l, *h->rhst[l], false);
      h->op(l)->assignLocal(*s->phi[l], *h->old[l]);
    }
    h->setAlphaBeta(1.0, -h->mu1 * dt);
    h->solve(false, i2);
  }
  s->ctx->c->sync();
  if (info) {
    info[0] = i1[0];
    info[1] = i1[1];
    info[2] = i2[0];
    info[3] = i2[1];
  }
  API_END
}
// ---- batched single-level solves ----
int amrmg_batch_create(amrmg_ctx* ctx, int nprob, const int* ispecs, const double* rspecs, int precision,
                       amrmg_batch** out) {
  return amrmg_batch_create_lanes(ctx, nprob, ispecs, rspecs, precision, 1, out);
}
int amrmg_batch_create_lanes(amrmg_ctx* ctx, int nprob, const int* ispecs, const double* rspecs, int precision,
                             int lanes, amrmg_batch** out) {
  API_BEGIN
  if (nprob <= 0 || nprob > (1 << 20)) throw std::runtime_error("batch: bad problem count");
  if (lanes < 1 || lanes > 3) throw std::runtime_error("batch: lanes must be 1, 2 or 3");
  std::vector<BatchProblemSpec> probs;
  const int* is = ispecs;
  for (int q = 0; q < nprob; q++) {
    if (is[0] != 1) throw std::runtime_error("batch: every problem must have one level");
    BatchProblemSpec P;
    double a, b, dx0;
    std::vector<DBL> g = parse_spec(is, rspecs + 9 * q, P.domain, P.bc, dx0, a, b);
    P.boxes = g[0];
    P.dx = dx0;
    P.alpha = a;
    P.beta = b;
    probs.push_back(P);
    is += 14 + 6 * is[13];
  }
  std::unique_ptr<amrmg_batch> bt(new amrmg_batch());
  bt->ctx = ctx;
  bt->nprob = nprob;
  // lanes by the size of the multigrid bottom: the problems whose bottom holds at
  // least 16^3 cells (their BiCGStab bottoms are the long pole of every V-cycle)
  // lockstep on a stream of their own, so the other problems' cycles fill the
  // device meanwhile; with three lanes the bottoms of at least 27^3 cells get a
  // third.  Empty lanes are dropped.
  const long long kBottomLane[2] = {4096, 19683};
  std::vector<int> want(nprob, 0);
  for (int q = 0; q < nprob; q++) {
    const long long bc = lanes > 1 ? batchBottomCells(probs[q]) : 0;
    if (bc >= kBottomLane[0]) want[q] = 1;
    if (lanes > 2 && bc >= kBottomLane[1]) want[q] = 2;
  }
  std::vector<int> remap(3, -1);
  int nl = 0;
  for (int l = 0; l < 3; l++)
    if (std::count(want.begin(), want.end(), l) > 0) remap[l] = nl++;
  std::vector<int> lane(nprob);
  for (int q = 0; q < nprob; q++) lane[q] = remap[want[q]];
  bt->lanes.resize(nl);
  bt->at.resize(nprob);
  std::vector<std::vector<BatchProblemSpec>> lp(nl);
  for (int q = 0; q < nprob; q++) {
    bt->at[q] = {lane[q], (int)lp[lane[q]].size()};
    lp[lane[q]].push_back(probs[q]);
    bt->lanes[lane[q]].probs.push_back(q);
  }
  for (int l = 0; l < nl; l++) {
    Context* c = ctx->c.get();
    if (l > 0) {
      bt->lanes[l].own.reset(new Context(ctx->c->device, ctx->c->dry));
      c = bt->lanes[l].own.get();
    }
    bt->lanes[l].solver.reset(new BatchSolver(c, std::move(lp[l]), precision));
  }
  *out = bt.release();
  API_END
}
int amrmg_batch_destroy(amrmg_batch* b) {
  API_BEGIN
  delete b;
  API_END
}
int amrmg_batch_set_params(amrmg_batch* b, const int* sp, const double* rp) {
  API_BEGIN
  for (auto& l : b->lanes) l.solver->setSolverParameters(sp[0], sp[1], sp[2], sp[3], sp[4], sp[5], rp[0], rp[1], rp[2]);
  API_END
}
int amrmg_batch_set_rhs(amrmg_batch* b, int q, const double* valid_packed) {
  API_BEGIN
  int k;
  BatchSolver& s = b->solver(q, k);
  s.setRhs(k, valid_packed);
  API_END
}
int amrmg_batch_get_phi(amrmg_batch* b, int q, double* valid_packed) {
  API_BEGIN
  int k;
  BatchSolver& s = b->solver(q, k);
  s.getPhi(k, valid_packed);
  API_END
}
int amrmg_batch_solve(amrmg_batch* b, int* iterations, int* exit_status) {
  API_BEGIN
  Context* main = b->ctx->c.get();
  if (b->lanes.size() == 1) {
    b->lanes[0].solver->solve();
  } else if (main->prof) {
    // with the event profiler on, the lanes run one after another on the profiled
    // stream, so every kernel is timed alone
    for (auto& l : b->lanes) {
      Context* own = l.solver->context();
      l.solver->setContext(main);
      try {
        l.solver->solve();
      } catch (...) {
        l.solver->setContext(own);
        throw;
      }
      l.solver->setContext(own);
    }
  } else {
    // each lane after the first is driven by a thread of its own on its own stream
    std::vector<std::string> err(b->lanes.size());
    std::vector<std::thread> th;
    for (size_t l = 1; l < b->lanes.size(); l++)
      th.emplace_back([&, l] {
        try {
          hip_check(hipSetDevice(main->device), "hipSetDevice");
          b->lanes[l].solver->solve();
        } catch (const std::exception& e) {
          err[l] = e.what();
        }
      });
    try {
      b->lanes[0].solver->solve();
    } catch (const std::exception& e) {
      err[0] = e.what();
    }
    for (auto& t : th) t.join();
    for (auto& e : err)
      if (!e.empty()) throw std::runtime_error(e);
  }
  for (int q = 0; q < b->nprob; q++) {
    int k;
    BatchSolver& s = b->solver(q, k);
    if (iterations) iterations[q] = s.iterations(k);
    if (exit_status) exit_status[q] = s.exitStatus(k);
  }
  API_END
}
int amrmg_batch_history(amrmg_batch* b, int q, double* hist, int cap, int* n) {
  API_BEGIN
  int k;
  BatchSolver& s = b->solver(q, k);
  const std::vector<double>& h = s.history(k);
  for (int i = 0; i < cap && i < (int)h.size(); i++) hist[i] = h[i];
  if (n) *n = (int)h.size();
  API_END
}
int amrmg_batch_set_relax_shape(amrmg_batch* b, int depth, int shape, int kc) {
  API_BEGIN
  if (kc < 0 || kc > 1024) throw std::runtime_error("batch: z-chunk out of range");
  for (auto& l : b->lanes) {
    if (depth < 0 || depth >= l.solver->numDepths()) throw std::runtime_error("batch: bad depth");
    l.solver->setRelaxShape(depth, shape, kc);
  }
  API_END
}
int amrmg_batch_tune_relax(amrmg_batch* b, double* ms, int cap, int* ndepth) {
  API_BEGIN
  // every lane is tuned; the reported times are the per-depth sums over the lanes
  std::vector<double> t;
  for (auto& l : b->lanes) {
    std::vector<double> tl = l.solver->tuneRelax();
    if (tl.size() > t.size()) t.resize(tl.size(), 0.0);
    for (size_t i = 0; i < tl.size(); i++) t[i] += tl[i];
  }
  for (int i = 0; i < cap && i < (int)t.size(); i++) ms[i] = t[i];
  if (ndepth) *ndepth = (int)t.size();
  API_END
}
int amrmg_batch_stats(amrmg_batch* b, double* cell_relax, long long* bottom_iters, long long* device_bytes) {
  API_BEGIN
  double cr = 0;
  long long bi = 0, db = 0;
  for (auto& l : b->lanes) {
    cr += l.solver->cellRelaxations();
    bi += l.solver->bottomIterations();
    db += l.solver->deviceBytes();
  }
  if (cell_relax) *cell_relax = cr;
  if (bottom_iters) *bottom_iters = bi;
  if (device_bytes) *device_bytes = db;
  API_END
}
int amrmg_batch_set_graph(amrmg_batch* b, int on, long long* launches) {
  API_BEGIN
  long long n = 0;
  for (auto& l : b->lanes) {
    if (on >= 0) l.solver->setGraph(on != 0);
    n += l.solver->graphLaunches();
  }
  if (launches) *launches = n;
  API_END
}
int amrmg_batch_set_reduction_order(amrmg_batch* b, int order) {
  API_BEGIN
  if (order != 0 && order != 1) throw std::runtime_error("batch: reduction order must be 0 or 1");
  for (auto& l : b->lanes) l.solver->setTreeOrder(order == 1);
  API_END
}
int amrmg_batch_lanes(amrmg_batch* b, int* lanes, int* lane_of_problem) {
  API_BEGIN
  if (lanes) *lanes = (int)b->lanes.size();
  if (lane_of_problem)
    for (int q = 0; q < b->nprob; q++) lane_of_problem[q] = b->at[q].first;
  API_END
}

// ---- serving: a request loop over device-resident hierarchies ----
namespace {
double serve_now(const std::chrono::steady_clock::time_point& t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
// HBM footprint of a solver for a spec: its layouts built on the host only (a dry
// context), counted as amrmg_solver_device_bytes counts a created solver's fields
long long serve_footprint(const std::vector<DBL>& g, const Box& dom, double dx0, int precision, const BCSpec& bc) {
  Context dry(0, true);
  DeviceHierarchy H;
  H.define(&dry, g, dom, dx0, true, &bc);
  return hierarchy_field_bytes(H, precision);
}
struct ServeInstance {
  amrmg_ctx* ctx = nullptr;  // its own stream and reduction scratch
  amrmg_solver* s = nullptr;
  std::string key;
  long long bytes = 0;
  double lastUse = 0;
  ~ServeInstance() {
    if (s) amrmg_solver_destroy(s);
    if (ctx) amrmg_ctx_destroy(ctx);
  }
};
struct ServeReq {
  long long id = 0;
  std::vector<int> ispec;
  std::vector<double> rspec, rhs, phi, hist;
  std::string key;
  long long bytes = 0, cells = 0;
  bool large = false;                        // runs in the lane of large requests
  double tArrival = 0, tAdmit = 0, tDone = 0;
  int iters = 0, exitStatus = -1;
  bool done = false, taken = false;
  int bypassed = 0;                          // times a later request was admitted before it
  int sp[6] = {0};                           // solver parameters, copied at admission
  double rp[3] = {0};
  std::string error;
};
// a request at the head of the queue that does not fit yet lets at most this many later
// (fitting) requests be admitted before it; then admission waits for it
constexpr int kServeMaxBypass = 8;
}  // namespace

struct amrmg_serve {
  int device = 0, precision = 0;
  long long budget = 0;
  int sp[6] = {2, 2, 2, 1, 100, 5};
  double rp[3] = {1e-10, 1e-15, 1e-30};
  std::mutex mu;
  std::condition_variable cv;
  std::deque<ServeReq*> queue;                              // waiting for admission, in arrival order
  std::unordered_map<std::string, long long> measured;     // device bytes of a created hierarchy, by spec
  std::unordered_map<long long, std::unique_ptr<ServeReq>> reqs;
  std::vector<std::unique_ptr<ServeInstance>> idle;        // device-resident, not in use
  long long bytesInUse = 0, bytesIdle = 0, maxBytes = 0;
  long long created = 0, reused = 0, evicted = 0;
  int running = 0, maxRunning = 0;
  // lanes: requests of at least largeCells cells are large, and at most maxLarge of
  // them run at once, so the other slots stay free for small requests (which then
  // never wait behind a large one for a slot)
  long long largeCells = 0;
  int maxLarge = 0, runningLarge = 0;
  bool stop = false;
  long long nextId = 1;
  std::vector<std::thread> workers;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  void worker() {
    for (;;) {
      ServeReq* r = nullptr;
      std::unique_ptr<ServeInstance> inst;
      std::vector<std::unique_ptr<ServeInstance>> evict;
      {
        std::unique_lock<std::mutex> lk(mu);
        // admission: the first request in arrival order whose footprint fits beside the
        // running ones and whose lane has room; a head that does not fit the budget is
        // passed at most kServeMaxBypass times (a head waiting for its lane only waits)
        size_t pick = 0;
        auto fits = [&](const ServeReq* q) { return bytesInUse + q->bytes <= budget; };
        auto lane_free = [&](const ServeReq* q) { return !q->large || runningLarge < maxLarge; };
        auto admissible = [&] {
          if (queue.empty()) return false;
          const ServeReq* h = queue.front();
          const size_t n = (!fits(h) && h->bypassed >= kServeMaxBypass) ? 1 : queue.size();
          for (size_t q = 0; q < n; q++)
            if (fits(queue[q]) && lane_free(queue[q])) {
              pick = q;
              return true;
            }
          return false;
        };
        cv.wait(lk, [&] { return stop || admissible(); });
        if (stop) return;
        r = queue[pick];
        if (pick > 0 && !fits(queue.front())) queue.front()->bypassed++;
        queue.erase(queue.begin() + pick);
        if (r->large) runningLarge++;
        std::memcpy(r->sp, sp, sizeof sp);
        std::memcpy(r->rp, rp, sizeof rp);
        for (size_t q = 0; q < idle.size(); q++)
          if (idle[q]->key == r->key) {
            inst = std::move(idle[q]);
            idle.erase(idle.begin() + q);
            bytesIdle -= inst->bytes;
            reused++;
            break;
          }
        // make room: evict the least recently used idle hierarchies
        while (bytesInUse + r->bytes + bytesIdle > budget && !idle.empty()) {
          auto lru = std::min_element(idle.begin(), idle.end(),
                                      [](const auto& a, const auto& b) { return a->lastUse < b->lastUse; });
          bytesIdle -= (*lru)->bytes;
          evict.push_back(std::move(*lru));
          idle.erase(lru);
          evicted++;
        }
        bytesInUse += r->bytes;
        maxBytes = std::max(maxBytes, bytesInUse + bytesIdle);
        running++;
        maxRunning = std::max(maxRunning, running);
        r->tAdmit = serve_now(t0);
      }
      evict.clear();
      try {
        if (!inst) {
          inst.reset(new ServeInstance());
          inst->key = r->key;
          inst->bytes = r->bytes;
          if (amrmg_ctx_create(device, &inst->ctx) || amrmg_solver_create(inst->ctx, r->ispec.data(), r->rspec.data(),
                                                                          precision, &inst->s))
            throw std::runtime_error(amrmg_last_error());
          // account the hierarchy by what it holds from now on (and admit later
          // requests of this spec by it)
          const long long mb = amrmg_solver_device_bytes(inst->s);
          std::lock_guard<std::mutex> lk(mu);
          created++;
          bytesInUse += mb - r->bytes;
          r->bytes = mb;
          inst->bytes = mb;
          measured[r->key] = mb;
          maxBytes = std::max(maxBytes, bytesInUse + bytesIdle);
        }
        double hist[256];
        int info[3];
        if (amrmg_solver_set_params(inst->s, r->sp, r->rp) || amrmg_solver_set_field(inst->s, 1, r->rhs.data()) ||
            amrmg_solver_solve(inst->s, 1, hist, 256, info))
          throw std::runtime_error(amrmg_last_error());
        r->phi.resize(r->rhs.size());
        if (amrmg_solver_get_field(inst->s, 0, r->phi.data())) throw std::runtime_error(amrmg_last_error());
        r->iters = info[0];
        r->exitStatus = info[1];
        r->hist.assign(hist, hist + std::min(256, info[0] + 1));
      } catch (const std::exception& e) {
        r->error = e.what();
        inst.reset();
      }
      std::lock_guard<std::mutex> lk(mu);
      r->tDone = serve_now(t0);
      r->done = true;
      r->rhs.clear();
      r->rhs.shrink_to_fit();
      bytesInUse -= r->bytes;
      running--;
      if (r->large) runningLarge--;
      if (inst) {
        inst->lastUse = r->tDone;
        bytesIdle += inst->bytes;
        idle.push_back(std::move(inst));
      }
      cv.notify_all();
    }
  }
};

int amrmg_serve_create(int device, int nslots, long long hbm_budget, int precision, amrmg_serve** out) {
  API_BEGIN
  if (nslots <= 0 || nslots > 64) throw std::runtime_error("serve: slots must be in 1..64");
  if (hbm_budget <= 0) throw std::runtime_error("serve: HBM budget must be positive");
  if (precision != 0 && precision != 1) throw std::runtime_error("serve: precision must be 0 or 1");
  std::unique_ptr<amrmg_serve> e(new amrmg_serve());
  e->device = device;
  e->precision = precision;
  e->budget = hbm_budget;
  e->maxLarge = nslots;  // one lane until amrmg_serve_set_lanes
  for (int q = 0; q < nslots; q++) e->workers.emplace_back([p = e.get()] { p->worker(); });
  *out = e.release();
  API_END
}
int amrmg_serve_set_params(amrmg_serve* e, const int* sp, const double* rp) {
  API_BEGIN
  std::lock_guard<std::mutex> lk(e->mu);
  std::memcpy(e->sp, sp, sizeof e->sp);
  std::memcpy(e->rp, rp, sizeof e->rp);
  API_END
}
int amrmg_serve_set_lanes(amrmg_serve* e, int max_large, long long large_cells) {
  API_BEGIN
  if (max_large < 1) throw std::runtime_error("serve: at least one large request must be able to run");
  if (large_cells < 0) throw std::runtime_error("serve: the large-request threshold must not be negative");
  std::lock_guard<std::mutex> lk(e->mu);
  e->maxLarge = max_large;
  e->largeCells = large_cells;
  for (ServeReq* q : e->queue) q->large = q->cells >= large_cells;
  e->cv.notify_all();
  API_END
}
int amrmg_serve_submit(amrmg_serve* e, const int* ispec, const double* rspec, const double* rhs, long long* id) {
  API_BEGIN
  Box dom;
  BCSpec bc;
  double dx0, a, b;
  std::vector<DBL> g = parse_spec(ispec, rspec, dom, bc, dx0, a, b);
  int nint = 13 + g.size();
  long long ncell = 0;
  for (auto& d : g) {
    nint += 6 * d.size();
    ncell += d.numCells();
  }
  std::unique_ptr<ServeReq> r(new ServeReq());
  r->ispec.assign(ispec, ispec + nint);
  r->rspec.assign(rspec, rspec + 9);
  r->rhs.assign(rhs, rhs + ncell);
  r->key.assign((const char*)ispec, sizeof(int) * nint);
  r->key.append((const char*)rspec, sizeof(double) * 9);
  r->bytes = serve_footprint(g, dom, dx0, e->precision, bc);
  r->cells = ncell;
  std::lock_guard<std::mutex> lk(e->mu);
  r->large = ncell >= e->largeCells;
  auto m = e->measured.find(r->key);
  if (m != e->measured.end()) r->bytes = m->second;
  if (r->bytes > e->budget) throw std::runtime_error("serve: request larger than the HBM budget");
  r->id = e->nextId++;
  r->tArrival = serve_now(e->t0);
  *id = r->id;
  e->queue.push_back(r.get());
  e->reqs[r->id] = std::move(r);
  e->cv.notify_all();
  API_END
}
// blocks until request id is done; phi (valid-packed, may be null), times =
// {arrival, admission, completion} in seconds since the engine started, info =
// {iterations, exit status}; the request is then forgotten
int amrmg_serve_wait(amrmg_serve* e, long long id, double* phi, double* times, int* info) {
  API_BEGIN
  std::unique_lock<std::mutex> lk(e->mu);
  auto it = e->reqs.find(id);
  if (it == e->reqs.end()) throw std::runtime_error("serve: unknown request id");
  ServeReq* r = it->second.get();
  if (r->taken) throw std::runtime_error("serve: another thread is already waiting for this request");
  r->taken = true;
  e->cv.wait(lk, [&] { return r->done; });
  if (!r->error.empty()) {
    std::string m = r->error;
    e->reqs.erase(it);
    throw std::runtime_error("serve: request failed: " + m);
  }
  if (phi) std::memcpy(phi, r->phi.data(), sizeof(double) * r->phi.size());
  if (times) { times[0] = r->tArrival; times[1] = r->tAdmit; times[2] = r->tDone; }
  if (info) { info[0] = r->iters; info[1] = r->exitStatus; }
  e->reqs.erase(it);
  API_END
}
// {max bytes in use (running + idle hierarchies), hierarchies created, reused,
//  evicted, most requests running at once}
int amrmg_serve_stats(amrmg_serve* e, long long* stats, int n) {
  API_BEGIN
  std::lock_guard<std::mutex> lk(e->mu);
  const long long v[5] = {e->maxBytes, e->created, e->reused, e->evicted, e->maxRunning};
  for (int q = 0; q < n && q < 5; q++) stats[q] = v[q];
  API_END
}
int amrmg_serve_destroy(amrmg_serve* e) {
  API_BEGIN
  {
    std::lock_guard<std::mutex> lk(e->mu);
    e->stop = true;
    e->cv.notify_all();
  }
  for (auto& t : e->workers) t.join();
  delete e;
  API_END
}

int amrmg_copy_probe(amrmg_ctx* ctx, long long nbytes, int iters, double* ms_out) {
  API_BEGIN
  if (nbytes <= 0 || iters <= 0) throw std::runtime_error("copy probe: nbytes and iters must be positive");
  *ms_out = copy_probe(ctx->c.get(), nbytes, iters);
  API_END
}

int amrmg_plan_check(const int* ispec, const double* rspec, long long* stats, int nstats) {
  return amrmg_plan_check_ratios(ispec, rspec, nullptr, stats, nstats);
}
int amrmg_plan_check_ratios(const int* ispec, const double* rspec, const int* ref_ratios, long long* stats,
                            int nstats) {
  API_BEGIN
  Context dry(0, true);
  Box dom;
  double dx0, a, b;
  BCSpec bc{};
  std::vector<DBL> g = parse_spec(ispec, rspec, dom, bc, dx0, a, b);
  DeviceHierarchy H;
  H.define(&dry, g, dom, dx0, true, &bc, ref_ratios);
  AMRPoissonOpFactory<double> fac;
  fac.define(&H, a, b, bc);
  BiCGStabSolver<LevelData<double>> bottom;
  AMRMultiGrid<double> amg;
  amg.define(fac, &bottom, H.nlev());
  long long st[8] = {H.nlev(), H.mgDepth(), 0, 0, 0, 0, 0, 0};
  for (auto& op : amg.m_op) {
    st[2] += (long long)op->m_ghost.h_patches.size();
    st[3] += op->n_cfc;
    st[4] += op->n_cff;
    st[5] += op->n_rcell;
    st[6] += op->m_lay->ntiles;
  }
  for (auto& l : H.mgLay) st[6] += l->ntiles;
  for (int l = 0; l < H.nlev(); l++) st[7] += H.amrLay[l]->valid_cells;
  // the multigrid operators of the levels above the base (solves with l_base > 0)
  for (int l = 1; l < H.nlev(); l++) {
    auto ops = fac.MGnewOpsLevel(l);
    for (auto& o : ops) st[2] += (long long)static_cast<AMRPoissonOp<double>*>(o.get())->m_ghost.h_patches.size();
  }
  for (int q = 0; q < nstats && q < 8; q++) stats[q] = st[q];
  API_END
}

int amrmg_profile_enable(amrmg_ctx* ctx, int class_mask) {
  API_BEGIN
  profiler_enable(ctx->c.get(), (unsigned)class_mask);
  API_END
}
int amrmg_profile_reset(amrmg_ctx* ctx) {
  API_BEGIN
  profiler_reset(ctx->c.get());
  API_END
}
int amrmg_profile_read(amrmg_ctx* ctx, int kind, double* ms, double* bytes, long long* launches) {
  API_BEGIN
  if (!profiler_read(ctx->c.get(), kind, ms, bytes, launches)) throw std::runtime_error("profiler off or bad class");
  API_END
}
int amrmg_profile_read_units(amrmg_ctx* ctx, int kind, double* units) {
  API_BEGIN
  double ms, bytes;
  long long n;
  if (!profiler_read(ctx->c.get(), kind, &ms, &bytes, &n, units)) throw std::runtime_error("profiler off or bad class");
  API_END
}

}  // extern "C"
