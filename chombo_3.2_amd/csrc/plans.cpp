// Host plan builders (see plans.hpp).  Each builder states the reference
// routine whose bookkeeping it reproduces.
#include "plans.hpp"

#include <algorithm>

#include <cmath>
#include <cstdint>
#include <map>
#include <stdexcept>

namespace amrmg {

// ---------------------------------------------------------------------------
// BoxFinder
// ---------------------------------------------------------------------------
BoxFinder::BoxFinder(const DBL& d) : dbl_(&d) {
  if (d.size() == 0) return;
  Box g = d[0];
  int mn = 1 << 30;
  for (auto& b : d.boxes) {
    for (int q = 0; q < 3; q++) {
      g.lo[q] = std::min(g.lo[q], b.lo[q]);
      g.hi[q] = std::max(g.hi[q], b.hi[q]);
      mn = std::min(mn, b.len(q));
    }
  }
  B_ = std::max(1, mn);
  grid_ = g;
  for (int q = 0; q < 3; q++) n_[q] = (g.len(q) + B_ - 1) / B_;
  long long nb = (long long)n_[0] * n_[1] * n_[2];
  std::vector<std::vector<int>> tmp((size_t)nb);
  for (int idx = 0; idx < d.size(); idx++) {
    const Box& b = d[idx];
    IV lo, hi;
    for (int q = 0; q < 3; q++) {
      lo[q] = (b.lo[q] - g.lo[q]) / B_;
      hi[q] = (b.hi[q] - g.lo[q]) / B_;
    }
    for (int k = lo[2]; k <= hi[2]; k++)
      for (int j = lo[1]; j <= hi[1]; j++)
        for (int i = lo[0]; i <= hi[0]; i++) tmp[((size_t)k * n_[1] + j) * n_[0] + i].push_back(idx);
  }
  start_.assign((size_t)nb + 1, 0);
  for (size_t q = 0; q < tmp.size(); q++) start_[q + 1] = start_[q] + (int)tmp[q].size();
  list_.reserve(start_.back());
  for (auto& v : tmp) list_.insert(list_.end(), v.begin(), v.end());
}
int BoxFinder::bucket(int i, int j, int k) const {
  int a = (i - grid_.lo[0]) / B_, b = (j - grid_.lo[1]) / B_, c = (k - grid_.lo[2]) / B_;
  return (c * n_[1] + b) * n_[0] + a;
}
int BoxFinder::find(const IV& p) const {
  if (dbl_->size() == 0 || !grid_.contains(p)) return -1;
  int bk = bucket(p[0], p[1], p[2]);
  for (int q = start_[bk]; q < start_[bk + 1]; q++)
    if ((*dbl_)[list_[q]].contains(p)) return list_[q];
  return -1;
}
std::vector<int> BoxFinder::overlapping(const Box& r0) const {
  std::vector<int> out;
  if (dbl_->size() == 0) return out;
  Box r = r0 & grid_;
  if (r.empty()) return out;
  IV lo, hi;
  for (int q = 0; q < 3; q++) {
    lo[q] = (r.lo[q] - grid_.lo[q]) / B_;
    hi[q] = (r.hi[q] - grid_.lo[q]) / B_;
  }
  for (int k = lo[2]; k <= hi[2]; k++)
    for (int j = lo[1]; j <= hi[1]; j++)
      for (int i = lo[0]; i <= hi[0]; i++) {
        int bk = (k * n_[1] + j) * n_[0] + i;
        for (int q = start_[bk]; q < start_[bk + 1]; q++)
          if (intersects((*dbl_)[list_[q]], r0)) out.push_back(list_[q]);
      }
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

static Patch make_patch(int kind, int dst, int src, int dir, int side, const Box& r, double val) {
  Patch p;
  p.kind = kind; p.dst = dst; p.src = src; p.dir = dir; p.side = side; p.val = val; p.voff = -1;
  for (int q = 0; q < 3; q++) { p.lo[q] = r.lo[q]; p.hi[q] = r.hi[q]; p.sh[q] = 0; }
  return p;
}

// ---------------------------------------------------------------------------
// ghost plan: Copier::exchangeDefine + trimEdges (Copier.cpp:1464-1541,142-177)
// restricted to face ghosts; ConstBCFunction (BCFunc.cpp:58-158); CFRegion /
// CFIVS (CFRegion.cpp:16-37, CFIVS.cpp:236-278) feeding homogeneousCFInterp
// (AMRPoissonOp.cpp:1598-1671).  Edge and corner ghosts are never read by the
// 7-point stencil and are not filled.
// ---------------------------------------------------------------------------
// With merged storage (one array per level) an exchange copies every cell onto
// itself, so no exchange patch is emitted.  In a periodic direction a face ghost
// outside the domain is filled from the periodic image, the boxes shifted by the
// domain length (the NeighborIterator images of Copier::exchangeDefine); such a
// level is stored per box.
ops::GhostPlan build_ghost_plan(const DBL& dbl, const Box& domain, const BCSpec& bc, bool withCF, bool merged) {
  ops::GhostPlan P;
  BoxFinder bf(dbl);
  for (int b = 0; b < dbl.size(); b++) {
    const Box& bx = dbl[b];
    for (int dir = 0; dir < 3; dir++)
      for (int side = 0; side < 2; side++) {
        Box R = adjCell(bx, dir, side, 1);
        int vend = side ? bx.hi[dir] : bx.lo[dir];
        int dend = side ? domain.hi[dir] : domain.lo[dir];
        if (vend == dend && bc.type[side][dir] == 3) {
          if (merged) throw std::runtime_error("ghost plan: a periodic face on a merged level");
          const int len = domain.len(dir), shift = side ? -len : len;
          Box img = R;
          img.lo[dir] += shift;
          img.hi[dir] += shift;
          for (int n : bf.overlapping(img)) {
            Box r = img & dbl[n];
            r.lo[dir] -= shift;
            r.hi[dir] -= shift;
            Patch pt = make_patch(PK_EXCHANGE, b, n, dir, side, r, 0.0);
            pt.sh[dir] = shift;
            P.h_patches.push_back(pt);
          }
          continue;
        }
        if (vend == dend) {
          int t = bc.type[side][dir];
          const int kind = t == 1 ? PK_DIRICHLET : (t == 2 ? PK_DIRICHLET2 : PK_NEUMANN);
          P.h_patches.push_back(make_patch(kind, b, -1, dir, side, R, bc.val[side][dir]));
          continue;
        }
        IVS cf(R & domain);
        for (int n : bf.overlapping(R)) {
          if (n == b) continue;
          Box r = R & dbl[n];
          if (!merged) P.h_patches.push_back(make_patch(PK_EXCHANGE, b, n, dir, side, r, 0.0));
          cf.subtract(dbl[n]);
        }
        if (!withCF || cf.numPts() == 0) continue;
        int kind = (bx.len(dir) == 1) ? PK_CFHOMOLIN : PK_CFHOMO;
        Box mb = cf.minBox();
        if (mb.npts() == cf.numPts()) {
          P.h_patches.push_back(make_patch(kind, b, -1, dir, side, mb, 0.0));
        } else {
          cf.forEach([&](const IV& p) { P.h_patches.push_back(make_patch(kind, b, -1, dir, side, pointBox(p), 0.0)); });
        }
      }
  }
  return P;
}

// DiriBC / NeumBC with a BCValueHolder (BCFunc.cpp:217-258, 291-354): the value of
// ghost cell ivTo is the holder's at the centre of the domain face of the valid cell
// next to it, getDomainFacePosition (BCFunc.cpp:187-202): dx*(iv+0.5) in every
// direction, moved half a cell outward in dir.
void attach_bc_values(ops::GhostPlan& P, double dx, const BCSpec& bc, int ncomp) {
  P.h_bcv.clear();
  P.nbcv = 0;
  if (ncomp < 1) throw std::runtime_error("BC values: ncomp must be at least 1");
  long long n = 0;  // face cells with a value (per component)
  for (Patch& p : P.h_patches) {
    p.voff = -1;
    if (!bc.fn || (p.kind != PK_DIRICHLET && p.kind != PK_DIRICHLET2 && p.kind != PK_NEUMANN)) continue;
    const long long np = (long long)(p.hi[0] - p.lo[0] + 1) * (p.hi[1] - p.lo[1] + 1) * (p.hi[2] - p.lo[2] + 1);
    if (n + np > (long long)INT32_MAX) throw std::runtime_error("BC values: too many face cells");
    p.voff = (int)n;
    n += np;
  }
  P.nbcv = n;
  P.h_bcv.assign((size_t)(n * ncomp), 0.0);
  std::vector<double> v((size_t)ncomp);
  for (const Patch& p : P.h_patches) {
    if (p.voff < 0) continue;
    long long q = p.voff;
    const int isign = p.side ? 1 : -1;
    for (int k = p.lo[2]; k <= p.hi[2]; k++)
      for (int j = p.lo[1]; j <= p.hi[1]; j++)
        for (int i = p.lo[0]; i <= p.hi[0]; i++, q++) {
          int iv[3] = {i, j, k};
          iv[p.dir] -= isign;  // ivClose
          double pos[3];
          for (int d = 0; d < 3; d++) pos[d] = dx * (iv[d] + 0.5);
          pos[p.dir] += 0.5 * double(isign) * dx;
          int dir = p.dir, side = p.side;
          std::fill(v.begin(), v.end(), 0.0);
          bc.fn(pos, &dir, &side, v.data(), bc.user);  // one value per component
          for (int c = 0; c < ncomp; c++) P.h_bcv[(size_t)(c * n + q)] = v[(size_t)c];
        }
  }
}

ops::MotionPlan build_copier(const DBL& src, const DBL& dst) {
  ops::MotionPlan P;
  BoxFinder bf(src);
  for (int t = 0; t < dst.size(); t++)
    for (int f : bf.overlapping(dst[t])) {
      Box r = dst[t] & src[f];
      Motion m;
      m.from = f; m.to = t;
      for (int q = 0; q < 3; q++) { m.lo[q] = r.lo[q]; m.hi[q] = r.hi[q]; }
      P.h_items.push_back(m);
    }
  return P;
}

// ---------------------------------------------------------------------------
// QuadCFInterp tables: CFStencil / QuadCFStencil::define (QuadCFStencil.cpp:455-660),
// buildStencils (:662-918), DerivStencil::accumulate, getPhiStar
// (QuadCFInterp.cpp:570-764), FORT_PHISTAR / FORT_QUADINTERP (QuadCFInterpF.ChF).
// ---------------------------------------------------------------------------
namespace {
struct Sten {
  std::vector<IV> iv;
  std::vector<double> w;
  void acc(const IV& p, double v) {  // DerivStencil::accumulate
    for (size_t q = 0; q < iv.size(); q++)
      if (iv[q] == p) { w[q] += v; return; }
    iv.push_back(p);
    w.push_back(v);
  }
};
// addFabToSten over a small box with values given by fn(cell), BoxIterator order
template <class F>
void addBox(Sten& s, const Box& b, F fn) {
  for (int k = b.lo[2]; k <= b.hi[2]; k++)
    for (int j = b.lo[1]; j <= b.hi[1]; j++)
      for (int i = b.lo[0]; i <= b.hi[0]; i++) s.acc(IV{i, j, k}, fn(IV{i, j, k}));
}
struct QStencils {
  bool dropOrd = false;
  Sten mixed;
  Sten first[3], second[3];
};
}  // namespace

CFPlanHost build_cf_plan(const DBL& fine, const DeviceLayout& fineLay, const DBL& coarse,
                         const DeviceLayout& coarseLay, const Box& fineDomain, double dxFine, int ref) {
  CFPlanHost out;
  if (ref != 2 && ref != 4) throw std::runtime_error("build_cf_plan: refinement ratio 2 or 4 only");
  // QuadCFInterp::define: no interpolation when the fine level covers the coarse one
  if (fine.numCells() / (ref * ref * ref) - coarse.numCells() == 0) return out;
  BoxFinder ff(fine), cfind(coarse);
  DBL cfineDbl = fine.coarsened(ref);
  BoxFinder cff(cfineDbl);
  const Box cdom = coarsen(fineDomain, ref);
  const double dxf = dxFine, dxc = ref * dxf;
  if (coarseLay.slab >= (1LL << 31)) throw std::runtime_error("coarse level too large for 32-bit CF offsets");

  auto coarseOff = [&](const IV& p) -> int {
    int b = cfind.find(p);
    if (b < 0) throw std::runtime_error("coarse-fine stencil outside the coarse level (proper nesting violated)");
    return (int)coarseLay.index(b, p[0], p[1], p[2]);
  };

  for (int fb = 0; fb < fine.size(); fb++) {
    const Box& grid = fine[fb];
    const Box cgrid = coarsen(grid, ref);
    const Box bufBox = grow(cgrid, 2);  // m_coarBoxes: coarsened fine box grown by 2
    for (int idir = 0; idir < 3; idir++)
      for (int side = 0; side < 2; side++) {
        // CFStencil: fine IVS and coarse IVS
        Box edge = adjCell(grid, idir, side, 1) & fineDomain;
        if (edge.empty()) continue;
        IVS fineIVS(edge);
        for (int n : ff.overlapping(edge)) fineIVS.subtract(fine[n]);
        if (fineIVS.numPts() == 0) continue;
        IVS coarIVS = fineIVS.coarsened(ref);
        coarIVS.intersect(coarIVS.minBox() & cdom);
        if (coarIVS.numPts() == 0) continue;
        Box packedBox = fineIVS.minBox();
        const bool packed = packedBox.npts() == fineIVS.numPts();

        int vt[2], nv = 0;
        for (int d = 0; d < 3; d++)
          if (d != idir) vt[nv++] = d;
        Box g2 = adjCell(cgrid, idir, side, 1), g1 = g2;
        for (int t = 0; t < 2; t++) { g2 = growDir(g2, vt[t], 2); g1 = growDir(g1, vt[t], 1); }
        g2 = g2 & cdom;
        g1 = g1 & cdom;
        // ivsAllGood: coarse cells of g2 covered by the coarse level and not by the fine one
        IVS allGood(g2);
        for (int n : cff.overlapping(g2)) allGood.subtract(cfineDbl[n]);
        {
          IVS eoe = allGood;
          for (int n : cfind.overlapping(g2)) eoe.subtract(coarse[n]);
          allGood.subtract(eoe);
        }
        IVS standard = allGood;
        standard.intersect(g1);
        standard.erode(vt[0]);
        standard.erode(vt[1]);

        // stencils for the non-standard coarse cells
        std::map<IV, QStencils> quadd;
        coarIVS.forEach([&](const IV& p) {
          if (!standard.has(p)) quadd[p];
        });
        const int t1 = vt[0], t2 = vt[1];
        for (auto& kv : quadd) {
          const IV& iv = kv.first;
          QStencils& st = kv.second;
          // mixed: four 2x2 quadrant stencils; lo corner offsets -e_t1, 0, -e_t2, -e_t1-e_t2
          const int offs[4][2] = {{-1, 0}, {0, 0}, {0, -1}, {-1, -1}};
          int imixed = 0;
          for (int c = 0; c < 4; c++) {
            Box fbx = pointBox(iv);
            fbx.lo[t1] += offs[c][0]; fbx.hi[t1] = fbx.lo[t1] + 1;
            fbx.lo[t2] += offs[c][1]; fbx.hi[t2] = fbx.lo[t2] + 1;
            if (!allGood.hasBox(fbx)) continue;
            IV lo = fbx.lo;
            addBox(st.mixed, fbx, [&](const IV& p) {
              int r1 = p[t1] - lo[t1], r2 = p[t2] - lo[t2];
              return (r1 + r2 == 1) ? 1.0 : -1.0;
            });
            imixed++;
          }
          if (imixed > 0)
            for (auto& w : st.mixed.w) w /= (double)imixed;
          st.dropOrd = (imixed == 0);
        }
        for (int ivec = 0; ivec < 2; ivec++) {
          const int d = vt[ivec];
          for (auto& kv : quadd) {
            const IV& iv = kv.first;
            QStencils& st = kv.second;
            if (st.dropOrd) continue;
            IV m = iv, p = iv, m2 = iv, p2 = iv;
            m[d] -= 1; p[d] += 1; m2[d] -= 2; p2[d] += 2;
            Box bsten = growDir(pointBox(iv), d, 1);
            if (allGood.hasBox(bsten)) {
              addBox(st.second[d], bsten, [&](const IV& q) { return q == iv ? -2.0 : 1.0; });
              addBox(st.first[d], bsten, [&](const IV& q) { return q == iv ? 0.0 : (q == m ? -0.5 : 0.5); });
              continue;
            }
            Box bs = shift(bsten, d, 1);
            if (allGood.hasBox(bs)) {
              addBox(st.second[d], bs, [&](const IV& q) { return q == p ? -2.0 : 1.0; });
              addBox(st.first[d], bs, [&](const IV& q) { return q == iv ? -3.0 / 2.0 : (q == p ? 4.0 / 2.0 : -1.0 / 2.0); });
              continue;
            }
            bs = shift(bs, d, -2);
            if (allGood.hasBox(bs)) {
              addBox(st.second[d], bs, [&](const IV& q) { return q == m ? -2.0 : 1.0; });
              addBox(st.first[d], bs, [&](const IV& q) { return q == iv ? 3.0 / 2.0 : (q == m ? -4.0 / 2.0 : 1.0 / 2.0); });
              continue;
            }
            st.dropOrd = true;
            if (allGood.has(p)) addBox(st.first[d], Box(iv, p), [&](const IV& q) { return q == iv ? -1.0 : 1.0; });
            else if (allGood.has(m)) addBox(st.first[d], Box(m, iv), [&](const IV& q) { return q == iv ? 1.0 : -1.0; });
            else addBox(st.first[d], pointBox(iv), [&](const IV&) { return 0.0; });
          }
        }

        // coarse records
        int you1, you2;
        if (idir == 0) { you1 = 1; you2 = 2; }
        else if (idir == 1) { you1 = 0; you2 = 2; }
        else { you1 = 0; you2 = 1; }
        int itran1, itran2;  // computeMixedDerivative naming
        if (idir == 0) { itran2 = 1; itran1 = 2; }
        else if (idir == 1) { itran2 = 0; itran1 = 2; }
        else { itran2 = 0; itran1 = 1; }
        std::map<IV, int> recOf;
        coarIVS.forEach([&](const IV& pc) {
          CFCoarse r;
          r.pc = coarseOff(pc);
          r.gen = -1;
          r.std = standard.has(pc) ? 1 : 0;
          if (r.std) {
            IV a = pc, b = pc, c = pc, e = pc;
            a[you1] -= 1; b[you1] += 1; c[you2] -= 1; e[you2] += 1;
            r.nbr[0] = coarseOff(a); r.nbr[1] = coarseOff(b); r.nbr[2] = coarseOff(c); r.nbr[3] = coarseOff(e);
            IV ur = pc, ul = pc, lr = pc, ll = pc;
            ur[itran2] += 1; ur[itran1] += 1;
            ul[itran2] -= 1; ul[itran1] += 1;
            lr[itran2] += 1; lr[itran1] -= 1;
            ll[itran2] -= 1; ll[itran1] -= 1;
            r.nbr[4] = coarseOff(ur); r.nbr[5] = coarseOff(ul); r.nbr[6] = coarseOff(lr); r.nbr[7] = coarseOff(ll);
          } else {
            const QStencils& st = quadd.at(pc);
            CFGen g;
            std::memset(&g, 0, sizeof g);
            // order: first(you1), second(you1), first(you2), second(you2), mixed
            const Sten* sl[5] = {&st.first[you1], &st.second[you1], &st.first[you2], &st.second[you2], &st.mixed};
            const bool zeroByDrop[5] = {false, st.dropOrd, false, st.dropOrd, st.dropOrd};
            for (int q = 0; q < 5; q++) {
              g.kind[q] = DK_GEN;
              g.n[q] = 0;
              if (zeroByDrop[q]) { g.kind[q] = DK_ZERO; continue; }
              const Sten& s = *sl[q];
              if (s.iv.size() > 9) throw std::runtime_error("CF stencil too long");
              bool keepzer = false;
              for (size_t t = 0; t < s.iv.size(); t++)
                if (!bufBox.contains(s.iv[t])) keepzer = true;
              if (keepzer) { g.kind[q] = DK_ZERO; continue; }
              for (size_t t = 0; t < s.iv.size(); t++) {
                int b = cfind.find(s.iv[t]);
                if (b < 0) {
                  if (s.w[t] != 0.0) throw std::runtime_error("CF stencil reads an uncovered coarse cell");
                  g.off[q][g.n[q]] = r.pc;
                } else {
                  g.off[q][g.n[q]] = (int)coarseLay.index(b, s.iv[t][0], s.iv[t][1], s.iv[t][2]);
                }
                g.w[q][g.n[q]] = s.w[t];
                g.n[q]++;
              }
            }
            r.gen = (int)out.gens.size();
            out.gens.push_back(g);
          }
          recOf[pc] = (int)out.coarse.size();
          out.coarse.push_back(r);
        });

        // fine records
        const int ihilo = side ? 1 : -1;
        Box psb = shift(packedBox, idir, ihilo);   // phistar box when packed
        IV ivs = psb.lo;
        IV civ{floordiv(ivs[0], ref), floordiv(ivs[1], ref), floordiv(ivs[2], ref)};
        fineIVS.forEach([&](const IV& g) {
          CFFine f;
          f.g = fineLay.index(fb, g[0], g[1], g[2]);
          IV a1 = g, a2 = g;
          a1[idir] -= ihilo; a2[idir] -= 2 * ihilo;
          f.a1 = fineLay.index(fb, a1[0], a1[1], a1[2]);
          f.a2 = fineLay.index(fb, a2[0], a2[1], a2[2]);
          f.packed = packed ? 1 : 0;
          IV pc{floordiv(g[0], ref), floordiv(g[1], ref), floordiv(g[2], ref)};
          if (packed) {
            IV ir{g[0] - ivs[0], g[1] - ivs[1], g[2] - ivs[2]};
            IV ic{ir[0] / ref, ir[1] / ref, ir[2] / ref};  // Fortran truncating division
            IV cg{ic[0] + civ[0], ic[1] + civ[1], ic[2] + civ[2]};
            if (!(cg == pc)) throw std::runtime_error("PHISTAR coarse index mismatch");
            double dxcf = ref * dxf;
            double xf1 = (ir[you1] + 0.5) * dxf, xc1 = (ic[you1] + 0.5) * dxcf;
            double xf2 = (ir[you2] + 0.5) * dxf, xc2 = (ic[you2] + 0.5) * dxcf;
            f.x1 = xf1 - xc1;
            f.x2 = xf2 - xc2;
          } else {
            double xf = (g[you1] + 0.5) * dxf, xc = (pc[you1] + 0.5) * dxc;
            f.x1 = xf - xc;
            xf = (g[you2] + 0.5) * dxf;
            xc = (pc[you2] + 0.5) * dxc;
            f.x2 = xf - xc;
          }
          auto it = recOf.find(pc);
          if (it == recOf.end()) throw std::runtime_error("CF fine cell without coarse record");
          f.coarse = it->second;
          out.fine.push_back(f);
        });
      }
  }
  return out;
}

// ---------------------------------------------------------------------------
// Reflux tables: LevelFluxRegister::define (LevelFluxRegister.cpp:101-262),
// incrementCoarse (:373-434), incrementFine (:452-586), reflux (:649-670),
// AMRPoissonOp::reflux / getFlux (AMRPoissonOp.cpp:1117-1201, 1797-1851).
// Coarse cells covered by the finer level are skipped: the reference overwrites
// (assignCopier) or zeroes (zeroCovered) them wherever they are used.
// ---------------------------------------------------------------------------
RefluxPlanHost build_reflux_plan(const DBL& coarse, const DeviceLayout& coarseLay, const DBL& fine,
                                 const DeviceLayout& fineLay, const Box& coarseDomain, int ref) {
  RefluxPlanHost out;
  if (ref != 2 && ref != 4) throw std::runtime_error("reflux: refinement ratio 2 or 4 only");
  DBL cfine = fine.coarsened(ref);
  BoxFinder cff(cfine), cfind(coarse);
  (void)coarseDomain;
  struct CellLists {
    std::vector<RefluxCoarse> c;
    std::vector<RefluxFine> f;
    std::vector<RefluxFace> cface, fface;
  };
  std::map<std::pair<int, long long>, CellLists> cells;  // (coarse box, cell index) -> contributions
  auto covered = [&](const IV& p) { return cff.find(p) >= 0; };

  // incrementCoarse: for each coarse box, direction, side, in reference order
  for (int C = 0; C < coarse.size(); C++) {
    const Box& cb = coarse[C];
    std::vector<Box> loc[3][2];
    for (int F : cff.overlapping(grow(cb, 1))) {
      const Box& fbx = cfine[F];
      for (int i = 0; i < 3; i++) {
        Box lo = adjCell(fbx, i, 0, 1), hi = adjCell(fbx, i, 1, 1);
        if (intersects(cb, lo)) loc[i][0].push_back(lo & cb);
        if (intersects(cb, hi)) loc[i][1].push_back(hi & cb);
      }
    }
    for (int idir = 0; idir < 3; idir++)
      for (int side = 0; side < 2; side++)
        for (const Box& bx : loc[idir][side])
          for (int k = bx.lo[2]; k <= bx.hi[2]; k++)
            for (int j = bx.lo[1]; j <= bx.hi[1]; j++)
              for (int i = bx.lo[0]; i <= bx.hi[0]; i++) {
                IV c{i, j, k};
                if (covered(c)) continue;
                IV n = c;
                if (side == 0) n[idir] += 1;  // shiftHalf: lo-side cell c <-> node c+1
                IV m1 = n;
                m1[idir] -= 1;
                RefluxCoarse x;
                x.a = coarseLay.index(C, n[0], n[1], n[2]);
                x.b = coarseLay.index(C, m1[0], m1[1], m1[2]);
                x.sign = side ? -1.0 : 1.0;
                RefluxFace rf{C, idir, {{n[0], n[1], n[2]}}};
                CellLists& cl = cells[{C, coarseLay.index(C, i, j, k)}];
                cl.c.push_back(x);
                cl.cface.push_back(rf);
              }
  }
  // fine registers copied through the reverse ghost copier, in its sorted order
  struct Item { int F, C; Box r; };
  std::vector<Item> items;
  for (int F = 0; F < cfine.size(); F++)
    for (int C : cfind.overlapping(grow(cfine[F], 1))) items.push_back(Item{F, C, grow(cfine[F], 1) & coarse[C]});
  std::sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.r < b.r; });
  for (const Item& it : items) {
    const Box& fbx = cfine[it.F];
    const Box fgrid = refine(fbx, ref);
    for (int k = it.r.lo[2]; k <= it.r.hi[2]; k++)
      for (int j = it.r.lo[1]; j <= it.r.hi[1]; j++)
        for (int i = it.r.lo[0]; i <= it.r.hi[0]; i++) {
          IV c{i, j, k};
          if (covered(c)) continue;
          // which face ghost layer of fbx holds c?
          int dir = -1, side = -1, nout = 0;
          for (int d = 0; d < 3; d++) {
            if (c[d] < fbx.lo[d]) { nout++; dir = d; side = 0; }
            else if (c[d] > fbx.hi[d]) { nout++; dir = d; side = 1; }
          }
          if (nout != 1) continue;  // interior or edge/corner ghost: register holds zero
          // sign(side)*a_scale/denom, denom = nRef^(D-1) (LevelFluxRegister.cpp:468-481)
          double denom = 1.0;
          for (int d = 0; d < 3; d++)
            if (d != dir) denom *= ref;
          const double scale = (side ? 1.0 : -1.0) * 1.0 / denom;
          // fine cells just outside fgrid under c, INCREMENTFINE loop order
          Box fcell = refine(pointBox(c), ref);
          std::vector<IV> nodes;
          for (int kk = fcell.lo[2]; kk <= fcell.hi[2]; kk++)
            for (int jj = fcell.lo[1]; jj <= fcell.hi[1]; jj++)
              for (int ii = fcell.lo[0]; ii <= fcell.hi[0]; ii++) {
                IV q{ii, jj, kk};
                if (side == 0 && q[dir] != fgrid.lo[dir] - 1) continue;
                if (side == 1 && q[dir] != fgrid.hi[dir] + 1) continue;
                IV n = q;
                if (side == 0) n[dir] += 1;
                nodes.push_back(n);
              }
          if ((int)nodes.size() != ref * ref) throw std::runtime_error("reflux: fine face must have ratio^2 cells");
          // the register as records of four fluxes, first..last
          CellLists& cl = cells[{it.C, coarseLay.index(it.C, i, j, k)}];
          for (size_t c0 = 0; c0 < nodes.size(); c0 += 4) {
            RefluxFine x;
            RefluxFace xf{it.F, dir, {}};
            x.scale = scale;
            x.first = c0 == 0;
            x.last = c0 + 4 == nodes.size();
            for (int t = 0; t < 4; t++) {
              const IV& n = nodes[c0 + t];
              IV m1 = n;
              m1[dir] -= 1;
              x.a[t] = fineLay.index(it.F, n[0], n[1], n[2]);
              x.b[t] = fineLay.index(it.F, m1[0], m1[1], m1[2]);
              for (int q3 = 0; q3 < 3; q3++) xf.iv[t][q3] = n[q3];
            }
            cl.f.push_back(x);
            cl.fface.push_back(xf);
          }
        }
  }
  for (auto& kv : cells) {
    RefluxCell rc;
    rc.u = kv.first.second;
    rc.sy = coarseLay.h_boxes[kv.first.first].sy;
    rc.sz = coarseLay.h_boxes[kv.first.first].sz;
    rc.c0 = (int)out.coarse.size();
    rc.nc = (int)kv.second.c.size();
    rc.f0 = (int)out.fine.size();
    rc.nf = (int)kv.second.f.size();
    out.coarse.insert(out.coarse.end(), kv.second.c.begin(), kv.second.c.end());
    out.fine.insert(out.fine.end(), kv.second.f.begin(), kv.second.f.end());
    out.coarseFace.insert(out.coarseFace.end(), kv.second.cface.begin(), kv.second.cface.end());
    out.fineFace.insert(out.fineFace.end(), kv.second.fface.begin(), kv.second.fface.end());
    out.cells.push_back(rc);
  }
  return out;
}

}  // namespace amrmg
