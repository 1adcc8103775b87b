/*
 * amr_oracle.c -- plain-C restatement of the FP64 path of Chombo 3.2's
 * AMRPoissonOp under AMRMultiGrid / MultiGrid with a BiCGStab bottom solver.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  It is loaded by tests/, by the smoke
 * check and by the CPU-baseline leg of bench.py, and never by the product.
 *
 * Every routine names the reference source it restates.  Arithmetic follows the
 * reference's expression order term by term and the file is compiled with
 * -ffp-contract=off, so that a restated kernel rounds exactly as the Fortran /
 * C++ it follows (Fortran evaluates a+b+c as (a+b)+c and never reassociates).
 * Paths are relative to the Chombo 3.2 source tree.
 *
 * Parity pins (see tests/test_oracle_*.py): each kernel is checked against
 * closed forms (operator exact on quadratics, discrete eigenfunctions, GSRB fixed
 * point, AVERAGE/PROLONG of polynomials), the reference's own known-answer tests
 * (lib/test/AMRTools/quadCFInterpTest.cpp quadratic reproduction at 1e-12,
 * lib/test/AMRElliptic/testAMRPoissonOp.H truncation), dense direct solves of
 * small systems, conservation of the refluxed composite operator and adjointness.
 */
#include "amr_oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static char g_err[512];
const char* oracle_last_error(void) { return g_err; }
static void set_err(const char* m) { snprintf(g_err, sizeof g_err, "%s", m); }

/* ======================================================================= */
/* Box (lib/src/BoxTools/Box.H, IntVect.H)                                 */
/* ======================================================================= */
typedef struct { int lo[3], hi[3]; } Box;

static int b_empty(const Box* b) {
  for (int d = 0; d < 3; d++) if (b->lo[d] > b->hi[d]) return 1;
  return 0;
}
static long b_npts(const Box* b) {
  if (b_empty(b)) return 0;
  return (long)(b->hi[0]-b->lo[0]+1)*(b->hi[1]-b->lo[1]+1)*(b->hi[2]-b->lo[2]+1);
}
static int b_len(const Box* b, int d) { return b->hi[d]-b->lo[d]+1; }
static Box b_grow(Box b, int g) { for (int d = 0; d < 3; d++) { b.lo[d]-=g; b.hi[d]+=g; } return b; }
static Box b_growd(Box b, int d, int g) { b.lo[d]-=g; b.hi[d]+=g; return b; }
static Box b_isect(Box a, const Box* b) {
  for (int d = 0; d < 3; d++) {
    if (b->lo[d] > a.lo[d]) a.lo[d] = b->lo[d];
    if (b->hi[d] < a.hi[d]) a.hi[d] = b->hi[d];
  }
  return a;
}
static int b_intersects(const Box* a, const Box* b) { Box c = b_isect(*a, b); return !b_empty(&c); }
/* IntVect::coarsen: floor division (IntVect.H:1288-1299) */
static int fdiv(int a, int r) { return a >= 0 ? a / r : -((-a + r - 1) / r); }
static Box b_coarsen(Box b, int r) {
  for (int d = 0; d < 3; d++) { b.lo[d] = fdiv(b.lo[d], r); b.hi[d] = fdiv(b.hi[d], r); }
  return b;
}
static Box b_refine(Box b, int r) {
  for (int d = 0; d < 3; d++) { b.lo[d] *= r; b.hi[d] = (b.hi[d]+1)*r - 1; }
  return b;
}
static int b_has(const Box* b, int i, int j, int k) {
  return i >= b->lo[0] && i <= b->hi[0] && j >= b->lo[1] && j <= b->hi[1] &&
         k >= b->lo[2] && k <= b->hi[2];
}
static int b_contains(const Box* a, const Box* b) {
  for (int d = 0; d < 3; d++) if (b->lo[d] < a->lo[d] || b->hi[d] > a->hi[d]) return 0;
  return 1;
}
static int b_eq(const Box* a, const Box* b) { return memcmp(a, b, sizeof(Box)) == 0; }
/* adjCellLo / adjCellHi (Box.H) : side 0 = lo, 1 = hi */
static Box b_adj(Box b, int d, int side, int n) {
  if (side == 0) { b.hi[d] = b.lo[d]-1; b.lo[d] = b.lo[d]-n; }
  else           { b.lo[d] = b.hi[d]+1; b.hi[d] = b.hi[d]+n; }
  return b;
}
static Box b_shift(Box b, int d, int s) { b.lo[d] += s; b.hi[d] += s; return b; }
/* IntVect::lexLT (IntVect.H:237) : dimension 0 most significant */
static int iv_lexlt(const int* a, const int* b) {
  for (int d = 0; d < 3; d++) { if (a[d] < b[d]) return 1; if (a[d] > b[d]) return 0; }
  return 0;
}
/* Box::operator< (Box.H:2223) */
static int b_lt(const Box* a, const Box* b) {
  if (memcmp(a->lo, b->lo, sizeof a->lo) == 0) return iv_lexlt(a->hi, b->hi);
  return iv_lexlt(a->lo, b->lo);
}
static Box mkbox(int l0, int l1, int l2, int h0, int h1, int h2) {
  Box b = {{l0,l1,l2},{h0,h1,h2}}; return b;
}

/* ======================================================================= */
/* Dense IntVectSet over a bounding box (DenseIntVectSet.cpp)              */
/* ======================================================================= */
typedef struct { Box mb; unsigned char* m; } IVS;

static void ivs_init_empty(IVS* s) { s->mb = mkbox(0,0,0,-1,-1,-1); s->m = NULL; }
static void ivs_free(IVS* s) { free(s->m); ivs_init_empty(s); }
static void ivs_define(IVS* s, Box b) {
  free(s->m); s->mb = b; s->m = NULL;
  if (!b_empty(&b)) { long n = b_npts(&b); s->m = malloc(n); memset(s->m, 1, n); }
}
static long ivs_off(const IVS* s, int i, int j, int k) {
  const Box* b = &s->mb;
  return ((long)(k-b->lo[2])*b_len(b,1) + (j-b->lo[1]))*b_len(b,0) + (i-b->lo[0]);
}
static int ivs_has(const IVS* s, int i, int j, int k) {
  if (!s->m || !b_has(&s->mb, i, j, k)) return 0;
  return s->m[ivs_off(s,i,j,k)];
}
static void ivs_sub_box(IVS* s, const Box* b) {
  if (!s->m) return;
  Box c = b_isect(s->mb, b);
  if (b_empty(&c)) return;
  for (int k = c.lo[2]; k <= c.hi[2]; k++)
    for (int j = c.lo[1]; j <= c.hi[1]; j++)
      for (int i = c.lo[0]; i <= c.hi[0]; i++) s->m[ivs_off(s,i,j,k)] = 0;
}
static void ivs_and_box(IVS* s, const Box* b) {
  if (!s->m) return;
  for (int k = s->mb.lo[2]; k <= s->mb.hi[2]; k++)
    for (int j = s->mb.lo[1]; j <= s->mb.hi[1]; j++)
      for (int i = s->mb.lo[0]; i <= s->mb.hi[0]; i++)
        if (!b_has(b, i, j, k)) s->m[ivs_off(s,i,j,k)] = 0;
}
static long ivs_numpts(const IVS* s) {
  if (!s->m) return 0;
  long n = b_npts(&s->mb), c = 0;
  for (long q = 0; q < n; q++) c += s->m[q];
  return c;
}
static Box ivs_minbox(const IVS* s) {
  Box r = mkbox(1<<30,1<<30,1<<30,-(1<<30),-(1<<30),-(1<<30));
  int any = 0;
  if (s->m)
    for (int k = s->mb.lo[2]; k <= s->mb.hi[2]; k++)
      for (int j = s->mb.lo[1]; j <= s->mb.hi[1]; j++)
        for (int i = s->mb.lo[0]; i <= s->mb.hi[0]; i++)
          if (s->m[ivs_off(s,i,j,k)]) {
            int iv[3] = {i,j,k}; any = 1;
            for (int d = 0; d < 3; d++) {
              if (iv[d] < r.lo[d]) r.lo[d] = iv[d];
              if (iv[d] > r.hi[d]) r.hi[d] = iv[d];
            }
          }
  if (!any) return mkbox(0,0,0,-1,-1,-1);
  return r;
}
static void ivs_copy(IVS* dst, const IVS* src) {
  ivs_free(dst);
  dst->mb = src->mb;
  if (src->m) { long n = b_npts(&src->mb); dst->m = malloc(n); memcpy(dst->m, src->m, n); }
}
/* DenseIntVectSet::coarsen */
static void ivs_coarsen(IVS* dst, const IVS* src, int r) {
  ivs_free(dst);
  if (!src->m) return;
  Box nb = b_coarsen(src->mb, r);
  dst->mb = nb; long n = b_npts(&nb); dst->m = calloc(n, 1);
  for (int k = src->mb.lo[2]; k <= src->mb.hi[2]; k++)
    for (int j = src->mb.lo[1]; j <= src->mb.hi[1]; j++)
      for (int i = src->mb.lo[0]; i <= src->mb.hi[0]; i++)
        if (src->m[ivs_off(src,i,j,k)])
          dst->m[ivs_off(dst, fdiv(i,r), fdiv(j,r), fdiv(k,r))] = 1;
}
/* DenseIntVectSet::grow(idir, -1): keep cells whose +-e neighbours are in the set */
static void ivs_erode(IVS* s, int d) {
  if (!s->m) return;
  long n = b_npts(&s->mb);
  unsigned char* o = malloc(n); memcpy(o, s->m, n);
  IVS t = { s->mb, o };
  for (int k = s->mb.lo[2]; k <= s->mb.hi[2]; k++)
    for (int j = s->mb.lo[1]; j <= s->mb.hi[1]; j++)
      for (int i = s->mb.lo[0]; i <= s->mb.hi[0]; i++) {
        int p[3] = {i,j,k}, q[3] = {i,j,k};
        p[d] -= 1; q[d] += 1;
        if (!(ivs_has(&t,p[0],p[1],p[2]) && ivs_has(&t,q[0],q[1],q[2])))
          s->m[ivs_off(s,i,j,k)] = 0;
      }
  free(o);
}
static int ivs_has_box(const IVS* s, const Box* b) {
  for (int k = b->lo[2]; k <= b->hi[2]; k++)
    for (int j = b->lo[1]; j <= b->hi[1]; j++)
      for (int i = b->lo[0]; i <= b->hi[0]; i++)
        if (!ivs_has(s,i,j,k)) return 0;
  return 1;
}
static void ivs_sub_ivs(IVS* s, const IVS* o) {
  if (!s->m || !o->m) return;
  for (int k = s->mb.lo[2]; k <= s->mb.hi[2]; k++)
    for (int j = s->mb.lo[1]; j <= s->mb.hi[1]; j++)
      for (int i = s->mb.lo[0]; i <= s->mb.hi[0]; i++)
        if (ivs_has(o,i,j,k)) s->m[ivs_off(s,i,j,k)] = 0;
}

/* ======================================================================= */
/* FArrayBox (Fortran order, one component; BaseFabImplem.H:291-304)       */
/* ======================================================================= */
typedef struct { Box b; long nx, ny; double* d; } Fab;

static void fab_alloc(Fab* f, Box b) {
  f->b = b; f->nx = b_len(&b,0); f->ny = b_len(&b,1);
  f->d = calloc((size_t)b_npts(&b), sizeof(double));
}
static void fab_free(Fab* f) { free(f->d); f->d = NULL; }
static inline double* P(const Fab* f, int i, int j, int k) {
  return f->d + ((long)(k-f->b.lo[2])*f->ny + (j-f->b.lo[1]))*f->nx + (i-f->b.lo[0]);
}
static void fab_setval(Fab* f, Box r, double v) {
  r = b_isect(r, &f->b);
  for (int k = r.lo[2]; k <= r.hi[2]; k++)
    for (int j = r.lo[1]; j <= r.hi[1]; j++)
      for (int i = r.lo[0]; i <= r.hi[0]; i++) *P(f,i,j,k) = v;
}
static void fab_copy_region(Fab* dst, const Fab* src, Box r) {
  for (int k = r.lo[2]; k <= r.hi[2]; k++)
    for (int j = r.lo[1]; j <= r.hi[1]; j++)
      for (int i = r.lo[0]; i <= r.hi[0]; i++) *P(dst,i,j,k) = *P(src,i,j,k);
}
/* the same from a periodic image: cell c of dst takes src's cell c - sh */
static void fab_copy_region_sh(Fab* dst, const Fab* src, Box r, const int sh[3]) {
  for (int k = r.lo[2]; k <= r.hi[2]; k++)
    for (int j = r.lo[1]; j <= r.hi[1]; j++)
      for (int i = r.lo[0]; i <= r.hi[0]; i++) *P(dst,i,j,k) = *P(src,i-sh[0],j-sh[1],k-sh[2]);
}

/* DisjointBoxLayout: boxes in Box::operator< order (BoxLayout::close sorts) */
typedef struct { int n; Box* b; } Layout;

static Layout lay_coarsen(const Layout* L, int r) {
  Layout c; c.n = L->n; c.b = malloc(sizeof(Box)*(L->n ? L->n : 1));
  for (int q = 0; q < L->n; q++) c.b[q] = b_coarsen(L->b[q], r);
  return c;
}
static int lay_coarsenable(const Layout* L, int r) {
  for (int q = 0; q < L->n; q++) {
    Box c = b_refine(b_coarsen(L->b[q], r), r);
    if (!b_eq(&c, &L->b[q])) return 0;
  }
  return 1;
}
static long lay_cells(const Layout* L) { long n = 0; for (int q = 0; q < L->n; q++) n += b_npts(&L->b[q]); return n; }

/* LevelData<FArrayBox> with ghost width g and nc components.  The FAB of box q,
   component c is f[c*L->n + q], so component c alone is the one-component LevelData
   COMP(x, c) over the same storage: every ChF kernel of the path loops
   "do n = 0, ncomp-1" around cell updates that never mix components, so running
   the one-component routine per component gives the same values. */
typedef struct { const Layout* L; int g; Fab* f; int nc, comp; } LD;
static LD ld_make_nc(const Layout* L, int g, int nc) {
  LD x; x.L = L; x.g = g; x.nc = nc; x.comp = 0;
  x.f = malloc(sizeof(Fab)*((long)L->n*nc > 0 ? (size_t)L->n*nc : 1));
  for (int c = 0; c < nc; c++)
    for (int q = 0; q < L->n; q++) fab_alloc(&x.f[(size_t)c*L->n + q], b_grow(L->b[q], g));
  return x;
}
static LD ld_make(const Layout* L, int g) { return ld_make_nc(L, g, 1); }
#define COMP(x, c) (&(LD){ (x)->L, (x)->g, (x)->f + (size_t)(c)*(x)->L->n, 1, (c) })
static void ld_free(LD* x) {
  if (!x->f) return;
  for (long q = 0; q < (long)x->L->n*x->nc; q++) fab_free(&x->f[q]);
  free(x->f); x->f = NULL;
}
/* valid-packed host order: box by box, in each box component by component (the
   order of an FArrayBox's data), cells x fastest */
static void ld_from_packed(LD* x, const double* v) {
  long o = 0;
  for (int q = 0; q < x->L->n; q++)
    for (int c = 0; c < x->nc; c++) {
      const Box* b = &x->L->b[q];
      Fab* f = &x->f[(size_t)c*x->L->n + q];
      for (int k = b->lo[2]; k <= b->hi[2]; k++)
        for (int j = b->lo[1]; j <= b->hi[1]; j++)
          for (int i = b->lo[0]; i <= b->hi[0]; i++) *P(f,i,j,k) = v[o++];
    }
}
static void ld_to_packed(const LD* x, double* v) {
  long o = 0;
  for (int q = 0; q < x->L->n; q++)
    for (int c = 0; c < x->nc; c++) {
      const Box* b = &x->L->b[q];
      const Fab* f = &x->f[(size_t)c*x->L->n + q];
      for (int k = b->lo[2]; k <= b->hi[2]; k++)
        for (int j = b->lo[1]; j <= b->hi[1]; j++)
          for (int i = b->lo[0]; i <= b->hi[0]; i++) v[o++] = *P(f,i,j,k);
    }
}
static void ld_from_ghosted(LD* x, const double* v) {
  long o = 0;
  for (int q = 0; q < x->L->n; q++) {
    long n = b_npts(&x->f[q].b);
    memcpy(x->f[q].d, v + o, sizeof(double)*n); o += n;
  }
}
static void ld_to_ghosted(const LD* x, double* v) {
  long o = 0;
  for (int q = 0; q < x->L->n; q++) {
    long n = b_npts(&x->f[q].b);
    memcpy(v + o, x->f[q].d, sizeof(double)*n); o += n;
  }
}

/* ======================================================================= */
/* ChF kernels, lib/src/AMRElliptic/AMRPoissonOpF.ChF                       */
/* ======================================================================= */
/* GSRBLAPLACIAN (AMRPoissonOpF.ChF:122-198) */
static void k_gsrb_lap(Fab* phi, const Fab* rhs, const Box* r, double dx, int redBlack) {
  double dxinv = 1.0/(dx*dx);
  double sum_b = 0.0;
  for (int d = 0; d < 3; d++) sum_b = sum_b + 2.0*dxinv;
  double lambda = -1.0/sum_b;
  for (int k = r->lo[2]; k <= r->hi[2]; k++)
    for (int j = r->lo[1]; j <= r->hi[1]; j++) {
      int imin = r->lo[0];
      int indtot = imin + j + k;
      imin = imin + abs((indtot + redBlack) % 2);
      for (int i = imin; i <= r->hi[0]; i += 2) {
        double lphi = (*P(phi,i+1,j,k) + *P(phi,i-1,j,k) + *P(phi,i,j+1,k)
                     + *P(phi,i,j-1,k) + *P(phi,i,j,k+1) + *P(phi,i,j,k-1)) * dxinv;
        *P(phi,i,j,k) = lambda*(*P(rhs,i,j,k) - lphi);
      }
    }
}
/* GSRBHELMHOLTZ (AMRPoissonOpF.ChF:53-120) */
static void k_gsrb_helm(Fab* phi, const Fab* rhs, const Box* r, double dx,
                        double alpha, double beta, int redBlack) {
  double dxinv = 1.0/(dx*dx);
  double sum_b = 0.0;
  for (int d = 0; d < 3; d++) sum_b = sum_b + 2.0*dxinv;
  double lambda = -1.0/(alpha - beta*sum_b);
  for (int k = r->lo[2]; k <= r->hi[2]; k++)
    for (int j = r->lo[1]; j <= r->hi[1]; j++) {
      int imin = r->lo[0];
      int indtot = imin + j + k;
      imin = imin + abs((indtot + redBlack) % 2);
      for (int i = imin; i <= r->hi[0]; i += 2) {
        double p = *P(phi,i,j,k);
        double lphi = (*P(phi,i+1,j,k) + *P(phi,i-1,j,k) + *P(phi,i,j+1,k)
                     + *P(phi,i,j-1,k) + *P(phi,i,j,k+1) + *P(phi,i,j,k-1)
                     - 6.0*p) * dxinv;
        double helmop = alpha*p + beta*lphi;
        *P(phi,i,j,k) = p + lambda*(helmop - *P(rhs,i,j,k));
      }
    }
}
/* The first cell of a coloured box: lo raised by one in each direction whose
   parity differs from the colour (AMRPoissonOp.cpp:1306-1313, 1556-1563). */
static int colored_box(const Box* v, const int color[3], Box* out) {
  *out = *v;
  for (int d = 0; d < 3; d++)
    if (abs(out->lo[d]) % 2 != color[d]) out->lo[d]++;
  return out->lo[0] <= out->hi[0] && out->lo[1] <= out->hi[1] && out->lo[2] <= out->hi[2];
}
/* GSMCAMRPOP (AMRPoissonOpF.ChF:10-51): a point update with the Laplacian only,
   phi += (rhs - lap phi) / (alpha - beta sum_b), over every second cell of the
   coloured box in each direction */
static void k_gsmc(Fab* phi, const Fab* rhs, const Box* cb, double dx, double alpha, double beta) {
  double dxinv = 1.0/(dx*dx);
  double sum_b = 0.0;
  for (int d = 0; d < 3; d++) sum_b = sum_b + 2.0*dxinv;
  double lambda = 1.0/(alpha - beta*sum_b);
  for (int k = cb->lo[2]; k <= cb->hi[2]; k += 2)
    for (int j = cb->lo[1]; j <= cb->hi[1]; j += 2)
      for (int i = cb->lo[0]; i <= cb->hi[0]; i += 2) {
        double p = *P(phi,i,j,k);
        double lphi = ((*P(phi,i+1,j,k) + *P(phi,i-1,j,k) - 2.0*p)
                     + (*P(phi,i,j+1,k) + *P(phi,i,j-1,k) - 2.0*p))
                     + (*P(phi,i,j,k+1) + *P(phi,i,j,k-1) - 2.0*p);
        lphi = lphi*dxinv;
        *P(phi,i,j,k) = p + lambda*(*P(rhs,i,j,k) - lphi);
      }
}
/* GSRBLAZY (AMRPoissonOpF.ChF:200-232): phi -= lambda (rhs - L phi) with
   lambda = -1 / (alpha - beta sum_b), over every second cell of the coloured box */
static void k_gsrb_lazy(Fab* phi, const Fab* lphi, const Fab* rhs, const Box* cb, double alpha, double beta,
                        double dx) {
  double dxinv = 1.0/(dx*dx);
  double sum_b = 0.0;
  for (int d = 0; d < 3; d++) sum_b = sum_b + 2.0*dxinv;
  double lambda = -1.0/(alpha - beta*sum_b);
  for (int k = cb->lo[2]; k <= cb->hi[2]; k += 2)
    for (int j = cb->lo[1]; j <= cb->hi[1]; j += 2)
      for (int i = cb->lo[0]; i <= cb->hi[0]; i += 2)
        *P(phi,i,j,k) = *P(phi,i,j,k) - lambda*(*P(rhs,i,j,k) - *P(lphi,i,j,k));
}
/* OPERATORLAP (AMRPoissonOpF.ChF:287-324) */
static void k_oplap(Fab* lofphi, const Fab* phi, const Box* r, double dx, double alpha, double beta) {
  double dxinv = 1.0/(dx*dx);
  for (int k = r->lo[2]; k <= r->hi[2]; k++)
    for (int j = r->lo[1]; j <= r->hi[1]; j++)
      for (int i = r->lo[0]; i <= r->hi[0]; i++) {
        double p = *P(phi,i,j,k);
        double lap = (*P(phi,i+1,j,k) + *P(phi,i-1,j,k) + *P(phi,i,j+1,k)
                    + *P(phi,i,j-1,k) + *P(phi,i,j,k+1) + *P(phi,i,j,k-1)
                    - 6.0*p) * dxinv;
        *P(lofphi,i,j,k) = alpha*p + beta*lap;
      }
}
/* OPERATORLAPRES (AMRPoissonOpF.ChF:341-377): r = -alpha*phi - beta*lap + rhs */
static void k_oplapres(Fab* res, const Fab* phi, const Fab* rhs, const Box* r,
                       double dx, double alpha, double beta) {
  double dxinv = 1.0/(dx*dx);
  for (int k = r->lo[2]; k <= r->hi[2]; k++)
    for (int j = r->lo[1]; j <= r->hi[1]; j++)
      for (int i = r->lo[0]; i <= r->hi[0]; i++) {
        double p = *P(phi,i,j,k);
        double lap = (*P(phi,i+1,j,k) + *P(phi,i-1,j,k) + *P(phi,i,j+1,k)
                    + *P(phi,i,j-1,k) + *P(phi,i,j,k+1) + *P(phi,i,j,k-1)
                    - 6.0*p) * dxinv;
        *P(res,i,j,k) = -(alpha*p) - beta*lap + *P(rhs,i,j,k);
      }
}
/* RESTRICTRES (AMRPoissonOpF.ChF:394-438) called with CHF_FRA_SHIFT as in
   AMRPoissonOp::restrictResidual (AMRPoissonOp.cpp:683-700): the coarse index is
   civ + (i - iv)/2 with iv = region.smallEnd(), civ = coarsen(iv, 2). */
static void k_restrictres(Fab* res, const Fab* phi, const Fab* rhs, double alpha, double beta,
                          const Box* r, double dx) {
  double dxinv = 1.0/(dx*dx);
  double denom = 8.0;
  int iv[3] = {r->lo[0], r->lo[1], r->lo[2]};
  int civ[3] = {fdiv(iv[0],2), fdiv(iv[1],2), fdiv(iv[2],2)};
  for (int k = r->lo[2]; k <= r->hi[2]; k++)
    for (int j = r->lo[1]; j <= r->hi[1]; j++)
      for (int i = r->lo[0]; i <= r->hi[0]; i++) {
        int ii = civ[0] + (i-iv[0])/2, jj = civ[1] + (j-iv[1])/2, kk = civ[2] + (k-iv[2])/2;
        double p = *P(phi,i,j,k);
        double lofphi = alpha*p
          + beta*(*P(phi,i+1,j,k) + *P(phi,i-1,j,k) + *P(phi,i,j+1,k)
                + *P(phi,i,j-1,k) + *P(phi,i,j,k+1) + *P(phi,i,j,k-1)
                - p*2.0*3.0) * dxinv;
        double* c = P(res,ii,jj,kk);
        *c = *c + (*P(rhs,i,j,k) - lofphi)/denom;
      }
}
/* PROLONG (AMRPoissonOpF.ChF:453-478) with the shifted indexing of
   AMRPoissonOp::prolongIncrement / AMRProlongS. */
static void k_prolong(Fab* phi, const Fab* coarse, const Box* r, int m) {
  int iv[3] = {r->lo[0], r->lo[1], r->lo[2]};
  int civ[3] = {fdiv(iv[0],m), fdiv(iv[1],m), fdiv(iv[2],m)};
  for (int k = r->lo[2]; k <= r->hi[2]; k++)
    for (int j = r->lo[1]; j <= r->hi[1]; j++)
      for (int i = r->lo[0]; i <= r->hi[0]; i++) {
        int ii = civ[0] + (i-iv[0])/m, jj = civ[1] + (j-iv[1])/m, kk = civ[2] + (k-iv[2])/m;
        *P(phi,i,j,k) = *P(phi,i,j,k) + *P(coarse,ii,jj,kk);
      }
}
/* average, refRatio 4 branch (lib/src/AMRTools/AverageF.ChF:60-134): the 64 children
   summed left to right, x fastest, then times refScale */
static void k_average4(Fab* coarse, const Fab* fine, const Box* b) {
  double refScale = 1.0/64.0;
  for (int k = b->lo[2]; k <= b->hi[2]; k++)
    for (int j = b->lo[1]; j <= b->hi[1]; j++)
      for (int i = b->lo[0]; i <= b->hi[0]; i++) {
        double sum = 0.0;
        int first = 1;
        for (int kk = 0; kk < 4; kk++)
          for (int jj = 0; jj < 4; jj++)
            for (int ii = 0; ii < 4; ii++) {
              double v = *P(fine, 4*i+ii, 4*j+jj, 4*k+kk);
              sum = first ? v : sum + v;
              first = 0;
            }
        *P(coarse,i,j,k) = refScale * sum;
      }
}
/* average, refRatio 2 branch (lib/src/AMRTools/AverageF.ChF:9-60) */
static void k_average2(Fab* coarse, const Fab* fine, const Box* b) {
  double refScale = 1.0/8.0;
  for (int k = b->lo[2]; k <= b->hi[2]; k++)
    for (int j = b->lo[1]; j <= b->hi[1]; j++)
      for (int i = b->lo[0]; i <= b->hi[0]; i++) {
        int a = 2*i, c = 2*j, e = 2*k;
        *P(coarse,i,j,k) = refScale *
          (*P(fine,a,c,e) + *P(fine,a+1,c,e) + *P(fine,a,c+1,e) + *P(fine,a+1,c+1,e)
         + *P(fine,a,c,e+1) + *P(fine,a+1,c,e+1) + *P(fine,a,c+1,e+1) + *P(fine,a+1,c+1,e+1));
      }
}
static void k_average(Fab* coarse, const Fab* fine, const Box* b, int ratio) {
  if (ratio == 4) k_average4(coarse, fine, b);
  else k_average2(coarse, fine, b);
}
/* INTERPHOMO (lib/src/AMRTools/InterpF.ChF:435-464) */
static void k_interphomo(Fab* phi, const Box* r, double x1, double dxCrse, int idir, int ihilo) {
  double c1 = 2.0*(dxCrse-x1)/(dxCrse+x1);
  double c2 = -((dxCrse-x1)/(dxCrse+3.0*x1));
  int s = -ihilo;
  for (int k = r->lo[2]; k <= r->hi[2]; k++)
    for (int j = r->lo[1]; j <= r->hi[1]; j++)
      for (int i = r->lo[0]; i <= r->hi[0]; i++) {
        int a[3] = {i,j,k}, b[3] = {i,j,k};
        a[idir] += s; b[idir] += 2*s;
        double pa = *P(phi,a[0],a[1],a[2]);
        double pb = *P(phi,b[0],b[1],b[2]);
        *P(phi,i,j,k) = c1*pa + c2*pb;
      }
}
/* INTERPHOMOLINEAR (InterpF.ChF:466-494) */
static void k_interphomolinear(Fab* phi, const Box* r, double x1, double dxCrse, int idir, int ihilo) {
  double factor = 1.0 - 2.0*x1/(x1+dxCrse);
  int s = -ihilo;
  for (int k = r->lo[2]; k <= r->hi[2]; k++)
    for (int j = r->lo[1]; j <= r->hi[1]; j++)
      for (int i = r->lo[0]; i <= r->hi[0]; i++) {
        int a[3] = {i,j,k}; a[idir] += s;
        *P(phi,i,j,k) = factor * *P(phi,a[0],a[1],a[2]);
      }
}

/* ======================================================================= */
/* ConstBCFunction (lib/src/AMRElliptic/BCFunc.cpp:58-158)                 */
/* ======================================================================= */
/* BCValueFunc (BCFunc.H:205-208) with a caller pointer */
typedef void (*BCValueFn)(double* pos, int* dir, int* side, double* value, void* user);
/* fn: when set, the inhomogeneous value of each ghost cell comes from fn in place of val */
typedef struct { int type[2][3]; double val[2][3]; BCValueFn fn; void* user; } BCSpec;

/* The BCValueHolder call of DiriBC / NeumBC (BCFunc.cpp:242-250, 309-313) for ghost
   cell (i,j,k) on face (idir, side): at getDomainFacePosition (BCFunc.cpp:187-202) of
   the valid cell next to it, dx*(iv+0.5) moved half a cell outward in idir.  The
   function fills one value per component (a_values[icomp]); comp picks one. */
#define ORACLE_MAX_NCOMP 16
static double face_value(const BCSpec* bc, int i, int j, int k, int idir, int side, double dx, int comp) {
  int isign = side ? 1 : -1;
  int iv[3] = {i, j, k};
  iv[idir] -= isign;
  double pos[3];
  for (int d = 0; d < 3; d++) pos[d] = dx*(iv[d] + 0.5);
  pos[idir] += 0.5*(double)isign*dx;
  int dir = idir, sd = side;
  double v[ORACLE_MAX_NCOMP] = {0.0};
  bc->fn(pos, &dir, &sd, v, bc->user);
  return v[comp];
}

static void apply_bc(Fab* st, const Box* valid, const Box* domain, double dx, int homog,
                     const BCSpec* bc, int comp) {
  for (int idir = 0; idir < 3; idir++)
    for (int side = 0; side < 2; side++) {
      int vend = side ? valid->hi[idir] : valid->lo[idir];
      int dend = side ? domain->hi[idir] : domain->lo[idir];
      if (vend != dend) continue;
      int type = bc->type[side][idir];
      if (type == 3) continue;   /* periodic direction: no BC (ghosts come from the exchange) */
      double value = bc->val[side][idir];
      int isign = side ? 1 : -1;
      Box to = b_isect(b_adj(*valid, idir, side, 1), &st->b);
      if (b_empty(&to)) continue;
      if (type == 0) {           /* Neumann */
        for (int k = to.lo[2]; k <= to.hi[2]; k++)
          for (int j = to.lo[1]; j <= to.hi[1]; j++)
            for (int i = to.lo[0]; i <= to.hi[0]; i++) {
              int c[3] = {i,j,k}; c[idir] -= isign;
              *P(st,i,j,k) = *P(st,c[0],c[1],c[2]);
            }
        if (!homog)
          for (int k = to.lo[2]; k <= to.hi[2]; k++)
            for (int j = to.lo[1]; j <= to.hi[1]; j++)
              for (int i = to.lo[0]; i <= to.hi[0]; i++) {
                double v = bc->fn ? face_value(bc, i, j, k, idir, side, dx, comp) : value;
                *P(st,i,j,k) += (double)isign*dx*v;
              }
      } else if (type == 2) {    /* DiriBC order 2, quadraticInterp (BCFunc.cpp:34-40, 291-354) */
        for (int k = to.lo[2]; k <= to.hi[2]; k++)
          for (int j = to.lo[1]; j <= to.hi[1]; j++)
            for (int i = to.lo[0]; i <= to.hi[0]; i++) {
              int c[3] = {i,j,k}; c[idir] -= isign;
              int f[3] = {c[0],c[1],c[2]}; f[idir] -= isign;
              double inhomogVal = 0.0;
              if (!homog) inhomogVal = bc->fn ? face_value(bc, i, j, k, idir, side, dx, comp) : value;
              double nearVal = *P(st,c[0],c[1],c[2]);
              double farVal = *P(st,f[0],f[1],f[2]);
              *P(st,i,j,k) = (8.0/3.0)*inhomogVal + (1.0/3.0)*(farVal) - 2*(nearVal);
            }
      } else {                   /* Dirichlet, linearInterp: 2*v - near */
        for (int k = to.lo[2]; k <= to.hi[2]; k++)
          for (int j = to.lo[1]; j <= to.hi[1]; j++)
            for (int i = to.lo[0]; i <= to.hi[0]; i++) {
              int c[3] = {i,j,k}; c[idir] -= isign;
              double inhomogVal = 0.0;
              if (!homog) inhomogVal = bc->fn ? face_value(bc, i, j, k, idir, side, dx, comp) : value;
              double nearVal = *P(st,c[0],c[1],c[2]);
              *P(st,i,j,k) = 2*inhomogVal - nearVal;
            }
      }
    }
}

/* ======================================================================= */
/* Copier motion plans (lib/src/BoxTools/Copier.cpp)                       */
/* ======================================================================= */
/* region: the cells written (toRegion); sh: the periodic shift of the source box,
   whose cells are region - sh (fromRegion = unshift(toRegion)); zero otherwise */
typedef struct { int from, to; Box region; int sh[3]; } Motion;
typedef struct { int n, cap; Motion* m; } Plan;

static void plan_push_sh(Plan* p, int from, int to, Box r, const int sh[3]) {
  if (p->n == p->cap) { p->cap = p->cap ? 2*p->cap : 16; p->m = realloc(p->m, sizeof(Motion)*p->cap); }
  p->m[p->n].from = from; p->m[p->n].to = to; p->m[p->n].region = r;
  for (int d = 0; d < 3; d++) p->m[p->n].sh[d] = sh[d];
  p->n++;
}
static void plan_push(Plan* p, int from, int to, Box r) {
  const int zero[3] = {0, 0, 0};
  plan_push_sh(p, from, to, r, zero);
}
static void plan_free(Plan* p) { free(p->m); p->m = NULL; p->n = p->cap = 0; }
/* MotionItemSorter (Copier.cpp:1547-1563): by toRegion, smallEnd then bigEnd */
static int motion_cmp(const void* a, const void* b) {
  const Motion* x = a; const Motion* y = b;
  if (b_lt(&x->region, &y->region)) return -1;
  if (b_lt(&y->region, &x->region)) return 1;
  return 0;
}
/* Copier::exchangeDefine + trimEdges (Copier.cpp:1464-1541, 142-177).  With a
   periodic domain the NeighborIterator also yields every box's periodic images
   (the box shifted by -L, 0 or +L in each periodic direction of domain length L,
   a box's own unshifted self excluded); an image's cells are copied from the box
   itself (nit.unshift).  periodic[d]: direction d is periodic. */
static Plan plan_exchange_p(const Layout* L, const Box* domain, const int periodic[3]) {
  Plan p = {0,0,NULL};
  int len[3], ns[3];
  for (int d = 0; d < 3; d++) {
    len[d] = domain->hi[d] - domain->lo[d] + 1;
    ns[d] = periodic[d] ? 3 : 1;
  }
  for (int t = 0; t < L->n; t++) {
    Box bg = b_grow(L->b[t], 1);
    for (int f = 0; f < L->n; f++)
      for (int a = 0; a < ns[0]; a++)
        for (int b = 0; b < ns[1]; b++)
          for (int c = 0; c < ns[2]; c++) {
            const int sh[3] = {periodic[0] ? (a - 1) * len[0] : 0, periodic[1] ? (b - 1) * len[1] : 0,
                               periodic[2] ? (c - 1) * len[2] : 0};
            if (f == t && sh[0] == 0 && sh[1] == 0 && sh[2] == 0) continue;
            Box img = L->b[f];
            for (int d = 0; d < 3; d++) { img.lo[d] += sh[d]; img.hi[d] += sh[d]; }
            Box r = b_isect(img, &bg);
            if (b_empty(&r)) continue;
            int keep = 0;
            for (int d = 0; d < 3 && !keep; d++) {
              Box g = b_growd(L->b[t], d, 1);
              if (b_intersects(&g, &r)) keep = 1;
            }
            if (keep) plan_push_sh(&p, f, t, r, sh);
          }
  }
  if (p.n > 1) qsort(p.m, p.n, sizeof(Motion), motion_cmp);
  return p;
}

/* Copier::define(src layout, dest layout, ghost) (Copier.cpp:903-1140): for each
   dest box grown by ghost, every intersecting src box */
static Plan plan_copy(const Layout* src, const Layout* dst, int dst_ghost) {
  Plan p = {0,0,NULL};
  for (int t = 0; t < dst->n; t++) {
    Box g = b_grow(dst->b[t], dst_ghost);
    for (int f = 0; f < src->n; f++) {
      Box r = b_isect(src->b[f], &g);
      if (!b_empty(&r)) plan_push(&p, f, t, r);
    }
  }
  if (p.n > 1) qsort(p.m, p.n, sizeof(Motion), motion_cmp);
  return p;
}
static void plan_reverse(Plan* p) {
  for (int q = 0; q < p->n; q++) { int t = p->m[q].from; p->m[q].from = p->m[q].to; p->m[q].to = t; }
  if (p->n > 1) qsort(p->m, p->n, sizeof(Motion), motion_cmp);
}
static void exec_exchange(LD* x, const Plan* p) {
  if (x->nc > 1) { for (int c = 0; c < x->nc; c++) exec_exchange(COMP(x, c), p); return; }
  for (int q = 0; q < p->n; q++)
    fab_copy_region_sh(&x->f[p->m[q].to], &x->f[p->m[q].from], p->m[q].region, p->m[q].sh);
}
/* LevelData::copyTo(..., copier) with plain copy */
static void exec_copy(LD* dst, const LD* src, const Plan* p) {
  if (dst->nc > 1) { for (int c = 0; c < dst->nc; c++) exec_copy(COMP(dst, c), COMP(src, c), p); return; }
  for (int q = 0; q < p->n; q++)
    fab_copy_region(&dst->f[p->m[q].to], &src->f[p->m[q].from], p->m[q].region);
}

/* ======================================================================= */
/* CFRegion / CFIVS (lib/src/AMRTools/CFRegion.cpp, CFIVS.cpp)             */
/* ======================================================================= */
typedef struct { int empty, packed; Box packedBox; IVS ivs; } CFIVS;

static void cfivs_define(CFIVS* c, const Box* domain, const Box* box, const Layout* L,
                         int dir, int side) {
  ivs_init_empty(&c->ivs);
  c->empty = 1; c->packed = 0; c->packedBox = mkbox(0,0,0,-1,-1,-1);
  Box edge = b_isect(b_adj(*box, dir, side, 1), domain);
  if (b_empty(&edge)) return;
  ivs_define(&c->ivs, edge);
  for (int q = 0; q < L->n; q++) ivs_sub_box(&c->ivs, &L->b[q]);
  /* packIVS */
  if (ivs_numpts(&c->ivs) == 0) { c->empty = 1; c->packed = 0; return; }
  c->empty = 0;
  c->packedBox = ivs_minbox(&c->ivs);
  c->packed = (b_npts(&c->packedBox) == ivs_numpts(&c->ivs));
}

/* ======================================================================= */
/* QuadCFStencil / QuadCFInterp (lib/src/AMRTools/QuadCFStencil.cpp,       */
/* QuadCFInterp.cpp, QuadCFInterpF.ChF, CFStencil.cpp, DerivStencil.cpp)   */
/* ======================================================================= */
#define MAXS 16
typedef struct { int n; int iv[MAXS][3]; double w[MAXS]; } DSten;

static void dsten_acc(DSten* s, int i, int j, int k, double w) {  /* DerivStencil::accumulate */
  for (int q = 0; q < s->n; q++)
    if (s->iv[q][0] == i && s->iv[q][1] == j && s->iv[q][2] == k) { s->w[q] += w; return; }
  s->iv[s->n][0] = i; s->iv[s->n][1] = j; s->iv[s->n][2] = k; s->w[s->n] = w; s->n++;
}
/* addFabToSten: BoxIterator order over the fab box, weight = fab value */
typedef struct { Box b; double v[27]; } SmallFab;
static void sfab_init(SmallFab* f, Box b, double v0) { f->b = b; for (int q = 0; q < 27; q++) f->v[q] = v0; }
static double* sfab(SmallFab* f, int i, int j, int k) {
  return &f->v[((k-f->b.lo[2])*b_len(&f->b,1) + (j-f->b.lo[1]))*b_len(&f->b,0) + (i-f->b.lo[0])];
}
static void add_fab_to_sten(SmallFab* f, DSten* s) {
  for (int k = f->b.lo[2]; k <= f->b.hi[2]; k++)
    for (int j = f->b.lo[1]; j <= f->b.hi[1]; j++)
      for (int i = f->b.lo[0]; i <= f->b.hi[0]; i++) dsten_acc(s, i, j, k, *sfab(f,i,j,k));
}

typedef struct {
  int defined_nonempty;        /* coarse IVS non-empty */
  int packed; Box packedBox;   /* m_baseCFS */
  IVS fine, coar;              /* m_baseCFS fine / coarse IVS */
  IVS standard, quadd;
  Box osbox;
  unsigned char* dropOrd;      /* over osbox */
  unsigned char* nmix;         /* over osbox: corner stencils averaged into the mixed derivative */
  DSten* mixed;                /* over osbox */
  DSten* first[3];             /* per derivative direction, over osbox */
  DSten* second[3];
  int dir;
} QCFS;

static long os_off(const QCFS* s, int i, int j, int k) {
  const Box* b = &s->osbox;
  return ((long)(k-b->lo[2])*b_len(b,1) + (j-b->lo[1]))*b_len(b,0) + (i-b->lo[0]);
}

/* QuadCFStencil::buildStencils (QuadCFStencil.cpp:662-918) */
static void qcfs_build_stencils(QCFS* s, const IVS* allGood) {
  int vt[2], n = 0;
  for (int d = 0; d < 3; d++) if (d != s->dir) vt[n++] = d;
  s->osbox = ivs_minbox(&s->quadd);
  long no = b_npts(&s->osbox);
  s->dropOrd = calloc(no, 1);
  s->nmix = calloc(no, 1);
  s->mixed = calloc(no, sizeof(DSten));
  for (int d = 0; d < 3; d++) { s->first[d] = calloc(no, sizeof(DSten)); s->second[d] = calloc(no, sizeof(DSten)); }
  int t1 = vt[0], t2 = vt[1];
  /* four 2x2 corner stencils: value -1 at lo corner and far corner, +1 at lo+e_t1, lo+e_t2 */
  for (int k = s->osbox.lo[2]; k <= s->osbox.hi[2]; k++)
    for (int j = s->osbox.lo[1]; j <= s->osbox.hi[1]; j++)
      for (int i = s->osbox.lo[0]; i <= s->osbox.hi[0]; i++) {
        if (!ivs_has(&s->quadd, i, j, k)) continue;
        int iv[3] = {i,j,k};
        /* lo-corner offsets of fabur, fabul, fablr, fabll (see the shifts at
           QuadCFStencil.cpp:700-722: fabur ends up shifted by -e_t1, fabul unshifted) */
        int offs[4][2] = {{-1,0},{0,0},{0,-1},{-1,-1}};
        DSten* ds = &s->mixed[os_off(s,i,j,k)];
        ds->n = 0;
        int imixed = 0, hasamixed = 0;
        for (int c = 0; c < 4; c++) {
          Box fb; for (int d = 0; d < 3; d++) { fb.lo[d] = iv[d]; fb.hi[d] = iv[d]; }
          fb.lo[t1] += offs[c][0]; fb.hi[t1] = fb.lo[t1] + 1;
          fb.lo[t2] += offs[c][1]; fb.hi[t2] = fb.lo[t2] + 1;
          if (!ivs_has_box(allGood, &fb)) continue;
          SmallFab f; sfab_init(&f, fb, -1.0);
          int a[3] = {fb.lo[0], fb.lo[1], fb.lo[2]}, b[3] = {fb.lo[0], fb.lo[1], fb.lo[2]};
          a[t1] += 1; b[t2] += 1;
          *sfab(&f, a[0], a[1], a[2]) = 1.0;
          *sfab(&f, b[0], b[1], b[2]) = 1.0;
          add_fab_to_sten(&f, ds);
          hasamixed = 1; imixed++;
        }
        if (imixed > 0) { double rn = imixed; for (int q = 0; q < ds->n; q++) ds->w[q] /= rn; }
        s->dropOrd[os_off(s,i,j,k)] = !hasamixed;
        s->nmix[os_off(s,i,j,k)] = (unsigned char)imixed;
      }
  for (int ivec = 0; ivec < 2; ivec++) {
    int idir = vt[ivec];
    for (int k = s->osbox.lo[2]; k <= s->osbox.hi[2]; k++)
      for (int j = s->osbox.lo[1]; j <= s->osbox.hi[1]; j++)
        for (int i = s->osbox.lo[0]; i <= s->osbox.hi[0]; i++) {
          if (!ivs_has(&s->quadd, i, j, k)) continue;
          long o = os_off(s,i,j,k);
          if (s->dropOrd[o]) continue;
          int iv[3] = {i,j,k};
          DSten* sec = &s->second[idir][o];
          DSten* fir = &s->first[idir][o];
          Box bsten; for (int d = 0; d < 3; d++) { bsten.lo[d] = iv[d]; bsten.hi[d] = iv[d]; }
          bsten = b_growd(bsten, idir, 1);
          int m[3] = {i,j,k}, p[3] = {i,j,k}, p2[3] = {i,j,k}, m2[3] = {i,j,k};
          m[idir] -= 1; p[idir] += 1; p2[idir] += 2; m2[idir] -= 2;
          if (ivs_has_box(allGood, &bsten)) {
            SmallFab f; sfab_init(&f, bsten, 0.0);
            *sfab(&f,m[0],m[1],m[2]) = 1.0; *sfab(&f,p[0],p[1],p[2]) = 1.0; *sfab(&f,i,j,k) = -2.0;
            sec->n = 0; add_fab_to_sten(&f, sec);
            *sfab(&f,m[0],m[1],m[2]) = -0.5; *sfab(&f,p[0],p[1],p[2]) = 0.5; *sfab(&f,i,j,k) = 0.0;
            fir->n = 0; add_fab_to_sten(&f, fir);
            continue;
          }
          Box bshift = b_shift(bsten, idir, 1);
          if (ivs_has_box(allGood, &bshift)) {
            SmallFab f; sfab_init(&f, bshift, 0.0);
            *sfab(&f,i,j,k) = 1.0; *sfab(&f,p[0],p[1],p[2]) = -2.0; *sfab(&f,p2[0],p2[1],p2[2]) = 1.0;
            sec->n = 0; add_fab_to_sten(&f, sec);
            *sfab(&f,i,j,k) = -3.0/2.0; *sfab(&f,p[0],p[1],p[2]) = 4.0/2.0; *sfab(&f,p2[0],p2[1],p2[2]) = -1.0/2.0;
            fir->n = 0; add_fab_to_sten(&f, fir);
            continue;
          }
          bshift = b_shift(bshift, idir, -2);
          if (ivs_has_box(allGood, &bshift)) {
            SmallFab f; sfab_init(&f, bshift, 0.0);
            *sfab(&f,i,j,k) = 1.0; *sfab(&f,m[0],m[1],m[2]) = -2.0; *sfab(&f,m2[0],m2[1],m2[2]) = 1.0;
            sec->n = 0; add_fab_to_sten(&f, sec);
            *sfab(&f,i,j,k) = 3.0/2.0; *sfab(&f,m[0],m[1],m[2]) = -4.0/2.0; *sfab(&f,m2[0],m2[1],m2[2]) = 1.0/2.0;
            fir->n = 0; add_fab_to_sten(&f, fir);
            continue;
          }
          s->dropOrd[o] = 1;
          if (ivs_has(allGood, p[0], p[1], p[2])) {
            Box bf; for (int d = 0; d < 3; d++) { bf.lo[d] = iv[d]; bf.hi[d] = p[d]; }
            SmallFab f; sfab_init(&f, bf, 0.0);
            *sfab(&f,i,j,k) = -1.0; *sfab(&f,p[0],p[1],p[2]) = 1.0;
            fir->n = 0; add_fab_to_sten(&f, fir);
          } else if (ivs_has(allGood, m[0], m[1], m[2])) {
            Box bf; for (int d = 0; d < 3; d++) { bf.lo[d] = m[d]; bf.hi[d] = iv[d]; }
            SmallFab f; sfab_init(&f, bf, 0.0);
            *sfab(&f,i,j,k) = 1.0; *sfab(&f,m[0],m[1],m[2]) = -1.0;
            fir->n = 0; add_fab_to_sten(&f, fir);
          } else {
            Box bf; for (int d = 0; d < 3; d++) { bf.lo[d] = iv[d]; bf.hi[d] = iv[d]; }
            SmallFab f; sfab_init(&f, bf, 0.0);
            fir->n = 0; add_fab_to_sten(&f, fir);
          }
        }
  }
}

/* CFStencil::define (CFStencil.cpp) + QuadCFStencil::define (QuadCFStencil.cpp:455-660) */
static void qcfs_define(QCFS* s, const Box* fineDomain, const Box* grid, const Layout* fineL,
                        const Layout* coarL, int ref, int dir, int side) {
  memset(s, 0, sizeof *s);
  ivs_init_empty(&s->fine); ivs_init_empty(&s->coar);
  ivs_init_empty(&s->standard); ivs_init_empty(&s->quadd);
  s->dir = dir;
  Box edge = b_isect(b_adj(*grid, dir, side, 1), fineDomain);
  if (b_empty(&edge)) return;
  ivs_define(&s->fine, edge);
  for (int q = 0; q < fineL->n; q++) ivs_sub_box(&s->fine, &fineL->b[q]);
  ivs_coarsen(&s->coar, &s->fine, ref);
  Box cdom = b_coarsen(*fineDomain, ref);
  Box dib = b_isect(ivs_minbox(&s->coar), &cdom);
  ivs_and_box(&s->coar, &dib);
  s->packedBox = ivs_minbox(&s->fine);
  s->packed = (ivs_numpts(&s->fine) > 0 && ivs_numpts(&s->fine) == b_npts(&s->packedBox));
  if (ivs_numpts(&s->coar) == 0) return;
  s->defined_nonempty = 1;

  int vt[2], n = 0;
  for (int d = 0; d < 3; d++) if (d != dir) vt[n++] = d;
  Box cgrid = b_coarsen(*grid, ref);
  Box g2 = b_adj(cgrid, dir, side, 1), g1 = g2;
  for (int t = 0; t < 2; t++) { g2 = b_growd(g2, vt[t], 2); g1 = b_growd(g1, vt[t], 1); }
  g2 = b_isect(g2, &cdom); g1 = b_isect(g1, &cdom);
  if (b_empty(&g2)) return;
  IVS allGood; ivs_init_empty(&allGood); ivs_define(&allGood, g2);
  for (int q = 0; q < fineL->n; q++) { Box cb = b_coarsen(fineL->b[q], ref); ivs_sub_box(&allGood, &cb); }
  /* ivsEdgeOfEarth = allGood minus coarse boxes; allGood -= edgeOfEarth */
  IVS eoe; ivs_init_empty(&eoe); ivs_copy(&eoe, &allGood);
  for (int q = 0; q < coarL->n; q++) ivs_sub_box(&eoe, &coarL->b[q]);
  ivs_sub_ivs(&allGood, &eoe);
  ivs_free(&eoe);
  ivs_copy(&s->standard, &allGood);
  ivs_and_box(&s->standard, &g1);
  for (int t = 0; t < 2; t++) ivs_erode(&s->standard, vt[t]);
  ivs_copy(&s->quadd, &s->coar);
  ivs_sub_ivs(&s->quadd, &s->standard);
  if (ivs_numpts(&s->quadd) > 0) qcfs_build_stencils(s, &allGood);
  ivs_free(&allGood);
}
static void qcfs_free(QCFS* s) {
  ivs_free(&s->fine); ivs_free(&s->coar); ivs_free(&s->standard); ivs_free(&s->quadd);
  free(s->dropOrd); free(s->nmix); free(s->mixed);
  for (int d = 0; d < 3; d++) { free(s->first[d]); free(s->second[d]); }
}

/* QuadCFStencil::computeFirstDerivative (QuadCFStencil.cpp:130-185) */
static double qcfs_first(const QCFS* s, const Fab* phic, int dd, int i, int j, int k, double dx) {
  if (ivs_has(&s->standard, i, j, k)) {
    int lo[3] = {i,j,k}, hi[3] = {i,j,k}; lo[dd] -= 1; hi[dd] += 1;
    double philo = *P(phic,lo[0],lo[1],lo[2]);
    double phihi = *P(phic,hi[0],hi[1],hi[2]);
    return (phihi - philo)/(2.0*dx);
  }
  const DSten* ds = &s->first[dd][os_off(s,i,j,k)];
  double firstd = 0.0; int keepzer = 0;
  for (int q = 0; q < ds->n; q++) {
    if (b_has(&phic->b, ds->iv[q][0], ds->iv[q][1], ds->iv[q][2]))
      firstd += ds->w[q] * *P(phic, ds->iv[q][0], ds->iv[q][1], ds->iv[q][2]);
    else keepzer = 1;
  }
  if (keepzer) firstd = 0.0;
  firstd /= dx;
  return firstd;
}
/* QuadCFStencil::computeSecondDerivative (QuadCFStencil.cpp:36-120) */
static double qcfs_second(const QCFS* s, const Fab* phic, int dd, int i, int j, int k, double dx) {
  if (ivs_has(&s->standard, i, j, k)) {
    int lo[3] = {i,j,k}, hi[3] = {i,j,k}; lo[dd] -= 1; hi[dd] += 1;
    double phicc = *P(phic,i,j,k);
    double philo = *P(phic,lo[0],lo[1],lo[2]);
    double phihi = *P(phic,hi[0],hi[1],hi[2]);
    return (phihi + philo - 2.0*phicc)/(dx*dx);
  }
  long o = os_off(s,i,j,k);
  if (s->dropOrd[o]) return 0.0;
  const DSten* ds = &s->second[dd][o];
  double secondd = 0.0; int keepzer = 0;
  for (int q = 0; q < ds->n; q++) {
    if (b_has(&phic->b, ds->iv[q][0], ds->iv[q][1], ds->iv[q][2]))
      secondd += ds->w[q] * *P(phic, ds->iv[q][0], ds->iv[q][1], ds->iv[q][2]);
    else keepzer = 1;
  }
  if (keepzer) secondd = 0.0;
  secondd /= (dx*dx);
  return secondd;
}
/* QuadCFStencil::computeMixedDerivative (QuadCFStencil.cpp:196-268) */
static double qcfs_mixed(const QCFS* s, const Fab* phic, int i, int j, int k, double dx) {
  int itran1, itran2;
  if (s->dir == 0) { itran2 = 1; itran1 = 2; }
  else if (s->dir == 1) { itran2 = 0; itran1 = 2; }
  else { itran2 = 0; itran1 = 1; }
  if (ivs_has(&s->standard, i, j, k)) {
    int ur[3] = {i,j,k}, ul[3] = {i,j,k}, lr[3] = {i,j,k}, ll[3] = {i,j,k};
    ur[itran2] += 1; ur[itran1] += 1;
    ul[itran2] -= 1; ul[itran1] += 1;
    lr[itran2] += 1; lr[itran1] -= 1;
    ll[itran2] -= 1; ll[itran1] -= 1;
    double phiur = *P(phic,ur[0],ur[1],ur[2]);
    double phiul = *P(phic,ul[0],ul[1],ul[2]);
    double philr = *P(phic,lr[0],lr[1],lr[2]);
    double phill = *P(phic,ll[0],ll[1],ll[2]);
    return (phiur+phill-philr-phiul)/(4.*dx*dx);
  }
  long o = os_off(s,i,j,k);
  if (s->dropOrd[o]) return 0.0;
  const DSten* ds = &s->mixed[o];
  double mixedd = 0.0; int keepzer = 0;
  for (int q = 0; q < ds->n; q++) {
    if (b_has(&phic->b, ds->iv[q][0], ds->iv[q][1], ds->iv[q][2]))
      mixedd += ds->w[q] * *P(phic, ds->iv[q][0], ds->iv[q][1], ds->iv[q][2]);
    else keepzer = 1;
  }
  if (keepzer) mixedd = 0.0;
  mixedd /= (dx*dx);
  return mixedd;
}

typedef struct {
  int level;               /* 0 when there is no coarser level or it is covered */
  double dxFine; int ref;
  Layout coarBoxes;        /* coarsen(fine box, ref) grown by 2, one per fine box */
  LD buf;                  /* coarse buffer on coarBoxes */
  Plan copier;             /* coarse grids -> coarBoxes */
  QCFS* st;                /* [fine box][dir][side] */
  const Layout* fineL;
} QCFI;

/* QuadCFInterp::define (QuadCFInterp.cpp:397-525) */
static void qcfi_define(QCFI* q, const Layout* fineL, const Layout* coarL, double dxFine, int ref,
                        const Box* fineDomain) {
  memset(q, 0, sizeof *q);
  q->dxFine = dxFine; q->ref = ref; q->fineL = fineL;
  long numPts = lay_cells(fineL)/(ref*ref*ref) - lay_cells(coarL);
  q->level = (numPts == 0) ? 0 : 1;
  if (!q->level) return;
  q->coarBoxes.n = fineL->n; q->coarBoxes.b = malloc(sizeof(Box)*fineL->n);
  for (int b = 0; b < fineL->n; b++) q->coarBoxes.b[b] = b_grow(b_coarsen(fineL->b[b], ref), 2);
  q->buf = ld_make(&q->coarBoxes, 0);
  q->copier = plan_copy(coarL, &q->coarBoxes, 0);
  q->st = calloc((size_t)fineL->n*6, sizeof(QCFS));
  for (int b = 0; b < fineL->n; b++)
    for (int d = 0; d < 3; d++)
      for (int s = 0; s < 2; s++)
        qcfs_define(&q->st[b*6+d*2+s], fineDomain, &fineL->b[b], fineL, coarL, ref, d, s);
}
static void qcfi_free(QCFI* q) {
  if (!q->level) return;
  for (int b = 0; b < q->fineL->n*6; b++) qcfs_free(&q->st[b]);
  free(q->st); ld_free(&q->buf); plan_free(&q->copier); free(q->coarBoxes.b);
}

/* getPhiStar (QuadCFInterp.cpp:570-764) + interpOnIVS (:775-839) for one face */
static void qcfi_face(const QCFI* q, Fab* phif, const Fab* phic, const QCFS* s, int side, int idir) {
  if (!s->defined_nonempty && ivs_numpts(&s->fine) == 0) return;
  if (ivs_numpts(&s->fine) == 0) return;
  double dxf = q->dxFine, dxc = q->ref*dxf;
  int you1, you2;
  if (idir == 0) { you1 = 1; you2 = 2; } else if (idir == 1) { you1 = 0; you2 = 2; } else { you1 = 0; you2 = 1; }
  int ihilo = side ? 1 : -1;
  Box cib = ivs_minbox(&s->coar);
  Box psb = b_shift(ivs_minbox(&s->fine), idir, ihilo);
  Fab phistar; fab_alloc(&phistar, psb);
  /* slopes on the coarse IVS */
  Fab slope[3], curva[3], mixed;
  int dd[2] = {you1, you2};
  for (int t = 0; t < 2; t++) { fab_alloc(&slope[dd[t]], cib); fab_alloc(&curva[dd[t]], cib); }
  fab_alloc(&mixed, cib);
  for (int k = cib.lo[2]; k <= cib.hi[2]; k++)
    for (int j = cib.lo[1]; j <= cib.hi[1]; j++)
      for (int i = cib.lo[0]; i <= cib.hi[0]; i++) {
        if (!ivs_has(&s->coar, i, j, k)) continue;
        *P(&slope[you1],i,j,k) = qcfs_first(s, phic, you1, i, j, k, dxc);
        *P(&curva[you1],i,j,k) = qcfs_second(s, phic, you1, i, j, k, dxc);
        *P(&slope[you2],i,j,k) = qcfs_first(s, phic, you2, i, j, k, dxc);
        *P(&curva[you2],i,j,k) = qcfs_second(s, phic, you2, i, j, k, dxc);
        *P(&mixed,i,j,k) = qcfs_mixed(s, phic, i, j, k, dxc);
      }
  if (s->packed) {
    /* FORT_PHISTAR (QuadCFInterpF.ChF:50-126) with CHF_*_SHIFT(iv / civ) */
    int iv[3] = {psb.lo[0], psb.lo[1], psb.lo[2]};
    int civ[3] = {fdiv(iv[0],q->ref), fdiv(iv[1],q->ref), fdiv(iv[2],q->ref)};
    const Box* r = &s->packedBox;
    int nRef = q->ref;
    double dxcf = nRef*dxf;
    for (int k = r->lo[2]; k <= r->hi[2]; k++)
      for (int j = r->lo[1]; j <= r->hi[1]; j++)
        for (int i = r->lo[0]; i <= r->hi[0]; i++) {
          int ir[3] = {i-iv[0], j-iv[1], k-iv[2]};
          int ic[3] = {ir[0]/nRef, ir[1]/nRef, ir[2]/nRef};   /* Fortran truncating division */
          int cg[3] = {ic[0]+civ[0], ic[1]+civ[1], ic[2]+civ[2]};
          double xf1 = (ir[you1]+0.5)*dxf;
          double xc1 = (ic[you1]+0.5)*dxcf;
          double xf2 = (ir[you2]+0.5)*dxf;
          double xc2 = (ic[you2]+0.5)*dxcf;
          double x1 = xf1-xc1, x2 = xf2-xc2;
          double aa = *P(phic, cg[0], cg[1], cg[2]);
          double update1 = x1* *P(&slope[you1],cg[0],cg[1],cg[2]) + 0.5*x1*x1* *P(&curva[you1],cg[0],cg[1],cg[2]);
          double update2 = x2* *P(&slope[you2],cg[0],cg[1],cg[2]) + 0.5*x2*x2* *P(&curva[you2],cg[0],cg[1],cg[2]);
          double update3 = x1*x2* *P(&mixed,cg[0],cg[1],cg[2]);
          int st[3] = {i,j,k}; st[idir] += ihilo;
          *P(&phistar, st[0], st[1], st[2]) = aa+update1+update2+update3;
        }
  } else {
    const Box mb = s->fine.mb;
    for (int k = mb.lo[2]; k <= mb.hi[2]; k++)
      for (int j = mb.lo[1]; j <= mb.hi[1]; j++)
        for (int i = mb.lo[0]; i <= mb.hi[0]; i++) {
          if (!ivs_has(&s->fine, i, j, k)) continue;
          int ivf[3] = {i,j,k};
          int ivc[3] = {fdiv(i,q->ref), fdiv(j,q->ref), fdiv(k,q->ref)};
          double pc = *P(phic, ivc[0], ivc[1], ivc[2]);
          int jf = ivf[you1], jc = ivc[you1];
          double xf = (jf+0.5)*dxf, xc = (jc+0.5)*dxc;
          double x1 = xf-xc;
          double update1 = x1* *P(&slope[you1],ivc[0],ivc[1],ivc[2]) + 0.5*x1*x1* *P(&curva[you1],ivc[0],ivc[1],ivc[2]);
          jf = ivf[you2]; jc = ivc[you2];
          xf = (jf+0.5)*dxf; xc = (jc+0.5)*dxc;
          double x2 = xf-xc;
          double update2 = x2* *P(&slope[you2],ivc[0],ivc[1],ivc[2]) + 0.5*x2*x2* *P(&curva[you2],ivc[0],ivc[1],ivc[2]);
          double update3 = x1*x2* *P(&mixed,ivc[0],ivc[1],ivc[2]);
          int st[3] = {i,j,k}; st[idir] += ihilo;
          *P(&phistar, st[0], st[1], st[2]) = pc+update1+update2+update3;
        }
  }
  /* interpOnIVS */
  int nref = q->ref;
  if (!s->packed) {
    const Box mb = s->fine.mb;
    for (int k = mb.lo[2]; k <= mb.hi[2]; k++)
      for (int j = mb.lo[1]; j <= mb.hi[1]; j++)
        for (int i = mb.lo[0]; i <= mb.hi[0]; i++) {
          if (!ivs_has(&s->fine, i, j, k)) continue;
          int a2[3] = {i,j,k}, a1[3] = {i,j,k}, sp[3] = {i,j,k};
          a2[idir] -= 2*ihilo; a1[idir] -= ihilo; sp[idir] += ihilo;
          double pa = *P(phif, a2[0], a2[1], a2[2]);
          double pb = *P(phif, a1[0], a1[1], a1[2]);
          double ps = *P(&phistar, sp[0], sp[1], sp[2]);
          double h = dxf;
          double a = (2./h/h)*(2.*ps + pa*(nref+1.0) -pb*(nref+3.0))/(nref*nref + 4*nref + 3.0);
          double b = (pb-pa)/h - a*h;
          double c = pa;
          double x = 2.*h;
          *P(phif, i, j, k) = a*x*x + b*x + c;
        }
  } else {
    /* FORT_QUADINTERP (QuadCFInterpF.ChF:9-47) */
    double h = dxf;
    double frac = 2.0/(h*h);
    double denom = nref*nref + 4*nref + 3;
    double mult = frac/denom;
    double invh = 1.0/h;
    double x = 2.0*h;
    double xsquared = 4.0*h*h;
    const Box* r = &s->packedBox;
    for (int k = r->lo[2]; k <= r->hi[2]; k++)
      for (int j = r->lo[1]; j <= r->hi[1]; j++)
        for (int i = r->lo[0]; i <= r->hi[0]; i++) {
          int a2[3] = {i,j,k}, a1[3] = {i,j,k}, sp[3] = {i,j,k};
          a2[idir] -= 2*ihilo; a1[idir] -= ihilo; sp[idir] += ihilo;
          double pa = *P(phif, a2[0], a2[1], a2[2]);
          double pb = *P(phif, a1[0], a1[1], a1[2]);
          double ps = *P(&phistar, sp[0], sp[1], sp[2]);
          double a = mult*(2.0*ps + (nref+1)*pa - (nref+3)*pb);
          double b = (pb-pa)*invh - a*h;
          *P(phif, i, j, k) = xsquared*a + b*x + pa;
        }
  }
  fab_free(&phistar);
  for (int t = 0; t < 2; t++) { fab_free(&slope[dd[t]]); fab_free(&curva[dd[t]]); }
  fab_free(&mixed);
}

/* QuadCFInterp::coarseFineInterp(LevelData&, const LevelData&) (QuadCFInterp.cpp:846-889) */
static void qcfi_interp(QCFI* q, LD* phif, const LD* phic) {
  if (phif->nc > 1) { for (int c = 0; c < phif->nc; c++) qcfi_interp(q, COMP(phif, c), COMP(phic, c)); return; }
  if (!q->level) return;
  exec_copy(&q->buf, phic, &q->copier);
  for (int idir = 0; idir < 3; idir++)
    for (int b = 0; b < phif->L->n; b++)
      for (int side = 0; side < 2; side++)
        qcfi_face(q, &phif->f[b], &q->buf.f[b], &q->st[b*6+idir*2+side], side, idir);
}

/* ======================================================================= */
/* LevelFluxRegister (lib/src/AMRTools/LevelFluxRegister.cpp)              */
/* ======================================================================= */
typedef struct { int n; Box* b; } BoxList;
typedef struct {
  const Layout* coarL; const Layout* fineL;
  Layout cfine;              /* coarsened fine layout */
  BoxList* loc;              /* [coarse box][dir][side] */
  LD coarFlux;               /* on coarse grids */
  LD fineFlux;               /* on coarsened fine grids, ghost 1 */
  Plan rev;                  /* fine registers (grown by 1) -> coarse, sorted by toRegion */
  int nRefine;
} FluxReg;

static void bl_push(BoxList* l, Box b) { l->b = realloc(l->b, sizeof(Box)*(l->n+1)); l->b[l->n++] = b; }

/* LevelFluxRegister::define (LevelFluxRegister.cpp:101-262), non-periodic */
static void fr_define(FluxReg* r, const Layout* fineL, const Layout* coarL, int nref) {
  memset(r, 0, sizeof *r);
  r->coarL = coarL; r->fineL = fineL; r->nRefine = nref;
  r->cfine = lay_coarsen(fineL, nref);
  r->coarFlux = ld_make(coarL, 0);
  r->fineFlux = ld_make(&r->cfine, 1);
  r->loc = calloc((size_t)coarL->n*6, sizeof(BoxList));
  for (int c = 0; c < coarL->n; c++) {
    const Box* cb = &coarL->b[c];
    for (int f = 0; f < r->cfine.n; f++) {
      const Box* fb = &r->cfine.b[f];
      if (fb->hi[0]+1 < cb->lo[0]) continue;
      if (fb->lo[0]-1 > cb->hi[0]) break;
      for (int i = 0; i < 3; i++) {
        Box lo = b_adj(*fb, i, 0, 1), hi = b_adj(*fb, i, 1, 1);
        if (b_intersects(cb, &lo)) bl_push(&r->loc[c*6+i*2+0], b_isect(lo, cb));
        if (b_intersects(cb, &hi)) bl_push(&r->loc[c*6+i*2+1], b_isect(hi, cb));
      }
    }
  }
  /* m_reverseCopier.ghostDefine(coarsenedFine, coarse, domain, Unit):
     Copier::define(coarse, coarsenedFine, ghost 1), reversed, sorted */
  r->rev = plan_copy(coarL, &r->cfine, 1);
  plan_reverse(&r->rev);
}
static void fr_free(FluxReg* r) {
  if (!r->loc) return;
  for (int q = 0; q < r->coarL->n*6; q++) free(r->loc[q].b);
  free(r->loc); ld_free(&r->coarFlux); ld_free(&r->fineFlux); plan_free(&r->rev); free(r->cfine.b);
  r->loc = NULL;
}

/* ======================================================================= */
/* AMRPoissonOp (lib/src/AMRElliptic/AMRPoissonOp.cpp)                     */
/* ======================================================================= */
typedef struct Op {
  Layout L; Box domain; double dx, dxCrse, alpha, beta; BCSpec bc;
  Plan ex; CFIVS* cf;
  int hasCoarser, hasFiner, refToCoarser, refToFiner;
  QCFI qcfi;          /* interpolation from the next coarser AMR level */
  FluxReg fr;         /* register with the next finer AMR level */
  /* VCAMRPoissonOp2 (variable coefficients): L phi = alpha a phi - beta div(b grad phi) */
  int vc;
  LD va, vlam;        /* aCoef and the relaxation coefficient lambda, valid cells */
  Fab* vb[3];         /* bCoef: per box, the faces normal to d (box grown by one face in d) */
  int relaxMode;      /* AMRPoissonOp::s_relaxMode: 1 levelGSRB, 4 levelJacobi */
} Op;

static void op_define(Op* op, const Layout* L, const Box* domain, double dx, double alpha,
                      double beta, const BCSpec* bc) {
  memset(op, 0, sizeof *op);
  op->L.n = L->n; op->L.b = malloc(sizeof(Box)*(L->n ? L->n : 1));
  memcpy(op->L.b, L->b, sizeof(Box)*L->n);
  op->domain = *domain; op->dx = dx; op->dxCrse = 2*dx; op->alpha = alpha; op->beta = beta;
  op->bc = *bc; op->refToCoarser = 2; op->refToFiner = 2; op->relaxMode = 1;
  {
    int per[3];
    for (int d = 0; d < 3; d++) per[d] = bc->type[0][d] == 3;
    op->ex = plan_exchange_p(&op->L, domain, per);
  }
  op->cf = calloc((size_t)L->n*6, sizeof(CFIVS));
  for (int b = 0; b < L->n; b++)
    for (int d = 0; d < 3; d++)
      for (int s = 0; s < 2; s++) cfivs_define(&op->cf[b*6+d*2+s], domain, &L->b[b], &op->L, d, s);
}
static void op_free(Op* op) {
  if (op->vc) {
    ld_free(&op->va); ld_free(&op->vlam);
    for (int d = 0; d < 3; d++) { for (int b = 0; b < op->L.n; b++) fab_free(&op->vb[d][b]); free(op->vb[d]); }
  }
  for (int b = 0; b < op->L.n*6; b++) ivs_free(&op->cf[b].ivs);
  free(op->cf); plan_free(&op->ex);
  if (op->hasCoarser) qcfi_free(&op->qcfi);
  if (op->hasFiner) fr_free(&op->fr);
  free(op->L.b);
}

static void op_exchange(const Op* op, LD* phi) { exec_exchange(phi, &op->ex); }
static void op_bc(const Op* op, LD* phi, int homog) {
  if (phi->nc > 1) { for (int c = 0; c < phi->nc; c++) op_bc(op, COMP(phi, c), homog); return; }
  for (int b = 0; b < op->L.n; b++) apply_bc(&phi->f[b], &op->L.b[b], &op->domain, op->dx, homog, &op->bc, phi->comp);
}
/* homogeneousCFInterp (AMRPoissonOp.cpp:1598-1794) */
static void op_homog_cf(const Op* op, LD* phi) {
  if (phi->nc > 1) { for (int c = 0; c < phi->nc; c++) op_homog_cf(op, COMP(phi, c)); return; }
  for (int b = 0; b < op->L.n; b++)
    for (int idir = 0; idir < 3; idir++)
      for (int side = 0; side < 2; side++) {
        const CFIVS* c = &op->cf[b*6+idir*2+side];
        Fab* f = &phi->f[b];
        int ihilo = side ? 1 : -1;
        if (c->empty) continue;
        if (c->packed) {
          if (b_len(&f->b, idir) == 3) k_interphomolinear(f, &c->packedBox, op->dx, op->dxCrse, idir, ihilo);
          else k_interphomo(f, &c->packedBox, op->dx, op->dxCrse, idir, ihilo);
        } else {
          const Box mb = c->ivs.mb;
          if (b_len(&f->b, idir) == 3) {
            double factor = 1-2*op->dx/(op->dx+op->dxCrse);
            for (int k = mb.lo[2]; k <= mb.hi[2]; k++)
              for (int j = mb.lo[1]; j <= mb.hi[1]; j++)
                for (int i = mb.lo[0]; i <= mb.hi[0]; i++) {
                  if (!ivs_has(&c->ivs, i, j, k)) continue;
                  int a[3] = {i,j,k}; a[idir] -= ihilo;
                  *P(f,i,j,k) = factor * *P(f,a[0],a[1],a[2]);
                }
          } else {
            double c1 = 2*(op->dxCrse-op->dx)/(op->dxCrse+op->dx);
            double c2 = -(op->dxCrse-op->dx)/(op->dxCrse+3*op->dx);
            for (int k = mb.lo[2]; k <= mb.hi[2]; k++)
              for (int j = mb.lo[1]; j <= mb.hi[1]; j++)
                for (int i = mb.lo[0]; i <= mb.hi[0]; i++) {
                  if (!ivs_has(&c->ivs, i, j, k)) continue;
                  int a[3] = {i,j,k}, bb[3] = {i,j,k};
                  a[idir] -= 2*ihilo; bb[idir] -= ihilo;
                  double pa = *P(f,a[0],a[1],a[2]);
                  double pb = *P(f,bb[0],bb[1],bb[2]);
                  *P(f,i,j,k) = c1*pb + c2*pa;
                }
          }
        }
      }
}
/* ---- VCAMRPoissonOp2 kernels (lib/src/AMRElliptic/VCAMRPoissonOpF.ChF) ---- */
/* the operator at (i,j,k): alpha*a*phi - beta*(sum of face fluxes)*dxinv, the six
   flux terms in the order of the CHF_DTERM expansion (VCAMRPoissonOpF.ChF:236-254) */
static inline double vc_lofphi(const Fab* phi, const Fab* a, Fab* const* b, int i, int j, int k,
                               double alpha, double beta, double dxinv) {
  const double p = *P(phi,i,j,k);
  const double s = *P(&b[0][0],i+1,j,k) * (*P(phi,i+1,j,k) - p)
                 - *P(&b[0][0],i,j,k) * (p - *P(phi,i-1,j,k))
                 + *P(&b[1][0],i,j+1,k) * (*P(phi,i,j+1,k) - p)
                 - *P(&b[1][0],i,j,k) * (p - *P(phi,i,j-1,k))
                 + *P(&b[2][0],i,j,k+1) * (*P(phi,i,j,k+1) - p)
                 - *P(&b[2][0],i,j,k) * (p - *P(phi,i,j,k-1));
  return alpha * *P(a,i,j,k) * p - beta * s * dxinv;
}
/* the three face arrays of box q */
#define VCB(op, q) ((Fab* const[3]){&(op)->vb[0][q], &(op)->vb[1][q], &(op)->vb[2][q]})
/* VCCOMPUTEOP (VCAMRPoissonOpF.ChF:197-265) */
static void k_vcop(Fab* lofphi, const Fab* phi, const Fab* a, Fab* const* b, const Box* r, double dx,
                   double alpha, double beta) {
  const double dxinv = 1.0/(dx*dx);
  for (int k = r->lo[2]; k <= r->hi[2]; k++)
    for (int j = r->lo[1]; j <= r->hi[1]; j++)
      for (int i = r->lo[0]; i <= r->hi[0]; i++)
        *P(lofphi,i,j,k) = vc_lofphi(phi, a, b, i, j, k, alpha, beta, dxinv);
}
/* VCCOMPUTERES (VCAMRPoissonOpF.ChF:311-382): res = rhs - L(phi) */
static void k_vcres(Fab* res, const Fab* phi, const Fab* rhs, const Fab* a, Fab* const* b, const Box* r,
                    double dx, double alpha, double beta) {
  const double dxinv = 1.0/(dx*dx);
  for (int k = r->lo[2]; k <= r->hi[2]; k++)
    for (int j = r->lo[1]; j <= r->hi[1]; j++)
      for (int i = r->lo[0]; i <= r->hi[0]; i++)
        *P(res,i,j,k) = *P(rhs,i,j,k) - vc_lofphi(phi, a, b, i, j, k, alpha, beta, dxinv);
}
/* GSRBHELMHOLTZVC (VCAMRPoissonOpF.ChF:57-155): phi -= lambda*(L(phi) - rhs) on one colour */
static void k_vcgsrb(Fab* phi, const Fab* rhs, const Fab* a, Fab* const* b, const Fab* lam, const Box* r,
                     double dx, double alpha, double beta, int redBlack) {
  const double dxinv = 1.0/(dx*dx);
  for (int k = r->lo[2]; k <= r->hi[2]; k++)
    for (int j = r->lo[1]; j <= r->hi[1]; j++) {
      int imin = r->lo[0];
      int indtot = imin + j + k;
      imin = imin + abs((indtot + redBlack) % 2);
      for (int i = imin; i <= r->hi[0]; i += 2) {
        const double lofphi = vc_lofphi(phi, a, b, i, j, k, alpha, beta, dxinv);
        *P(phi,i,j,k) = *P(phi,i,j,k) - *P(lam,i,j,k) * (lofphi - *P(rhs,i,j,k));
      }
    }
}
/* RESTRICTRESVC (VCAMRPoissonOpF.ChF:422-496): the coarse residual accumulates
   (rhs - L(phi))/8 over the fine region in Fortran order (res zeroed first) */
static void k_vcrestrictres(Fab* res, const Fab* phi, const Fab* rhs, const Fab* a, Fab* const* b,
                            const Box* r, double dx, double alpha, double beta) {
  const double dxinv = 1.0/(dx*dx);
  const double denom = 2*2*2;
  for (int k = r->lo[2]; k <= r->hi[2]; k++)
    for (int j = r->lo[1]; j <= r->hi[1]; j++)
      for (int i = r->lo[0]; i <= r->hi[0]; i++) {
        const double lofphi = vc_lofphi(phi, a, b, i, j, k, alpha, beta, dxinv);
        double* c = P(res, fdiv(r->lo[0], 2) + (i - r->lo[0])/2, fdiv(r->lo[1], 2) + (j - r->lo[1])/2,
                      fdiv(r->lo[2], 2) + (k - r->lo[2])/2);
        *c = *c + (*P(rhs,i,j,k) - lofphi) / denom;
      }
}
/* resetLambda (VCAMRPoissonOp2.cpp:283-318) with SUMFACES (VCAMRPoissonOpF.ChF:498-525):
   lambda = 1 / (alpha*a + sum_d scale*beta*(b_d(i+e_d) + b_d(i))) */
static void k_vclambda(Fab* lam, const Fab* a, Fab* const* b, const Box* r, double dx, double alpha,
                       double beta) {
  const double scale = 1.0/(dx*dx);
  for (int k = r->lo[2]; k <= r->hi[2]; k++)
    for (int j = r->lo[1]; j <= r->hi[1]; j++)
      for (int i = r->lo[0]; i <= r->hi[0]; i++) *P(lam,i,j,k) = *P(a,i,j,k) * alpha;
  for (int d = 0; d < 3; d++) {
    const int e0 = d == 0, e1 = d == 1, e2 = d == 2;
    for (int k = r->lo[2]; k <= r->hi[2]; k++)
      for (int j = r->lo[1]; j <= r->hi[1]; j++)
        for (int i = r->lo[0]; i <= r->hi[0]; i++) {
          const double sumVal = *P(b[d],i+e0,j+e1,k+e2) + *P(b[d],i,j,k);
          *P(lam,i,j,k) = *P(lam,i,j,k) + scale * beta * sumVal;
        }
  }
  for (int k = r->lo[2]; k <= r->hi[2]; k++)
    for (int j = r->lo[1]; j <= r->hi[1]; j++)
      for (int i = r->lo[0]; i <= r->hi[0]; i++) *P(lam,i,j,k) = 1.0 / *P(lam,i,j,k);
}
/* CoarseAverage::averageToCoarse -> average (AverageF.ChF:9-162): the r^3 children
   summed with i fastest; refRatio 2 and 4 are unrolled sums multiplied by refScale,
   other ratios a loop from zero times refScale.  harmonic: averageToCoarseHarmonic ->
   AVERAGEHARMONIC (AverageF.ChF:165-314), the same sums of 1/fine and the reciprocal
   of refScale times the sum (0 + x == x, so both branches give 1/(refScale*sum)). */
static void k_vcavg_cells(Fab* coarse, const Fab* fine, const Box* cb, int r, int harmonic) {
  const double refScale = 1.0/(double)(r*r*r);
  for (int K = cb->lo[2]; K <= cb->hi[2]; K++)
    for (int J = cb->lo[1]; J <= cb->hi[1]; J++)
      for (int I = cb->lo[0]; I <= cb->hi[0]; I++) {
        double sum = 0.0;
        int first = 1;
        for (int kk = 0; kk < r; kk++)
          for (int jj = 0; jj < r; jj++)
            for (int ii = 0; ii < r; ii++) {
              const double v = *P(fine, I*r+ii, J*r+jj, K*r+kk);
              const double f = harmonic ? 1.0/v : v;
              if (first && r <= 4) { sum = f; first = 0; }
              else sum = sum + f;
            }
        if (harmonic) *P(coarse,I,J,K) = (r <= 4) ? 1.0/(refScale * sum) : 1.0/(sum * refScale);
        else *P(coarse,I,J,K) = (r <= 4) ? refScale * sum : sum * refScale;
      }
}
/* CoarseAverageFace::averageGridData -> AVERAGEFACE (AverageFaceF.ChF): each coarse
   face sums the r^2 fine faces on it (i fastest) from zero, times (1/r)^2.
   harmonic: AVERAGEFACEHARMONIC (AverageFaceF.ChF:84-128), the sum of 1/fine and
   1/(ref_scale*sum). */
static void k_vcavg_faces(Fab* coarse, const Fab* fine, const Box* cfb, int dir, int r, int harmonic) {
  const double ref_scale = (1.0/r)*(1.0/r);
  const int n0 = dir == 0 ? 1 : r, n1 = dir == 1 ? 1 : r, n2 = dir == 2 ? 1 : r;
  for (int K = cfb->lo[2]; K <= cfb->hi[2]; K++)
    for (int J = cfb->lo[1]; J <= cfb->hi[1]; J++)
      for (int I = cfb->lo[0]; I <= cfb->hi[0]; I++) {
        double crseSum = 0;
        for (int kk = 0; kk < n2; kk++)
          for (int jj = 0; jj < n1; jj++)
            for (int ii = 0; ii < n0; ii++) {
              const double v = *P(fine, I*r+ii, J*r+jj, K*r+kk);
              crseSum = crseSum + (harmonic ? 1.0/v : v);
            }
        *P(coarse,I,J,K) = harmonic ? 1.0/(ref_scale*crseSum) : ref_scale*crseSum;
      }
}
static Box vc_facebox(Box b, int d) { b.hi[d] += 1; return b; }
/* define the VC data of op (coefficients given) and its lambda */
static void op_vc_alloc(Op* op) {
  op->vc = 1;
  op->va = ld_make(&op->L, 0);
  op->vlam = ld_make(&op->L, 0);
  for (int d = 0; d < 3; d++) {
    op->vb[d] = calloc(op->L.n ? op->L.n : 1, sizeof(Fab));
    for (int q = 0; q < op->L.n; q++) fab_alloc(&op->vb[d][q], vc_facebox(op->L.b[q], d));
  }
}
static void op_vc_lambda(Op* op) {
  for (int q = 0; q < op->L.n; q++)
    k_vclambda(&op->vlam.f[q], &op->va.f[q], VCB(op, q), &op->L.b[q], op->dx, op->alpha, op->beta);
}

/* residualI (AMRPoissonOp.cpp:260-297; VCAMRPoissonOp2.cpp:30-87) */
static void op_residualI(const Op* op, LD* lhs, LD* phi, const LD* rhs, int homog) {
  if (phi->nc > 1) { for (int c = 0; c < phi->nc; c++) op_residualI(op, COMP(lhs, c), COMP(phi, c), COMP(rhs, c), homog); return; }
  op_exchange(op, phi);
  op_bc(op, phi, homog);
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < op->L.n; b++) {
    if (op->vc)
      k_vcres(&lhs->f[b], &phi->f[b], &rhs->f[b], &op->va.f[b], VCB(op, b), &op->L.b[b], op->dx, op->alpha,
              op->beta);
    else
      k_oplapres(&lhs->f[b], &phi->f[b], &rhs->f[b], &op->L.b[b], op->dx, op->alpha, op->beta);
  }
}
static void op_residual(const Op* op, LD* lhs, LD* phi, const LD* rhs, int homog) {
  if (phi->nc > 1) { for (int c = 0; c < phi->nc; c++) op_residual(op, COMP(lhs, c), COMP(phi, c), COMP(rhs, c), homog); return; }
  if (homog) op_homog_cf(op, phi);
  op_residualI(op, lhs, phi, rhs, homog);
}
/* applyOpI (AMRPoissonOp.cpp:343-385) */
static void op_applyI(const Op* op, LD* lhs, LD* phi, int homog) {
  if (phi->nc > 1) { for (int c = 0; c < phi->nc; c++) op_applyI(op, COMP(lhs, c), COMP(phi, c), homog); return; }
  op_exchange(op, phi);
  op_bc(op, phi, homog);
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < op->L.n; b++) {
    if (op->vc)
      k_vcop(&lhs->f[b], &phi->f[b], &op->va.f[b], VCB(op, b), &op->L.b[b], op->dx, op->alpha, op->beta);
    else
      k_oplap(&lhs->f[b], &phi->f[b], &op->L.b[b], op->dx, op->alpha, op->beta);
  }
}
static void op_apply(const Op* op, LD* lhs, LD* phi, int homog) {
  if (phi->nc > 1) { for (int c = 0; c < phi->nc; c++) op_apply(op, COMP(lhs, c), COMP(phi, c), homog); return; }
  op_homog_cf(op, phi);
  op_applyI(op, lhs, phi, homog);
}
/* cell-relaxations performed (one red+black iteration of one cell), for timing */
static long long g_cell_relax = 0;
long long oracle_cell_relax_count(void) { return g_cell_relax; }
void oracle_reset_counters(void) { g_cell_relax = 0; }

/* levelGSRB (AMRPoissonOp.cpp:1216-1281) */
static void op_levelGSRB(const Op* op, LD* phi, const LD* rhs) {
  g_cell_relax += lay_cells(&op->L);
  for (int pass = 0; pass <= 1; pass++) {
    op_homog_cf(op, phi);
    op_exchange(op, phi);
    #pragma omp parallel for schedule(static)
    for (int b = 0; b < op->L.n; b++) {
      apply_bc(&phi->f[b], &op->L.b[b], &op->domain, op->dx, 1, &op->bc, phi->comp);
      if (op->vc)  /* VCAMRPoissonOp2::levelGSRB (VCAMRPoissonOp2.cpp:547-626) */
        k_vcgsrb(&phi->f[b], &rhs->f[b], &op->va.f[b], VCB(op, b), &op->vlam.f[b], &op->L.b[b], op->dx,
                 op->alpha, op->beta, pass);
      else if (op->alpha == 0.0 && op->beta == 1.0)
        k_gsrb_lap(&phi->f[b], &rhs->f[b], &op->L.b[b], op->dx, pass);
      else
        k_gsrb_helm(&phi->f[b], &rhs->f[b], &op->L.b[b], op->dx, op->alpha, op->beta, pass);
    }
  }
}
static void ldo_incr(LD* lhs, const LD* x, double scale);
/* levelJacobi (AMRPoissonOp.cpp:1579-1596; VCAMRPoissonOp2.cpp:762-786): the
   homogeneous residual, weighted, added to phi */
static void op_levelJacobi(const Op* op, LD* phi, const LD* rhs) {
  g_cell_relax += lay_cells(&op->L);
  LD resid = ld_make(&op->L, 0);
  op_residual(op, &resid, phi, rhs, 1);
  if (op->vc) {
    for (int b = 0; b < op->L.n; b++) {
      const Box* v = &op->L.b[b];
      for (int k = v->lo[2]; k <= v->hi[2]; k++)
        for (int j = v->lo[1]; j <= v->hi[1]; j++)
          for (int i = v->lo[0]; i <= v->hi[0]; i++)
            *P(&resid.f[b],i,j,k) = *P(&resid.f[b],i,j,k) * *P(&op->vlam.f[b],i,j,k);
    }
    ldo_incr(phi, &resid, 0.5);
  } else {
    double weight = op->alpha - 2.0*3 * op->beta / (op->dx*op->dx);
    ldo_incr(phi, &resid, 0.666/weight);
  }
  ld_free(&resid);
}
/* one GSRB pass of box b over region r: GSRBLAPLACIAN when lap, else the
   alpha/beta selection of levelGSRB (AMRPoissonOp.cpp:1260) */
static void box_gsrb(const Op* op, LD* phi, const LD* rhs, int b, const Box* r, int lap, int pass) {
  if (r->lo[0] > r->hi[0] || r->lo[1] > r->hi[1] || r->lo[2] > r->hi[2]) return;
  if (lap || (op->alpha == 0.0 && op->beta == 1.0))
    k_gsrb_lap(&phi->f[b], &rhs->f[b], r, op->dx, pass);
  else
    k_gsrb_helm(&phi->f[b], &rhs->f[b], r, op->dx, op->alpha, op->beta, pass);
}
/* looseGSRB (AMRPoissonOp.cpp:1333-1413): one ghost refresh (homogeneous CF
   interpolation, exchange, homogeneous BC), then both passes box by box with
   those ghost values (no refresh between the colours) */
static void op_looseGSRB(const Op* op, LD* phi, const LD* rhs) {
  g_cell_relax += lay_cells(&op->L);
  op_homog_cf(op, phi);
  op_exchange(op, phi);
  for (int b = 0; b < op->L.n; b++) {
    apply_bc(&phi->f[b], &op->L.b[b], &op->domain, op->dx, 1, &op->bc, phi->comp);
    box_gsrb(op, phi, rhs, b, &op->L.b[b], 0, 0);
    box_gsrb(op, phi, rhs, b, &op->L.b[b], 0, 1);
  }
}
/* overlapGSRB (AMRPoissonOp.cpp:1416-1521) without MPI, where exchangeBegin has
   already copied every ghost (LevelDataI.H:533-538): looseGSRB unless the level
   is one box or (the reference's test) alpha != 1 or beta != 0, then exchange,
   homogeneous CF interpolation, per box the homogeneous BC and GSRBLAPLACIAN
   (both passes) on the box grown by -1, then per box the six one-cell face
   layers (x lo, x hi, y lo, ..., each shrunk by one in the lower directions),
   both passes each, with GSRBLAPLACIAN whatever alpha and beta */
static void op_overlapGSRB(const Op* op, LD* phi, const LD* rhs) {
  if (op->L.n == 1 || op->alpha != 1 || op->beta != 0) op_looseGSRB(op, phi, rhs);
  g_cell_relax += lay_cells(&op->L);
  op_exchange(op, phi);
  op_homog_cf(op, phi);
  for (int b = 0; b < op->L.n; b++) {
    apply_bc(&phi->f[b], &op->L.b[b], &op->domain, op->dx, 1, &op->bc, phi->comp);
    Box r = op->L.b[b];
    for (int d = 0; d < 3; d++) { r.lo[d]++; r.hi[d]--; }
    box_gsrb(op, phi, rhs, b, &r, 1, 0);
    box_gsrb(op, phi, rhs, b, &r, 1, 1);
  }
  for (int b = 0; b < op->L.n; b++)
    for (int dir = 0; dir < 3; dir++)
      for (int side = 0; side < 2; side++) {
        Box r = op->L.b[b];
        if (side == 0) r.hi[dir] = r.lo[dir];
        else r.lo[dir] = r.hi[dir];
        for (int e = 0; e < dir; e++) { r.lo[e]++; r.hi[e]--; }
        box_gsrb(op, phi, rhs, b, &r, 1, 0);
        box_gsrb(op, phi, rhs, b, &r, 1, 1);
      }
}
/* levelGSRBLazy (AMRPoissonOp.cpp:1524-1578): eight colours, the parities of
   (i, j, k) in the order nextColorLoc steps them (x fastest); for each, the whole
   homogeneous operator (applyOp, with its CF interpolation, exchange and BC) and
   then GSRBLAZY on the cells of that colour */
static void op_levelGSRBLazy(const Op* op, LD* phi, const LD* rhs) {
  g_cell_relax += lay_cells(&op->L);
  LD lphi = ld_make(&op->L, 0);
  for (int c = 0; c < 8; c++) {
    const int color[3] = {c & 1, (c >> 1) & 1, (c >> 2) & 1};
    op_apply(op, &lphi, phi, 1);
    for (int b = 0; b < op->L.n; b++) {
      Box cb;
      if (colored_box(&op->L.b[b], color, &cb))
        k_gsrb_lazy(&phi->f[b], &lphi.f[b], &rhs->f[b], &cb, op->alpha, op->beta, op->dx);
    }
  }
  ld_free(&lphi);
}
/* amrpgetMultiColors (AMRPoissonOp.cpp:36-56), 3D order */
static const int k_multi_colors[8][3] = {{0,0,0},{1,1,0},{0,1,1},{1,0,1},{0,1,0},{1,0,0},{0,0,1},{1,1,1}};
/* levelMultiColor (AMRPoissonOp.cpp:1284-1331): for each of the eight colours,
   homogeneous CF interpolation, exchange, homogeneous BC, then GSMCAMRPOP */
static void op_levelMultiColor(const Op* op, LD* phi, const LD* rhs) {
  g_cell_relax += lay_cells(&op->L);
  for (int c = 0; c < 8; c++) {
    op_homog_cf(op, phi);
    op_exchange(op, phi);
    for (int b = 0; b < op->L.n; b++) {
      Box cb;
      apply_bc(&phi->f[b], &op->L.b[b], &op->domain, op->dx, 1, &op->bc, phi->comp);
      if (colored_box(&op->L.b[b], k_multi_colors[c], &cb))
        k_gsmc(&phi->f[b], &rhs->f[b], &cb, op->dx, op->alpha, op->beta);
    }
  }
}
/* relax (AMRPoissonOp.cpp:600-632): s_relaxMode 0 (looseGSRB), 1 (levelGSRB),
   2 (overlapGSRB), 3 (levelGSRBLazy), 4 (levelJacobi) or 5 (levelMultiColor) */
static void op_relax(const Op* op, LD* e, const LD* r, int n) {
  if (e->nc > 1) { for (int c = 0; c < e->nc; c++) op_relax(op, COMP(e, c), COMP(r, c), n); return; }
  for (int i = 0; i < n; i++) {
    if (op->relaxMode == 0) op_looseGSRB(op, e, r);
    else if (op->relaxMode == 2) op_overlapGSRB(op, e, r);
    else if (op->relaxMode == 4) op_levelJacobi(op, e, r);
    else if (op->relaxMode == 3) op_levelGSRBLazy(op, e, r);
    else if (op->relaxMode == 5) op_levelMultiColor(op, e, r);
    else op_levelGSRB(op, e, r);
  }
}
/* restrictResidual (AMRPoissonOp.cpp:651-703) */
static void op_restrict(const Op* op, LD* resC, LD* phi, const LD* rhs) {
  if (phi->nc > 1) { for (int c = 0; c < phi->nc; c++) op_restrict(op, COMP(resC, c), COMP(phi, c), COMP(rhs, c)); return; }
  op_homog_cf(op, phi);
  op_exchange(op, phi);
  op_bc(op, phi, 1);
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < op->L.n; b++) {
    fab_setval(&resC->f[b], resC->f[b].b, 0.0);
    if (op->vc)  /* VCAMRPoissonOp2::restrictResidual (VCAMRPoissonOp2.cpp:195-256) */
      k_vcrestrictres(&resC->f[b], &phi->f[b], &rhs->f[b], &op->va.f[b], VCB(op, b), &op->L.b[b], op->dx,
                      op->alpha, op->beta);
    else
      k_restrictres(&resC->f[b], &phi->f[b], &rhs->f[b], op->alpha, op->beta, &op->L.b[b], op->dx);
  }
}
/* prolongIncrement (AMRPoissonOp.cpp:706-733) */
static void op_prolong(const Op* op, LD* phi, const LD* coarse) {
  if (phi->nc > 1) { for (int c = 0; c < phi->nc; c++) op_prolong(op, COMP(phi, c), COMP(coarse, c)); return; }
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < op->L.n; b++) k_prolong(&phi->f[b], &coarse->f[b], &op->L.b[b], 2);
}

/* ---- LevelDataOps<FArrayBox> (LevelDataOps.H:93-294) ---- */
static void ldo_setzero(LD* x) {
  if (x->nc > 1) { for (int c = 0; c < x->nc; c++) ldo_setzero(COMP(x, c)); return; }
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < x->L->n; b++) memset(x->f[b].d, 0, sizeof(double)*b_npts(&x->f[b].b));
}
/* incr: lhs += scale*x on valid boxes (BaseFab::plus) */
static void ldo_incr(LD* lhs, const LD* x, double scale) {
  if (lhs->nc > 1) { for (int c = 0; c < lhs->nc; c++) ldo_incr(COMP(lhs, c), COMP(x, c), scale); return; }
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < lhs->L->n; b++) {
    const Box* r = &lhs->L->b[b];
    for (int k = r->lo[2]; k <= r->hi[2]; k++)
      for (int j = r->lo[1]; j <= r->hi[1]; j++)
        for (int i = r->lo[0]; i <= r->hi[0]; i++) *P(&lhs->f[b],i,j,k) += *P(&x->f[b],i,j,k)*scale;
  }
}
/* axby on whole FABs: copy x, mult a, plus(y, b) */
static void ldo_axby(LD* lhs, const LD* x, const LD* y, double a, double bb) {
  if (lhs->nc > 1) { for (int c = 0; c < lhs->nc; c++) ldo_axby(COMP(lhs, c), COMP(x, c), COMP(y, c), a, bb); return; }
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < lhs->L->n; b++) {
    Fab* f = &lhs->f[b];
    Box r = b_isect(f->b, &x->f[b].b);
    fab_copy_region(f, &x->f[b], r);
    long n = b_npts(&f->b);
    for (long q = 0; q < n; q++) f->d[q] *= a;
    Box ry = b_isect(f->b, &y->f[b].b);
    for (int k = ry.lo[2]; k <= ry.hi[2]; k++)
      for (int j = ry.lo[1]; j <= ry.hi[1]; j++)
        for (int i = ry.lo[0]; i <= ry.hi[0]; i++) *P(f,i,j,k) += *P(&y->f[b],i,j,k)*bb;
  }
}
static void ldo_scale(LD* x, double s) {
  if (x->nc > 1) { for (int c = 0; c < x->nc; c++) ldo_scale(COMP(x, c), s); return; }
  for (int b = 0; b < x->L->n; b++) { long n = b_npts(&x->f[b].b); for (long q = 0; q < n; q++) x->f[b].d[q] *= s; }
}
/* assignLocal: FArrayBox::copy over the intersection of the two FAB boxes */
static void ldo_assign_local(LD* lhs, const LD* rhs) {
  if (lhs->nc > 1) { for (int c = 0; c < lhs->nc; c++) ldo_assign_local(COMP(lhs, c), COMP(rhs, c)); return; }
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < lhs->L->n; b++) {
    Box r = b_isect(lhs->f[b].b, &rhs->f[b].b);
    fab_copy_region(&lhs->f[b], &rhs->f[b], r);
  }
}
/* FArrayBox::dotProduct over the valid box (ForAllThisCBNNXC: serial Fortran order);
   LevelDataOps::dotProduct sums boxes in layout order */
static double ldo_dot(const LD* a, const LD* c) {
  double val = 0.0;
  const int n = a->L->n;
  for (int b = 0; b < n; b++) {
    const Box* r = &a->L->b[b];
    double dot = 0.0;
    for (int m = 0; m < a->nc; m++) {      /* components outermost, one running sum */
      const Fab* fa = &a->f[(size_t)m*n + b];
      const Fab* fc = &c->f[(size_t)m*n + b];
      for (int k = r->lo[2]; k <= r->hi[2]; k++)
        for (int j = r->lo[1]; j <= r->hi[1]; j++)
          for (int i = r->lo[0]; i <= r->hi[0]; i++) dot += *P(fa,i,j,k) * *P(fc,i,j,k);
    }
    val += dot;
  }
  return val;
}
/* FArrayBox::norm(box, 0) : max |x| ; order independent */
static double fab_maxnorm(const Fab* f, const Box* r) {
  double m = 0.0;
  for (int k = r->lo[2]; k <= r->hi[2]; k++)
    for (int j = r->lo[1]; j <= r->hi[1]; j++)
      for (int i = r->lo[0]; i <= r->hi[0]; i++) { double a = fabs(*P(f,i,j,k)); if (a > m) m = a; }
  return m;
}
/* FArrayBox::sumPow(box, 2) (FArrayBox.cpp:165-230; ABS: FArrayBox::norm(box, 1),
   :110-141) of box b over every component: the pencils of ForAllThisCPencil, component
   outermost, then z, then y, accumulate into one per-column array (the first pencil
   sets it), then the columns are summed left to right */
static double ld_box_sumpow(const LD* x, int b, int absval) {
  const Box* r = &x->L->b[b];
  int nx = b_len(r, 0);
  double* tmp = malloc(sizeof(double)*nx);
  int first = 1;
  for (int m = 0; m < x->nc; m++) {
    const Fab* f = &x->f[(size_t)m*x->L->n + b];
    for (int k = r->lo[2]; k <= r->hi[2]; k++)
      for (int j = r->lo[1]; j <= r->hi[1]; j++) {
        const double* row = P(f, r->lo[0], j, k);
        if (first) {
          for (int i = 0; i < nx; i++) tmp[i] = absval ? fabs(row[i]) : row[i]*row[i];
          first = 0;
        } else {
          for (int i = 0; i < nx; i++) tmp[i] += absval ? fabs(row[i]) : row[i]*row[i];
        }
      }
  }
  double sum = tmp[0];
  for (int i = 1; i < nx; i++) sum += tmp[i];
  free(tmp);
  return sum;
}
static double ld_box_maxnorm(const LD* x, int b) {
  double m = 0.0;
  for (int c = 0; c < x->nc; c++) {
    double v = fab_maxnorm(&x->f[(size_t)c*x->L->n + b], &x->L->b[b]);
    if (v > m) m = v;
  }
  return m;
}
/* norm(BoxLayoutData, interval, p) (BoxLayoutData.cpp:23-92), p = 0, 1 or 2 */
static double ldo_norm(const LD* x, int p) {
  double norm = 0;
  if (p == 0) {
    for (int b = 0; b < x->L->n; b++) { double m = ld_box_maxnorm(x, b); norm = norm > m ? norm : m; }
    return norm;
  }
  if (p == 1) {
    for (int b = 0; b < x->L->n; b++) norm += ld_box_sumpow(x, b, 1);
    return norm;
  }
  for (int b = 0; b < x->L->n; b++) norm += ld_box_sumpow(x, b, 0);
  return sqrt(norm);
}
/* AMRPoissonOp::localMaxNorm */
static double op_localMaxNorm(const LD* x) {
  double m = 0;
  #pragma omp parallel for reduction(max:m) schedule(static)
  for (int b = 0; b < x->L->n; b++) { double v = ld_box_maxnorm(x, b); if (v > m) m = v; }
  return m;
}
/* preCond (AMRPoissonOp.cpp:304-329) */
static void op_precond(const Op* op, LD* phi, const LD* rhs) {
  if (phi->nc > 1) { for (int c = 0; c < phi->nc; c++) op_precond(op, COMP(phi, c), COMP(rhs, c)); return; }
  if (op->vc) {  /* VCAMRPoissonOp2::preCond (VCAMRPoissonOp2.cpp:93-127): phi = rhs * lambda */
    for (int b = 0; b < op->L.n; b++) {
      Box r = b_isect(phi->f[b].b, &rhs->f[b].b);
      fab_copy_region(&phi->f[b], &rhs->f[b], r);
      const Box* v = &op->L.b[b];
      for (int k = v->lo[2]; k <= v->hi[2]; k++)
        for (int j = v->lo[1]; j <= v->hi[1]; j++)
          for (int i = v->lo[0]; i <= v->hi[0]; i++) *P(&phi->f[b],i,j,k) = *P(&phi->f[b],i,j,k) * *P(&op->vlam.f[b],i,j,k);
    }
    op_relax(op, phi, rhs, 2);
    return;
  }
  double mult = 1.0/(op->alpha - 2.0*3*op->beta/(op->dx*op->dx));
  for (int b = 0; b < op->L.n; b++) {
    Box r = b_isect(phi->f[b].b, &rhs->f[b].b);
    fab_copy_region(&phi->f[b], &rhs->f[b], r);
    long n = b_npts(&phi->f[b].b);
    for (long q = 0; q < n; q++) phi->f[b].d[q] *= mult;
  }
  op_relax(op, phi, rhs, 2);
}

/* reflux (AMRPoissonOp.cpp:1117-1201) + NEWGETFLUX + LevelFluxRegister */
static void op_reflux(Op* cop, Op* fop, LD* phiFine, LD* phi, LD* res) {
  if (phi->nc > 1) { for (int c = 0; c < phi->nc; c++) op_reflux(cop, fop, COMP(phiFine, c), COMP(phi, c), COMP(res, c)); return; }
  FluxReg* r = &cop->fr;
  ldo_setzero(&r->coarFlux);
  ldo_setzero(&r->fineFlux);
  /* incrementCoarse for every coarse box, direction, side */
  for (int c = 0; c < cop->L.n; c++) {
    const Fab* cf = &phi->f[c];
    for (int idir = 0; idir < 3; idir++) {
      double scale_flux = cop->beta*1/cop->dx;            /* getFlux(..., a_ref = 1) */
      for (int side = 0; side < 2; side++) {
        const BoxList* bl = &r->loc[c*6+idir*2+side];
        double sc = side ? -1.0 : 1.0;                     /* -sign(side)*1 */
        for (int q = 0; q < bl->n; q++) {
          const Box* bx = &bl->b[q];
          for (int k = bx->lo[2]; k <= bx->hi[2]; k++)
            for (int j = bx->lo[1]; j <= bx->hi[1]; j++)
              for (int i = bx->lo[0]; i <= bx->hi[0]; i++) {
                int n0[3] = {i,j,k};
                if (side == 0) n0[idir] += 1;                /* shiftHalf(-1): node n -> cell n-1 */
                int m1[3] = {n0[0],n0[1],n0[2]}; m1[idir] -= 1;
                double F = (*P(cf,n0[0],n0[1],n0[2]) - *P(cf,m1[0],m1[1],m1[2]))*scale_flux;
                /* VCAMRPoissonOp2::getFlux (VCAMRPoissonOp2.cpp:787-837): -b(face) * gradphi */
                if (cop->vc) F = -*P(&cop->vb[idir][c],n0[0],n0[1],n0[2]) * F;
                double* reg = P(&r->coarFlux.f[c],i,j,k);
                *reg += F*sc;
              }
        }
      }
    }
  }
  /* CF ghosts of the finer level from this level */
  qcfi_interp(&fop->qcfi, phiFine, phi);
  /* incrementFine for every fine box, direction, side */
  for (int f = 0; f < phiFine->L->n; f++) {
    const Box* gb = &phiFine->L->b[f];
    const Fab* pf = &phiFine->f[f];
    for (int idir = 0; idir < 3; idir++) {
      double scale_flux = cop->beta*cop->refToFiner/cop->dx;
      for (int side = 0; side < 2; side++) {
        /* sign(side)*a_scale/denom, denom = nRef^(D-1) (LevelFluxRegister.cpp:468-481) */
        const int nr = r->nRefine;
        double denom = 1.0;
        for (int d = 0; d < 3; d++) if (d != idir) denom *= nr;
        double scale = (side ? 1.0 : -1.0)*1.0/denom;
        Box clip = b_refine(r->cfine.b[f], nr);
        Box fineBox = b_adj(clip, idir, side, 1);  /* cells outside the box on this side */
        int iv[3] = {fineBox.lo[0], fineBox.lo[1], fineBox.lo[2]};
        int civ[3] = {fdiv(iv[0],nr), fdiv(iv[1],nr), fdiv(iv[2],nr)};
        for (int k = fineBox.lo[2]; k <= fineBox.hi[2]; k++)
          for (int j = fineBox.lo[1]; j <= fineBox.hi[1]; j++)
            for (int i = fineBox.lo[0]; i <= fineBox.hi[0]; i++) {
              int n0[3] = {i,j,k};
              if (side == 0) n0[idir] += 1;
              int m1[3] = {n0[0],n0[1],n0[2]}; m1[idir] -= 1;
              double F = (*P(pf,n0[0],n0[1],n0[2]) - *P(pf,m1[0],m1[1],m1[2]))*scale_flux;
              if (fop->vc) F = -*P(&fop->vb[idir][f],n0[0],n0[1],n0[2]) * F;
              int ii = civ[0] + (i-iv[0])/nr, jj = civ[1] + (j-iv[1])/nr, kk = civ[2] + (k-iv[2])/nr;
              double* cfin = P(&r->fineFlux.f[f], ii, jj, kk);
              *cfin = *cfin + scale*F;
            }
        (void)gb;
      }
    }
  }
  /* LevelFluxRegister::reflux(res, 1/dx) */
  double a_scale = 1.0/cop->dx;
  for (int c = 0; c < cop->L.n; c++) {
    Fab* u = &res->f[c];
    const Fab* reg = &r->coarFlux.f[c];
    Box bx = b_isect(u->b, &reg->b);
    for (int k = bx.lo[2]; k <= bx.hi[2]; k++)
      for (int j = bx.lo[1]; j <= bx.hi[1]; j++)
        for (int i = bx.lo[0]; i <= bx.hi[0]; i++) *P(u,i,j,k) += *P(reg,i,j,k)*(-a_scale);
  }
  for (int q = 0; q < r->rev.n; q++) {
    const Motion* m = &r->rev.m[q];
    Fab* u = &res->f[m->to];
    const Fab* src = &r->fineFlux.f[m->from];
    const Box* bx = &m->region;
    for (int k = bx->lo[2]; k <= bx->hi[2]; k++)
      for (int j = bx->lo[1]; j <= bx->hi[1]; j++)
        for (int i = bx->lo[0]; i <= bx->hi[0]; i++) *P(u,i,j,k) += *P(src,i,j,k)*(-a_scale);
  }
}

/* ======================================================================= */
/* Hierarchy: AMRPoissonOpFactory + AMRMultiGrid + MultiGrid + BiCGStab    */
/* ======================================================================= */
struct OHier {
  int vcHarmonic;     /* VCAMRPoissonOp2Factory's coefficient average type: 0 arithmetic, 1 harmonic */
  int nlev;
  int* ratio;         /* ratio[l]: refinement ratio between level l-1 and l (2 or 4), l >= 1 */
  long* mgBottomCells; /* m_mg[l]'s m_bottomCells (MultiGrid.H:477-508), for AMRMultiGrid::relax */
  Layout* grids;      /* per AMR level */
  Box* domains;
  double* dx;
  double alpha, beta;
  BCSpec bc;
  Op* op;             /* AMR level ops */
  /* MultiGrid of AMR level 0 */
  int mgD;            /* depth count (m_depth) */
  Layout* mgL;        /* mgL[0] = grids[0] */
  Op* mgop;           /* mgop[0] aliases op[0] */
  long bottomCells;
  /* MultiGrid of AMR level l > 0 (m_mg[l] at its full depth, for solves with
     l_base = l), made on first use: depth count, layouts, ops (index 0 unused) */
  int* lmgD;
  Layout** lmgL;
  Op** lmgop;
  /* AMRMultiGrid copiers */
  Layout* cfL;        /* coarsened fine layouts (resC) per level >= 1 */
  Plan* resCopier;    /* resC[l] -> grids[l-1] */
  Plan* revCopier;    /* grids[l-1] -> resC[l] */
  /* BiCGStabSolver's public parameters (BiCGStabSolver.H:105-153) for every bottom solve */
  int bic_imax, bic_numRestarts;
  double bic_reps, bic_hang, bic_small;
  /* statistics */
  int bottom_calls;
  int nc;             /* components of the solution, rhs and every solver field */
};

/* ratios: nlev-1 refinement ratios (2 or 4) between consecutive levels (NULL: all 2),
   AMRPoissonOpFactory::define's refRatios (AMRPoissonOp.cpp:1857-1903) */
OHier* oracle_create_r(const int* is, const double* rs, const int* ratios) {
  OHier* h = calloc(1, sizeof(OHier));
  int nlev = is[0];
  if (nlev < 1 || nlev > 8) { set_err("bad nlev"); free(h); return NULL; }
  h->nlev = nlev;
  h->nc = 1;
  h->ratio = calloc(nlev, sizeof(int));
  h->mgBottomCells = calloc(nlev, sizeof(long));
  for (int l = 1; l < nlev; l++) {
    h->ratio[l] = ratios ? ratios[l-1] : 2;
    if (h->ratio[l] != 2 && h->ratio[l] != 4) { set_err("refinement ratio must be 2 or 4"); return NULL; }
  }
  h->grids = calloc(nlev, sizeof(Layout));
  h->lmgD = calloc(nlev, sizeof(int)); h->lmgL = calloc(nlev, sizeof(Layout*)); h->lmgop = calloc(nlev, sizeof(Op*));
  h->domains = calloc(nlev, sizeof(Box));
  h->dx = calloc(nlev, sizeof(double));
  h->domains[0] = mkbox(is[1], is[2], is[3], is[4], is[5], is[6]);
  for (int d = 0; d < 3; d++) { h->bc.type[0][d] = is[7+d]; h->bc.type[1][d] = is[10+d]; }
  h->dx[0] = rs[0]; h->alpha = rs[1]; h->beta = rs[2];
  /* BiCGStabSolver constructor defaults (BiCGStabSolver.H:166-183) */
  h->bic_imax = 80; h->bic_numRestarts = 5; h->bic_reps = 1e-12; h->bic_hang = 1e-8; h->bic_small = 1e-30;
  for (int d = 0; d < 3; d++) { h->bc.val[0][d] = rs[3+d]; h->bc.val[1][d] = rs[6+d]; }
  const int* bp = is + 13 + nlev;
  for (int l = 0; l < nlev; l++) {
    int n = is[13+l];
    h->grids[l].n = n; h->grids[l].b = malloc(sizeof(Box)*(n ? n : 1));
    for (int q = 0; q < n; q++) { h->grids[l].b[q] = mkbox(bp[0],bp[1],bp[2],bp[3],bp[4],bp[5]); bp += 6; }
    for (int q = 1; q < n; q++)
      if (!b_lt(&h->grids[l].b[q-1], &h->grids[l].b[q])) { set_err("boxes not sorted"); return NULL; }
    if (l > 0) { h->domains[l] = b_refine(h->domains[l-1], h->ratio[l]); h->dx[l] = h->dx[l-1]/h->ratio[l]; }
  }
  /* periodic directions (BC type 3 on both faces): the levels above the base stay at
     least 2 coarse cells away from the periodic faces, so their coarse-fine stencils
     and flux registers never reach across them */
  for (int d = 0; d < 3; d++) {
    if ((h->bc.type[0][d] == 3) != (h->bc.type[1][d] == 3)) { set_err("periodic on one face only"); return NULL; }
    if (h->bc.type[0][d] != 3) continue;
    for (int l = 1; l < nlev; l++)
      for (int q = 0; q < h->grids[l].n; q++) {
        const Box* b = &h->grids[l].b[q];
        const int m = 2*h->ratio[l];
        if (b->lo[d] < h->domains[l].lo[d] + m || b->hi[d] > h->domains[l].hi[d] - m) {
          set_err("a refined level comes within 2 coarse cells of a periodic face");
          return NULL;
        }
      }
  }
  /* AMRPoissonOpFactory::AMRnewOp */
  h->op = calloc(nlev, sizeof(Op));
  for (int l = 0; l < nlev; l++) {
    Op* op = &h->op[l];
    op_define(op, &h->grids[l], &h->domains[l], h->dx[l], h->alpha, h->beta, &h->bc);
    op->dxCrse = (l > 0) ? h->dx[l-1] : -1.0;
    if (l > 0) {
      op->hasCoarser = 1; op->refToCoarser = h->ratio[l];
      qcfi_define(&op->qcfi, &h->grids[l], &h->grids[l-1], h->dx[l], h->ratio[l], &h->domains[l]);
    }
    if (l < nlev-1) {
      op->hasFiner = 1; op->refToFiner = h->ratio[l+1];
      fr_define(&op->fr, &h->grids[l+1], &h->grids[l], h->ratio[l+1]);
    }
    /* m_mg[l]'s bottom cell count: the level's domain over the multigrid depths its
       layout allows (MGnewOp: coarsenable by coarsening*2) */
    {
      long cells = b_npts(&h->domains[l]);
      for (int c = 2; lay_coarsenable(&h->grids[l], c*2); c *= 2) cells /= 8;
      h->mgBottomCells[l] = cells;
    }
  }
  /* MultiGrid::define on level 0 via MGnewOp (AMRPoissonOp.cpp:1923-1989) */
  int D = 1;
  while (1) {
    int coarsening = 1 << D;
    if (!lay_coarsenable(&h->grids[0], coarsening*2)) break;
    D++;
  }
  h->mgD = D;
  h->mgL = calloc(D, sizeof(Layout));
  h->mgop = calloc(D, sizeof(Op));
  h->mgL[0] = h->grids[0];
  h->bottomCells = b_npts(&h->domains[0]);
  for (int d = 1; d < D; d++) {
    h->mgL[d] = lay_coarsen(&h->grids[0], 1 << d);
    Box dom = b_coarsen(h->domains[0], 1 << d);
    op_define(&h->mgop[d], &h->mgL[d], &dom, h->dx[0]*(1 << d), h->alpha, h->beta, &h->bc);
    h->mgop[d].dxCrse = -1.0;
    h->bottomCells /= 8;
  }
  /* AMRMultiGrid::init copiers */
  h->cfL = calloc(nlev, sizeof(Layout));
  h->resCopier = calloc(nlev, sizeof(Plan));
  h->revCopier = calloc(nlev, sizeof(Plan));
  for (int l = 1; l < nlev; l++) {
    h->cfL[l] = lay_coarsen(&h->grids[l], h->ratio[l]);
    h->resCopier[l] = plan_copy(&h->cfL[l], &h->grids[l-1], 0);
    h->revCopier[l] = h->resCopier[l];
    h->revCopier[l].m = malloc(sizeof(Motion)*(h->resCopier[l].n ? h->resCopier[l].n : 1));
    if (h->resCopier[l].n) memcpy(h->revCopier[l].m, h->resCopier[l].m, sizeof(Motion)*h->resCopier[l].n);
    h->revCopier[l].cap = h->revCopier[l].n;
    plan_reverse(&h->revCopier[l]);
  }
  return h;
}

/* drops the MultiGrids of the levels above the base (remade on next use with the
   operators' current coefficients, BC and relaxation mode) */
static void mg_levels_free(OHier* h) {
  for (int l = 1; l < h->nlev; l++) {
    if (!h->lmgD[l]) continue;
    for (int d = 1; d < h->lmgD[l]; d++) { op_free(&h->lmgop[l][d]); free(h->lmgL[l][d].b); }
    free(h->lmgL[l]); free(h->lmgop[l]);
    h->lmgD[l] = 0; h->lmgL[l] = NULL; h->lmgop[l] = NULL;
  }
}
void oracle_destroy(OHier* h) {
  if (!h) return;
  mg_levels_free(h);
  free(h->lmgD); free(h->lmgL); free(h->lmgop);
  for (int l = 0; l < h->nlev; l++) op_free(&h->op[l]);
  for (int d = 1; d < h->mgD; d++) { op_free(&h->mgop[d]); free(h->mgL[d].b); }
  for (int l = 1; l < h->nlev; l++) { plan_free(&h->resCopier[l]); plan_free(&h->revCopier[l]); free(h->cfL[l].b); }
  for (int l = 0; l < h->nlev; l++) free(h->grids[l].b);
  free(h->grids); free(h->domains); free(h->dx); free(h->op); free(h->mgL); free(h->mgop);
  free(h->cfL); free(h->resCopier); free(h->revCopier); free(h->ratio); free(h->mgBottomCells); free(h);
}
OHier* oracle_create(const int* is, const double* rs) { return oracle_create_r(is, rs, NULL); }
/* VCAMRPoissonOp2Factory::define + MGnewOp (VCAMRPoissonOp2.cpp:868-1070) on a
   one-level hierarchy: the given coefficients on level 0, and on MG depth d the
   coefficients averaged from level 0 with ratio 2^d (CoarseAverage, CoarseAverageFace,
   arithmetic), each op with its lambda. */
static const Op* mg_op(const OHier* h, int d);
int oracle_vc_set(OHier* h, const double* a, const double* b0, const double* b1, const double* b2) {
  if (h->nc > 1) { set_err("variable coefficients with ncomp > 1 are not supported"); return -1; }
  mg_levels_free(h);
  const double* bs[3] = {b0, b1, b2};
  long ao = 0, bo[3] = {0, 0, 0};
  if (h->op[0].relaxMode != 1 && h->op[0].relaxMode != 4) {
    set_err("only relax modes 1 and 4 are implemented for variable coefficients");
    return -1;
  }
  /* every AMR level: its own coefficients (VCAMRPoissonOp2Factory::AMRnewOp, :1074-1171) */
  for (int l = 0; l < h->nlev; l++) {
    Op* op = &h->op[l];
    if (!op->vc) op_vc_alloc(op);
    ld_from_packed(&op->va, a + ao);
    ao += lay_cells(&op->L);
    for (int d = 0; d < 3; d++)
      for (int q = 0; q < op->L.n; q++) {
        Fab* f = &op->vb[d][q];
        long n = b_npts(&f->b);
        memcpy(f->d, bs[d] + bo[d], sizeof(double)*n); bo[d] += n;
      }
    op_vc_lambda(op);
  }
  Op* op0 = &h->op[0];
  for (int dd = 1; dd < h->mgD; dd++) {
    Op* op = &h->mgop[dd];
    if (!op->vc) op_vc_alloc(op);
    const int r = 1 << dd;
    for (int q = 0; q < op->L.n; q++) {
      k_vcavg_cells(&op->va.f[q], &op0->va.f[q], &op->L.b[q], r, h->vcHarmonic);
      for (int d = 0; d < 3; d++) {
        Box cfb = vc_facebox(op->L.b[q], d);
        k_vcavg_faces(&op->vb[d][q], &op0->vb[d][q], &cfb, d, r, h->vcHarmonic);
      }
    }
    op_vc_lambda(op);
  }
  return 0;
}
/* m_coefficient_average_type of VCAMRPoissonOp2Factory (VCAMRPoissonOp2.cpp:1048-1063):
   0 CoarseAverage::arithmetic, 1 harmonic; used by the next oracle_vc_set */
/* DiriBC / NeumBC with a BCValueHolder on every operator (fn NULL: the face constants) */
int oracle_set_bc_values(OHier* h, BCValueFn fn, void* user) {
  mg_levels_free(h);
  h->bc.fn = fn; h->bc.user = user;
  for (int l = 0; l < h->nlev; l++) { h->op[l].bc.fn = fn; h->op[l].bc.user = user; }
  for (int d = 1; d < h->mgD; d++) { h->mgop[d].bc.fn = fn; h->mgop[d].bc.user = user; }
  return 0;
}
int oracle_set_vc_average(OHier* h, int type) {
  if (type != 0 && type != 1) { set_err("coefficient average type must be 0 (arithmetic) or 1 (harmonic)"); return -1; }
  h->vcHarmonic = type;
  return 0;
}
/* AMRPoissonOp::s_relaxMode (AMRPoissonOp.cpp:31, 600-632) for every operator */
int oracle_set_relax_mode(OHier* h, int mode) {
  mg_levels_free(h);
  if (mode < 0 || mode > 5) {
    set_err("relax mode must be 0 (looseGSRB), 1 (levelGSRB), 2 (overlapGSRB), 3 (levelGSRBLazy), "
            "4 (levelJacobi) or 5 (levelMultiColor)");
    return -1;
  }
  /* VCAMRPoissonOp2 aborts in looseGSRB, overlapGSRB, levelGSRBLazy and levelMultiColor
     (VCAMRPoissonOp2.cpp:628-640, 748-760) */
  if (mode != 1 && mode != 4 && h->op[0].vc) {
    set_err("only relax modes 1 and 4 are implemented for variable coefficients");
    return -1;
  }
  for (int l = 0; l < h->nlev; l++) h->op[l].relaxMode = mode;
  for (int d = 1; d < h->mgD; d++) h->mgop[d].relaxMode = mode;
  return 0;
}
/* setAlphaAndBeta (AMRPoissonOp.cpp:1103-1108, VCAMRPoissonOp2.cpp:258-266) on every
   operator of the hierarchy (AMRMultiGrid::getAllOperators), lambda recomputed */
int oracle_set_alpha_beta(OHier* h, double alpha, double beta) {
  mg_levels_free(h);
  h->alpha = alpha; h->beta = beta;
  for (int l = 0; l < h->nlev; l++) {
    h->op[l].alpha = alpha; h->op[l].beta = beta;
    if (h->op[l].vc) op_vc_lambda(&h->op[l]);
  }
  for (int d = 1; d < h->mgD; d++) {
    h->mgop[d].alpha = alpha; h->mgop[d].beta = beta;
    if (h->mgop[d].vc) op_vc_lambda(&h->mgop[d]);
  }
  return 0;
}
long oracle_face_cells(const OHier* h, int lev, int dir) {
  long n = 0;
  for (int q = 0; q < h->grids[lev].n; q++) { Box f = vc_facebox(h->grids[lev].b[q], dir); n += b_npts(&f); }
  return n;
}
/* the lambda of MG depth d, valid-packed (a check of the coefficient coarsening) */
int oracle_vc_lambda(OHier* h, int depth, double* out) {
  const Op* op = mg_op(h, depth);
  if (!op->vc) { set_err("no variable coefficients"); return -1; }
  ld_to_packed(&op->vlam, out);
  return 0;
}
long oracle_level_cells(const OHier* h, int lev) { return lay_cells(&h->grids[lev]); }
/* norm(BoxLayoutData, interval, p) on MG depth d of level 0 (p = 0, 1, 2), over the
   solver's components (x valid-packed with them) */
int oracle_level_norm(OHier* h, int d, const double* x_in, int p, double* out) {
  if (d < 0 || d >= h->mgD) { set_err("bad depth"); return -1; }
  if (p < 0 || p > 2) { set_err("norm order must be 0, 1 or 2"); return -1; }
  const Op* op = mg_op(h, d);
  LD x = ld_make_nc(&op->L, 0, h->nc);
  ld_from_packed(&x, x_in);
  *out = ldo_norm(&x, p);
  ld_free(&x);
  return 0;
}
/* LevelDataOps::dotProduct on MG depth d of level 0, over the solver's components */
int oracle_level_dot(OHier* h, int d, const double* a_in, const double* b_in, double* out) {
  if (d < 0 || d >= h->mgD) { set_err("bad depth"); return -1; }
  const Op* op = mg_op(h, d);
  LD a = ld_make_nc(&op->L, 0, h->nc), b = ld_make_nc(&op->L, 0, h->nc);
  ld_from_packed(&a, a_in);
  ld_from_packed(&b, b_in);
  *out = ldo_dot(&a, &b);
  ld_free(&a); ld_free(&b);
  return 0;
}
/* number of components of the solver's fields (AMRPoissonOp acts on LevelData with any
   nComp; every component shares the operator and its boundary conditions) */
int oracle_set_ncomp(OHier* h, int nc) {
  if (nc < 1 || nc > ORACLE_MAX_NCOMP) { set_err("ncomp must be 1..16"); return -1; }
  if (nc > 1 && h->op[0].vc) { set_err("ncomp > 1 with variable coefficients is not supported"); return -1; }
  h->nc = nc;
  return 0;
}
int oracle_num_mg_depths(const OHier* h) { return h->mgD; }
long oracle_mg_cells(const OHier* h, int d) { return lay_cells(&h->mgL[d]); }
long oracle_ghosted_cells(const OHier* h, int lev) {
  long n = 0; for (int q = 0; q < h->grids[lev].n; q++) { Box g = b_grow(h->grids[lev].b[q], 1); n += b_npts(&g); }
  return n;
}
static const Op* mg_op(const OHier* h, int d) { return d == 0 ? &h->op[0] : &h->mgop[d]; }

/* ---- BiCGStabSolver<T>::solve (BiCGStabSolver.H:199-519) ---- */
typedef struct { double eps, reps, hang, small, metric; int imax, numRestarts, normType; } BiCG;

static int bicgstab_solve(const Op* op, LD* phi, const LD* rhs, const BiCG* p, int* iters_out) {
  const int nc = rhs->nc;
  LD r = ld_make_nc(rhs->L, rhs->g, nc), r_tilde = ld_make_nc(rhs->L, rhs->g, nc), e = ld_make_nc(phi->L, phi->g, nc),
     pp = ld_make_nc(rhs->L, rhs->g, nc), p_tilde = ld_make_nc(phi->L, phi->g, nc),
     s_tilde = ld_make_nc(phi->L, phi->g, nc), t = ld_make_nc(rhs->L, rhs->g, nc), v = ld_make_nc(rhs->L, rhs->g, nc);
  int recount = 0, status = -1;
  ldo_setzero(&r);
  op_residual(op, &r, phi, rhs, 1);
  ldo_assign_local(&r_tilde, &r);
  ldo_setzero(&e); ldo_setzero(&p_tilde); ldo_setzero(&s_tilde);
  int iter = 0;
  double rho[4] = {0,0,0,0}, norm[2];
  norm[0] = ldo_norm(&r, p->normType);
  double initial_norm = norm[0], initial_residual = norm[0];
  norm[1] = norm[0];
  double alpha[2] = {0,0}, beta[2] = {0,0}, omega[2] = {0,0};
  int init = 1, restarts = 0;
  if (p->metric > 0) initial_norm = p->metric;
  while ((iter < p->imax && norm[0] > p->eps*norm[1]) && (norm[1] > 0)) {
    iter++;
    norm[1] = norm[0]; alpha[1] = alpha[0]; beta[1] = beta[0]; omega[1] = omega[0];
    rho[3] = rho[2]; rho[2] = rho[1];
    rho[1] = ldo_dot(&r_tilde, &r);
    if (rho[1] == 0.0) { ldo_incr(phi, &e, 1.0); status = 2; goto done; }
    if (init) { ldo_assign_local(&pp, &r); init = 0; }
    else {
      beta[1] = (rho[1]/rho[2])*(alpha[1]/omega[1]);
      ldo_scale(&pp, beta[1]);
      ldo_incr(&pp, &v, -beta[1]*omega[1]);
      ldo_incr(&pp, &r, 1.0);
    }
    op_precond(op, &p_tilde, &pp);
    ldo_setzero(&v);
    op_apply(op, &v, &p_tilde, 1);
    double m = ldo_dot(&r_tilde, &v);
    alpha[0] = rho[1]/m;
    if (fabs(m) > p->small*fabs(rho[1])) {
      ldo_incr(&r, &v, -alpha[0]);
      norm[0] = ldo_norm(&r, p->normType);
      ldo_incr(&e, &p_tilde, alpha[0]);
    } else { ldo_setzero(&r); norm[0] = 0.0; }
    if (norm[0] > p->eps*initial_norm && norm[0] > p->reps*initial_residual) {
      op_precond(op, &s_tilde, &r);
      ldo_setzero(&t);
      op_apply(op, &t, &s_tilde, 1);
      omega[0] = ldo_dot(&t, &r)/ldo_dot(&t, &t);
      ldo_incr(&e, &s_tilde, omega[0]);
      ldo_incr(&r, &t, -omega[0]);
      norm[0] = ldo_norm(&r, p->normType);
    }
    if (norm[0] <= p->eps*initial_norm || norm[0] <= p->reps*initial_residual) { status = 1; break; }
    if (omega[0] == 0.0 || norm[0] > (1-p->hang)*norm[1]) {
      if (recount == 0) recount = 1;
      else {
        recount = 0;
        ldo_incr(phi, &e, 1.0);
        if (restarts == p->numRestarts) { status = 3; goto done; }
        op_residual(op, &r, phi, rhs, 1);
        norm[0] = ldo_norm(&r, p->normType);
        rho[1] = 0.0; rho[2] = 0.0; rho[3] = 0.0;
        alpha[0] = 0; beta[0] = 0; omega[0] = 0;
        ldo_assign_local(&r_tilde, &r);
        ldo_setzero(&e);
        restarts++;
        init = 1;
      }
    }
  }
  ldo_incr(phi, &e, 1.0);
done:
  ld_free(&r); ld_free(&r_tilde); ld_free(&e); ld_free(&pp); ld_free(&p_tilde);
  ld_free(&s_tilde); ld_free(&t); ld_free(&v);
  if (iters_out) *iters_out = iter;
  return status;
}

/* ---- MultiGrid<T>::cycle (MultiGrid.H:630-696), V/W cycle ---- */
typedef struct {
  OHier* h; int pre, post, bottom, cycles;
  int lev, D;                     /* the AMR level whose MultiGrid this is, its depth count */
  const Layout* L;                /* layouts per depth (L[0] = the level's grids) */
  Op* ops;                        /* ops per depth >= 1 (depth 0 is the AMR level's op) */
  long bottomCells;
  LD* mgres; LD* mgcor;           /* per depth >= 1 */
  BiCG bic;
} MGState;

static const Op* mgs_op(const MGState* s, int depth) {
  return depth == 0 ? &s->h->op[s->lev] : &s->ops[depth];
}
static void mg_cycle(MGState* s, int depth, LD* corr, const LD* res) {
  OHier* h = s->h;
  const Op* op = mgs_op(s, depth);
  if (depth == s->D-1) {
    if (s->bottomCells == 1) op_relax(op, corr, res, 1);
    else {
      op_relax(op, corr, res, s->bottom);
      bicgstab_solve(op, corr, res, &s->bic, NULL);
      h->bottom_calls++;
    }
    return;
  }
  if (s->cycles < 0) {  /* F cycle (MultiGrid.H:655-681) */
    int cycles = -s->cycles;
    op_restrict(op, &s->mgres[depth+1], corr, res);
    ldo_setzero(&s->mgcor[depth+1]);
    mg_cycle(s, depth+1, &s->mgcor[depth+1], &s->mgres[depth+1]);
    op_prolong(op, corr, &s->mgcor[depth+1]);
    op_relax(op, corr, res, s->pre);
    for (int img = 0; img < cycles; img++) {
      op_restrict(op, &s->mgres[depth+1], corr, res);
      ldo_setzero(&s->mgcor[depth+1]);
      s->cycles = 1;                 /* "hack to get a V-cycle" */
      mg_cycle(s, depth+1, &s->mgcor[depth+1], &s->mgres[depth+1]);
      s->cycles = -cycles;
      op_prolong(op, corr, &s->mgcor[depth+1]);
    }
    op_relax(op, corr, res, s->post);
    return;
  }
  op_relax(op, corr, res, s->pre);
  op_restrict(op, &s->mgres[depth+1], corr, res);
  ldo_setzero(&s->mgcor[depth+1]);
  for (int c = 0; c < s->cycles; c++) mg_cycle(s, depth+1, &s->mgcor[depth+1], &s->mgres[depth+1]);
  op_prolong(op, corr, &s->mgcor[depth+1]);
  op_relax(op, corr, res, s->post);
}

/* ---- AMRMultiGrid ---- */
typedef struct {
  OHier* h; MGState mg;
  LD *corr, *resid, *resC;        /* m_correction, m_residual, m_resC */
} AMRState;

/* computeAMRResidualLevel (AMRMultiGrid.H:837-892) */
static void amr_residual_level(AMRState* a, LD* res, LD* phi, const LD* rhs, int lmax, int lbase,
                               int ilev, int homog) {
  OHier* h = a->h;
  if (lmax != lbase) {
    if (ilev == lmax) {
      Op* op = &h->op[lmax];               /* AMRResidualNF */
      qcfi_interp(&op->qcfi, &phi[lmax], &phi[lmax-1]);
      op_residualI(op, &res[lmax], &phi[lmax], &rhs[lmax], homog);
    } else {
      Op* op = &h->op[ilev];                /* AMRResidual / AMRResidualNC */
      if (ilev > 0) qcfi_interp(&op->qcfi, &phi[ilev], &phi[ilev-1]);
      op_applyI(op, &res[ilev], &phi[ilev], homog);
      op_reflux(op, &h->op[ilev+1], &phi[ilev+1], &phi[ilev], &res[ilev]);
      ldo_axby(&res[ilev], &res[ilev], &rhs[ilev], -1.0, 1.0);
    }
  } else if (lbase == 0) {
    op_residual(&h->op[0], &res[0], &phi[0], &rhs[0], homog);
  } else {                                  /* one level above the base: AMRResidualNF */
    Op* op = &h->op[lmax];
    qcfi_interp(&op->qcfi, &phi[lmax], &phi[lmax-1]);
    op_residualI(op, &res[lmax], &phi[lmax], &rhs[lmax], homog);
  }
}
/* zeroCovered -> LevelDataOps::copyToZero over the copier's toRegions */
static void zero_covered(OHier* h, LD* res, int lev) {
  if (res->nc > 1) { for (int c = 0; c < res->nc; c++) zero_covered(h, COMP(res, c), lev); return; }
  const Plan* p = &h->resCopier[lev+1];
  for (int q = 0; q < p->n; q++) fab_setval(&res->f[p->m[q].to], p->m[q].region, 0.0);
}
/* computeAMRResidual (AMRMultiGrid.H:683-736) */
static double amr_residual(AMRState* a, LD* res, LD* phi, const LD* rhs, int lmax, int lbase, int homog) {
  double rnorm = 0;
  for (int ilev = lbase; ilev <= lmax; ilev++) {
    amr_residual_level(a, res, phi, rhs, lmax, lbase, ilev, homog);
    double local;
    if (ilev == lmax) local = op_localMaxNorm(&res[ilev]);
    else { zero_covered(a->h, &res[ilev], ilev); local = op_localMaxNorm(&res[ilev]); }
    rnorm = local > rnorm ? local : rnorm;
  }
  return rnorm;
}
/* AMRRestrictS (AMRPoissonOp.cpp:877-919) */
static void amr_restrictS(AMRState* a, int l, LD* resC, const LD* residual, LD* correction,
                          const LD* coarseCorrection, LD* scratch) {
  if (resC->nc > 1) {
    for (int c = 0; c < resC->nc; c++)
      amr_restrictS(a, l, COMP(resC, c), COMP(residual, c), COMP(correction, c), COMP(coarseCorrection, c),
                    COMP(scratch, c));
    return;
  }
  OHier* h = a->h; Op* op = &h->op[l];
  qcfi_interp(&op->qcfi, correction, coarseCorrection);      /* AMRResidualNF */
  op_residualI(op, scratch, correction, residual, 1);
  for (int b = 0; b < op->L.n; b++) k_average(&resC->f[b], &scratch->f[b], &h->cfL[l].b[b], h->ratio[l]);
}
/* AMRProlongS (AMRPoissonOp.cpp:959-990) */
static void amr_prolongS(AMRState* a, int l, LD* correction, const LD* coarseCorrection, LD* temp) {
  if (correction->nc > 1) {
    for (int c = 0; c < correction->nc; c++)
      amr_prolongS(a, l, COMP(correction, c), COMP(coarseCorrection, c), COMP(temp, c));
    return;
  }
  OHier* h = a->h; Op* op = &h->op[l];
  exec_copy(temp, coarseCorrection, &h->revCopier[l]);
  for (int b = 0; b < op->L.n; b++) k_prolong(&correction->f[b], &temp->f[b], &op->L.b[b], h->ratio[l]);
}
/* AMRMultiGrid::relax (AMRMultiGrid.H:633-666): with ratio 4 to the coarser level,
   m_mg[l] runs MultiGrid::cycle with m_depth 1 (set by init, :1224-1236), which is
   its bottom branch (MultiGrid.H:634-646): one relaxation when m_bottomCells is 1,
   else m_bottom relaxations (the bottom solver is the no-op one) */
static void amr_relax(AMRState* a, int l, LD* corr, const LD* res, int n) {
  OHier* h = a->h;
  if (h->ratio[l] > 2) n = (h->mgBottomCells[l] == 1) ? 1 : a->mg.bottom;
  op_relax(&h->op[l], corr, res, n);
}
/* AMRVCycle (AMRMultiGrid.H:1342-1419) */
static void amr_vcycle(AMRState* a, LD* uberCorr, LD* uberRes, int ilev, int lmax, int lbase) {
  OHier* h = a->h;
  if (ilev == lmax)
    for (int level = lbase; level <= lmax; level++) {
      ldo_assign_local(&a->resid[level], &uberRes[level]);
      ldo_setzero(&a->corr[level]);
    }
  if (lmax == lbase) { mg_cycle(&a->mg, 0, &uberCorr[ilev], &uberRes[ilev]); return; }
  if (ilev == lbase) {
    mg_cycle(&a->mg, 0, &a->corr[ilev], &a->resid[ilev]);
    ldo_incr(&uberCorr[ilev], &a->corr[ilev], 1.0);
    return;
  }
  Op* op = &h->op[ilev];
  amr_relax(a, ilev, &a->corr[ilev], &a->resid[ilev], a->mg.pre);
  ldo_incr(&uberCorr[ilev], &a->corr[ilev], 1.0);
  ldo_setzero(&a->corr[ilev-1]);
  amr_residual_level(a, a->resid, uberCorr, uberRes, lmax, lbase, ilev-1, 1);
  amr_restrictS(a, ilev, &a->resC[ilev], &a->resid[ilev], &a->corr[ilev], &a->corr[ilev-1], &uberCorr[ilev]);
  exec_copy(&a->resid[ilev-1], &a->resC[ilev], &h->resCopier[ilev]);
  for (int img = 0; img < a->mg.cycles; img++) amr_vcycle(a, uberCorr, uberRes, ilev-1, lmax, lbase);
  amr_prolongS(a, ilev, &a->corr[ilev], &a->corr[ilev-1], &a->resC[ilev]);
  /* AMRUpdateResidual: AMRResidualNF(res, corr, corrC, res, true) in place */
  qcfi_interp(&op->qcfi, &a->corr[ilev], &a->corr[ilev-1]);
  op_residualI(op, &a->resid[ilev], &a->corr[ilev], &a->resid[ilev], 1);
  LD* dCorr = &uberCorr[ilev];
  ldo_setzero(dCorr);
  amr_relax(a, ilev, dCorr, &a->resid[ilev], a->mg.post);
  ldo_incr(&a->corr[ilev], dCorr, 1.0);
  ldo_assign_local(&uberCorr[ilev], &a->corr[ilev]);
}

/* MultiGrid::define on AMR level l > 0 via MGnewOp (AMRPoissonOp.cpp:1923-1989):
   coarsenings of the level's grids while coarsenable by coarsening*2, each op with
   the level's coarser dx as dxCrse (its coarse-fine faces keep homogeneous
   interpolation), and m_bottomCells from the level's domain */
static void mg_level_define(OHier* h, int l) {
  if (h->lmgD[l]) return;
  int D = 1;
  while (lay_coarsenable(&h->grids[l], (1 << D)*2)) D++;
  h->lmgD[l] = D;
  h->lmgL[l] = calloc(D, sizeof(Layout));
  h->lmgop[l] = calloc(D, sizeof(Op));
  h->lmgL[l][0] = h->grids[l];
  for (int d = 1; d < D; d++) {
    h->lmgL[l][d] = lay_coarsen(&h->grids[l], 1 << d);
    Box dom = b_coarsen(h->domains[l], 1 << d);
    op_define(&h->lmgop[l][d], &h->lmgL[l][d], &dom, h->dx[l]*(1 << d), h->alpha, h->beta, &h->bc);
    h->lmgop[l][d].dxCrse = h->dx[l-1];
    h->lmgop[l][d].relaxMode = h->op[l].relaxMode;
  }
}
static void amr_state_make(AMRState* a, OHier* h, int pre, int post, int bottom, int numMG, int lbase) {
  a->h = h;
  a->corr = calloc(h->nlev, sizeof(LD)); a->resid = calloc(h->nlev, sizeof(LD)); a->resC = calloc(h->nlev, sizeof(LD));
  for (int l = 0; l < h->nlev; l++) {
    a->corr[l] = ld_make_nc(&h->grids[l], 1, h->nc);
    a->resid[l] = ld_make_nc(&h->grids[l], 0, h->nc);
    if (l > 0) a->resC[l] = ld_make_nc(&h->cfL[l], 0, h->nc);
  }
  MGState* m = &a->mg;
  m->h = h; m->pre = pre; m->post = post; m->bottom = bottom; m->cycles = numMG;
  m->lev = lbase;
  if (lbase == 0) {
    m->D = h->mgD; m->L = h->mgL; m->ops = h->mgop; m->bottomCells = h->bottomCells;
  } else {
    mg_level_define(h, lbase);
    m->D = h->lmgD[lbase]; m->L = h->lmgL[lbase]; m->ops = h->lmgop[lbase];
    m->bottomCells = h->mgBottomCells[lbase];
  }
  m->mgres = calloc(m->D, sizeof(LD)); m->mgcor = calloc(m->D, sizeof(LD));
  for (int d = 1; d < m->D; d++) {
    m->mgres[d] = ld_make_nc(&m->L[d], 0, h->nc);
    m->mgcor[d] = ld_make_nc(&m->L[d], 1, h->nc);
  }
  m->bic.eps = 1e-6; m->bic.reps = h->bic_reps; m->bic.hang = h->bic_hang; m->bic.small = h->bic_small;
  m->bic.metric = -1.0; m->bic.imax = h->bic_imax; m->bic.numRestarts = h->bic_numRestarts; m->bic.normType = 2;
}
static void amr_state_free(AMRState* a) {
  OHier* h = a->h;
  for (int l = 0; l < h->nlev; l++) { ld_free(&a->corr[l]); ld_free(&a->resid[l]); if (l > 0) ld_free(&a->resC[l]); }
  for (int d = 1; d < a->mg.D; d++) { ld_free(&a->mg.mgres[d]); ld_free(&a->mg.mgcor[d]); }
  free(a->corr); free(a->resid); free(a->resC); free(a->mg.mgres); free(a->mg.mgcor);
}

static LD* levels_make(OHier* h, int g) {
  LD* x = calloc(h->nlev, sizeof(LD));
  for (int l = 0; l < h->nlev; l++) x[l] = ld_make_nc(&h->grids[l], g, h->nc);
  return x;
}
static void levels_free(OHier* h, LD* x) { for (int l = 0; l < h->nlev; l++) ld_free(&x[l]); free(x); }
static void levels_from_packed(OHier* h, LD* x, const double* v) {
  long o = 0; for (int l = 0; l < h->nlev; l++) { ld_from_packed(&x[l], v + o); o += lay_cells(&h->grids[l])*h->nc; }
}
static void levels_to_packed(OHier* h, const LD* x, double* v) {
  long o = 0; for (int l = 0; l < h->nlev; l++) { ld_to_packed(&x[l], v + o); o += lay_cells(&h->grids[l])*h->nc; }
}

/* AMRMultiGrid::solve(phi, rhs, l_max, l_base) -> solveNoInit -> solveNoInitResid
   (AMRMultiGrid.H:895-1124).  phi and rhs are valid-packed over all levels; only
   levels l_base..l_max are solved for, phi of level l_base-1 gives the coarse-fine
   boundary values of level l_base, and the other levels are returned unchanged. */
int oracle_solve_levels(OHier* h, const int* sp, const double* rp, const double* rhs_in, double* phi_io,
                        int zero_phi, int lmax, int lbase, double* hist, int hist_cap, int* info) {
  int pre = sp[0], post = sp[1], bottom = sp[2], numMG = sp[3], iterMax = sp[4], imin = sp[5];
  double eps = rp[0], hang = rp[1], normThresh = rp[2];
  if (lbase < 0 || lbase > lmax || lmax >= h->nlev) { set_err("bad l_base / l_max"); return -1; }
  if (lbase > 0 && h->op[lbase].vc) { set_err("l_base > 0 with variable coefficients"); return -1; }
  AMRState a; amr_state_make(&a, h, pre, post, bottom, numMG, lbase);
  LD* phi = levels_make(h, 1);
  LD* rhs = levels_make(h, 0);
  LD* uberRes = levels_make(h, 0);
  LD* uberCorr = levels_make(h, 1);
  levels_from_packed(h, rhs, rhs_in);
  levels_from_packed(h, phi, phi_io);
  if (zero_phi) for (int l = lbase; l <= lmax; l++) ldo_setzero(&phi[l]);
  h->bottom_calls = 0;
  double initial = amr_residual(&a, uberRes, phi, rhs, lmax, lbase, 0);
  double rnorm = initial, norm_last = 2*initial;
  a.mg.bic.metric = initial; a.mg.bic.eps = 1.0*eps;     /* setConvergenceMetrics */
  int iter = 0;
  if (hist && hist_cap > 0) hist[0] = rnorm;
  int goNorm = rnorm > normThresh, goRedu = rnorm > eps*initial, goIter = iter < iterMax,
      goHang = iter < imin || rnorm < (1-hang)*norm_last;
  while (goIter && goRedu && goHang && goNorm) {
    norm_last = rnorm;
    amr_vcycle(&a, uberCorr, uberRes, lmax, lmax, lbase);
    for (int l = lbase; l <= lmax; l++) { ldo_incr(&phi[l], &uberCorr[l], 1.0); ldo_setzero(&uberCorr[l]); }
    rnorm = amr_residual(&a, uberRes, phi, rhs, lmax, lbase, 0);
    iter++;
    if (hist && iter < hist_cap) hist[iter] = rnorm;
    goNorm = rnorm > normThresh; goRedu = rnorm > eps*initial; goIter = iter < iterMax;
    goHang = iter < imin || rnorm < (1-hang)*norm_last;
  }
  int exitStatus = (!goRedu) + (!goIter)*2 + (!goHang)*4 + (!goNorm)*8;
  levels_to_packed(h, phi, phi_io);
  if (info) { info[0] = iter; info[1] = exitStatus; info[2] = h->bottom_calls; }
  levels_free(h, phi); levels_free(h, rhs); levels_free(h, uberRes); levels_free(h, uberCorr);
  amr_state_free(&a);
  return 0;
}
int oracle_solve(OHier* h, const int* sp, const double* rp, const double* rhs_in, double* phi_io,
                 int zero_phi, double* hist, int hist_cap, int* info) {
  return oracle_solve_levels(h, sp, rp, rhs_in, phi_io, zero_phi, h->nlev-1, 0, hist, hist_cap, info);
}

double oracle_amr_residual(OHier* h, const double* phi_in, const double* rhs_in, double* res_out, int homog) {
  AMRState a; amr_state_make(&a, h, 2, 2, 2, 1, 0);
  LD* phi = levels_make(h, 1); LD* rhs = levels_make(h, 0); LD* res = levels_make(h, 0);
  levels_from_packed(h, phi, phi_in); levels_from_packed(h, rhs, rhs_in);
  double n = amr_residual(&a, res, phi, rhs, h->nlev-1, 0, homog);
  levels_to_packed(h, res, res_out);
  levels_free(h, phi); levels_free(h, rhs); levels_free(h, res); amr_state_free(&a);
  return n;
}

/* computeAMROperator (AMRMultiGrid.H:790-833) */
int oracle_amr_operator(OHier* h, const double* phi_in, double* out, int homog) {
  LD* phi = levels_make(h, 1); LD* lph = levels_make(h, 0);
  levels_from_packed(h, phi, phi_in);
  int lmax = h->nlev-1, lbase = 0;
  for (int ilev = lbase; ilev <= lmax; ilev++) {
    if (lmax != lbase) {
      Op* op = &h->op[ilev];
      if (ilev == lmax) {
        qcfi_interp(&op->qcfi, &phi[lmax], &phi[lmax-1]);
        op_applyI(op, &lph[lmax], &phi[lmax], homog);
      } else {
        if (ilev > 0) qcfi_interp(&op->qcfi, &phi[ilev], &phi[ilev-1]);
        op_applyI(op, &lph[ilev], &phi[ilev], homog);
        op_reflux(op, &h->op[ilev+1], &phi[ilev+1], &phi[ilev], &lph[ilev]);
      }
    } else op_apply(&h->op[0], &lph[0], &phi[0], homog);
  }
  levels_to_packed(h, lph, out);
  levels_free(h, phi); levels_free(h, lph);
  return 0;
}

/* ---- single-level entry points on the MG levels of AMR level 0 ---- */
int oracle_level_apply(OHier* h, int d, const double* phi_in, double* out, int homog) {
  if (d < 0 || d >= h->mgD) { set_err("bad depth"); return -1; }
  const Op* op = mg_op(h, d);
  LD phi = ld_make(&h->mgL[d], 1), l = ld_make(&h->mgL[d], 0);
  ld_from_packed(&phi, phi_in);
  op_apply(op, &l, &phi, homog);
  ld_to_packed(&l, out);
  ld_free(&phi); ld_free(&l);
  return 0;
}
int oracle_level_residual(OHier* h, int d, const double* phi_in, const double* rhs_in, double* out, int homog) {
  if (d < 0 || d >= h->mgD) { set_err("bad depth"); return -1; }
  const Op* op = mg_op(h, d);
  LD phi = ld_make(&h->mgL[d], 1), rhs = ld_make(&h->mgL[d], 0), r = ld_make(&h->mgL[d], 0);
  ld_from_packed(&phi, phi_in); ld_from_packed(&rhs, rhs_in);
  op_residual(op, &r, &phi, &rhs, homog);
  ld_to_packed(&r, out);
  ld_free(&phi); ld_free(&rhs); ld_free(&r);
  return 0;
}
int oracle_level_relax(OHier* h, int d, double* phi_io, const double* rhs_in, int iters) {
  if (d < 0 || d >= h->mgD) { set_err("bad depth"); return -1; }
  const Op* op = mg_op(h, d);
  LD phi = ld_make(&h->mgL[d], 1), rhs = ld_make(&h->mgL[d], 0);
  ld_from_packed(&phi, phi_io); ld_from_packed(&rhs, rhs_in);
  op_relax(op, &phi, &rhs, iters);
  ld_to_packed(&phi, phi_io);
  ld_free(&phi); ld_free(&rhs);
  return 0;
}
int oracle_level_restrict(OHier* h, int d, const double* phi_in, const double* rhs_in, double* resc) {
  if (d < 0 || d+1 >= h->mgD) { set_err("bad depth"); return -1; }
  const Op* op = mg_op(h, d);
  LD phi = ld_make(&h->mgL[d], 1), rhs = ld_make(&h->mgL[d], 0), rc = ld_make(&h->mgL[d+1], 0);
  ld_from_packed(&phi, phi_in); ld_from_packed(&rhs, rhs_in);
  op_restrict(op, &rc, &phi, &rhs);
  ld_to_packed(&rc, resc);
  ld_free(&phi); ld_free(&rhs); ld_free(&rc);
  return 0;
}
int oracle_level_prolong(OHier* h, int d, double* phi_io, const double* ec) {
  if (d < 0 || d+1 >= h->mgD) { set_err("bad depth"); return -1; }
  const Op* op = mg_op(h, d);
  LD phi = ld_make(&h->mgL[d], 1), c = ld_make(&h->mgL[d+1], 1);
  ld_from_packed(&phi, phi_io); ld_from_packed(&c, ec);
  op_prolong(op, &phi, &c);
  ld_to_packed(&phi, phi_io);
  ld_free(&phi); ld_free(&c);
  return 0;
}
int oracle_kernel_gsrb(OHier* h, int lev, double* phi_g, const double* rhs_in, int color) {
  const Op* op = &h->op[lev];
  LD phi = ld_make(&h->grids[lev], 1), rhs = ld_make(&h->grids[lev], 0);
  ld_from_ghosted(&phi, phi_g); ld_from_packed(&rhs, rhs_in);
  for (int b = 0; b < op->L.n; b++) {
    if (op->alpha == 0.0 && op->beta == 1.0) k_gsrb_lap(&phi.f[b], &rhs.f[b], &op->L.b[b], op->dx, color);
    else k_gsrb_helm(&phi.f[b], &rhs.f[b], &op->L.b[b], op->dx, op->alpha, op->beta, color);
  }
  ld_to_ghosted(&phi, phi_g);
  ld_free(&phi); ld_free(&rhs);
  return 0;
}
int oracle_homog_cf(OHier* h, int lev, double* phi_g) {
  LD phi = ld_make(&h->grids[lev], 1);
  ld_from_ghosted(&phi, phi_g);
  op_homog_cf(&h->op[lev], &phi);
  ld_to_ghosted(&phi, phi_g);
  ld_free(&phi);
  return 0;
}
int oracle_quad_cf(OHier* h, int lev, double* phi_g, const double* coarse) {
  if (lev < 1 || lev >= h->nlev) { set_err("bad level"); return -1; }
  LD phi = ld_make(&h->grids[lev], 1), c = ld_make(&h->grids[lev-1], 1);
  ld_from_ghosted(&phi, phi_g); ld_from_packed(&c, coarse);
  qcfi_interp(&h->op[lev].qcfi, &phi, &c);
  ld_to_ghosted(&phi, phi_g);
  ld_free(&phi); ld_free(&c);
  return 0;
}
int oracle_average(OHier* h, int lev, const double* fine, double* out) {
  if (lev < 1 || lev >= h->nlev) { set_err("bad level"); return -1; }
  LD f = ld_make(&h->grids[lev], 0), c = ld_make(&h->cfL[lev], 0);
  ld_from_packed(&f, fine);
  for (int b = 0; b < h->grids[lev].n; b++) k_average(&c.f[b], &f.f[b], &h->cfL[lev].b[b], h->ratio[lev]);
  ld_to_packed(&c, out);
  ld_free(&f); ld_free(&c);
  return 0;
}
/* AMRProlongS's PROLONG (AMRPoissonOp.cpp:959-990) on AMR level lev: fine (valid-packed,
   in/out) += the coarsened-fine-layout field coarse (valid-packed over coarsen(grids, ratio)) */
int oracle_amr_prolong(OHier* h, int lev, double* fine_io, const double* coarse) {
  if (lev < 1 || lev >= h->nlev) { set_err("bad level"); return -1; }
  LD f = ld_make(&h->grids[lev], 0), c = ld_make(&h->cfL[lev], 0);
  ld_from_packed(&f, fine_io);
  ld_from_packed(&c, coarse);
  for (int b = 0; b < h->grids[lev].n; b++) k_prolong(&f.f[b], &c.f[b], &h->grids[lev].b[b], h->ratio[lev]);
  ld_to_packed(&f, fine_io);
  ld_free(&f); ld_free(&c);
  return 0;
}
int oracle_bottom_solve(OHier* h, double* corr_io, const double* res_in, int nbottom, double metric,
                        double eps, int* info) {
  int d = h->mgD-1;
  const Op* op = mg_op(h, d);
  LD c = ld_make(&h->mgL[d], 1), r = ld_make(&h->mgL[d], 0);
  ld_from_packed(&c, corr_io); ld_from_packed(&r, res_in);
  BiCG b = {eps, h->bic_reps, h->bic_hang, h->bic_small, metric, h->bic_imax, h->bic_numRestarts, 2};
  op_relax(op, &c, &r, nbottom);
  int it = 0;
  int st = bicgstab_solve(op, &c, &r, &b, &it);
  ld_to_packed(&c, corr_io);
  if (info) { info[0] = st; info[1] = it; }
  ld_free(&c); ld_free(&r);
  return 0;
}

/* BiCGStabSolver's public parameters (BiCGStabSolver.H:105-153) for every later
   bottom solve of h: oracle_bottom_solve and the multigrid bottom of oracle_solve. */
int oracle_set_bottom_params(OHier* h, int imax, int numRestarts, double reps, double hang, double small) {
  if (imax < 0 || numRestarts < 0) { set_err("bad BiCGStab parameters"); return -1; }
  h->bic_imax = imax; h->bic_numRestarts = numRestarts;
  h->bic_reps = reps; h->bic_hang = hang; h->bic_small = small;
  return 0;
}

/* ---- operator applications on AMR level lev with its own CFRegion: the level's
   AMRPoissonOp (homogeneousCFInterp with dxCrse = dx of level lev-1, exchange, BC) ---- */
static int chk_lev(const OHier* h, int lev) {
  if (lev < 0 || lev >= h->nlev) { set_err("bad level"); return -1; }
  return 0;
}
/* AMRPoissonOp::applyOp (AMRPoissonOp.cpp:332-340) */
int oracle_amrlevel_apply(OHier* h, int lev, const double* phi_in, double* out, int homog) {
  if (chk_lev(h, lev)) return -1;
  LD phi = ld_make(&h->grids[lev], 1), l = ld_make(&h->grids[lev], 0);
  ld_from_packed(&phi, phi_in);
  op_apply(&h->op[lev], &l, &phi, homog);
  ld_to_packed(&l, out);
  ld_free(&phi); ld_free(&l);
  return 0;
}
/* AMRPoissonOp::residual (AMRPoissonOp.cpp:224-236) */
int oracle_amrlevel_residual(OHier* h, int lev, const double* phi_in, const double* rhs_in, double* out, int homog) {
  if (chk_lev(h, lev)) return -1;
  LD phi = ld_make(&h->grids[lev], 1), rhs = ld_make(&h->grids[lev], 0), r = ld_make(&h->grids[lev], 0);
  ld_from_packed(&phi, phi_in); ld_from_packed(&rhs, rhs_in);
  op_residual(&h->op[lev], &r, &phi, &rhs, homog);
  ld_to_packed(&r, out);
  ld_free(&phi); ld_free(&rhs); ld_free(&r);
  return 0;
}
/* AMRPoissonOp::relax (AMRPoissonOp.cpp:600-632) */
int oracle_amrlevel_relax(OHier* h, int lev, double* phi_io, const double* rhs_in, int iters) {
  if (chk_lev(h, lev)) return -1;
  LD phi = ld_make(&h->grids[lev], 1), rhs = ld_make(&h->grids[lev], 0);
  ld_from_packed(&phi, phi_io); ld_from_packed(&rhs, rhs_in);
  op_relax(&h->op[lev], &phi, &rhs, iters);
  ld_to_packed(&phi, phi_io);
  ld_free(&phi); ld_free(&rhs);
  return 0;
}

/* Which QuadCFStencil branch (QuadCFStencil.cpp:662-918) each tangential derivative
   of level lev's coarse-fine stencils took, summed over every fine box and face:
   counts[0] standard cells (centred formulas throughout), [1] cells with special
   stencils (m_ivsQuadd), [2] no admissible mixed stencil (order dropped),
   [3] centred, [4] shifted toward +, [5] shifted toward -, [6] dropped order with a
   one-sided first derivative toward +, [7] toward -, [8] no admissible first
   derivative (zero), [9] mixed stencils averaged over 1, [10] 2, [11] 3, [12] 4
   corners.  Branches [3..8] count (cell, tangential direction) pairs. */
int oracle_cf_branch_counts(OHier* h, int lev, long* counts) {
  if (chk_lev(h, lev) || lev == 0) { set_err("bad level"); return -1; }
  for (int q = 0; q < 13; q++) counts[q] = 0;
  const QCFI* qi = &h->op[lev].qcfi;
  if (!qi->level) return 0;
  for (int b = 0; b < h->grids[lev].n; b++)
    for (int f = 0; f < 6; f++) {
      const QCFS* s = &qi->st[b*6+f];
      if (!s->defined_nonempty) continue;
      counts[0] += ivs_numpts(&s->standard);
      if (ivs_numpts(&s->quadd) == 0) continue;
      int vt[2], n = 0;
      for (int d = 0; d < 3; d++) if (d != s->dir) vt[n++] = d;
      for (int k = s->osbox.lo[2]; k <= s->osbox.hi[2]; k++)
        for (int j = s->osbox.lo[1]; j <= s->osbox.hi[1]; j++)
          for (int i = s->osbox.lo[0]; i <= s->osbox.hi[0]; i++) {
            if (!ivs_has(&s->quadd, i, j, k)) continue;
            counts[1]++;
            long o = os_off(s, i, j, k);
            const int nm = s->nmix[o];
            if (nm == 0) { counts[2]++; continue; }
            counts[8+nm]++;
            int iv[3] = {i, j, k};
            for (int t = 0; t < 2; t++) {
              int dd = vt[t];
              const DSten* sec = &s->second[dd][o];
              const DSten* fir = &s->first[dd][o];
              if (sec->n == 3) {
                int lo = 1 << 30, hi = -(1 << 30);
                for (int q = 0; q < 3; q++) { int off = sec->iv[q][dd] - iv[dd]; if (off < lo) lo = off; if (off > hi) hi = off; }
                if (lo == -1 && hi == 1) counts[3]++;
                else if (lo == 0 && hi == 2) counts[4]++;
                else if (lo == -2 && hi == 0) counts[5]++;
              } else {
                int plus = 0, minus = 0;
                for (int q = 0; q < fir->n; q++) {
                  int off = fir->iv[q][dd] - iv[dd];
                  if (off == 1 && fir->w[q] != 0.0) plus = 1;
                  if (off == -1 && fir->w[q] != 0.0) minus = 1;
                }
                if (plus) counts[6]++;
                else if (minus) counts[7]++;
                else counts[8]++;
              }
            }
          }
    }
  return 0;
}

/* OpenMP threads of every later oracle call (the CPU baseline times all cores and one) */
void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
int oracle_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
