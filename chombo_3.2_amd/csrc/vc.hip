// Variable-coefficient kernels (VCAMRPoissonOp2) and their coefficient storage.
//
// Every kernel restates the reference's Fortran expression in its order (the
// library is built without FMA contraction), so a level operation is bitwise the
// oracle's.  Cells are enumerated by the level's tile table as in kernels.hip; the
// tile's box index selects both the cell box and the face box.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

#include "fused.hpp"
#include "ops.hpp"
#include "vc.hpp"

namespace amrmg {

// ---------------------------------------------------------------------------
// storage
// ---------------------------------------------------------------------------
template <class Real>
VCCoefs<Real>::VCCoefs(const DeviceLayout* L) : ctx(L->ctx), lay(L) {
  a.define(L);
  lambda.define(L);
  const int nent = (int)L->h_boxes.size();
  for (int d = 0; d < 3; d++) {
    std::vector<DBox>& fb = h_fbox[d];
    fb.resize(nent);
    long long off = 0;
    auto face_array = [&](const DBox& x) {  // a face array of its own over box x
      DBox f = x;
      const int nx = x.hi[0] - x.lo[0] + 1 + (d == 0), ny = x.hi[1] - x.lo[1] + 1 + (d == 1);
      const int nz = x.hi[2] - x.lo[2] + 1 + (d == 2);
      f.off = off;
      f.sy = nx;
      f.sz = nx * ny;
      off += (long long)nx * ny * nz;
      return f;
    };
    if (L->merged) {
      const DBox r = face_array(L->h_boxes[L->nbox]);
      fb[L->nbox] = r;
      for (int q = 0; q < L->nbox; q++) {
        DBox f = L->h_boxes[q];
        f.sy = r.sy;
        f.sz = r.sz;
        f.off = r.off + (f.lo[0] - r.lo[0]) + (long long)(f.lo[1] - r.lo[1]) * r.sy +
                (long long)(f.lo[2] - r.lo[2]) * r.sz;
        fb[q] = f;
      }
    } else {
      for (int q = 0; q < nent; q++) fb[q] = face_array(L->h_boxes[q]);
    }
    fslab[d] = off;
    dev_alloc(ctx, (void**)&b[d], sizeof(Real) * (size_t)off, "hipMalloc bCoef");
    dev_alloc(ctx, (void**)&d_fbox[d], sizeof(DBox) * fb.size(), "hipMalloc face table");
    dev_h2d(ctx, d_fbox[d], fb.data(), sizeof(DBox) * fb.size(), "copy face table");
  }
}
template <class Real>
VCCoefs<Real>::~VCCoefs() {
  if (ctx && !ctx->dry) (void)hipStreamSynchronize(ctx->stream);
  for (int d = 0; d < 3; d++) {
    dev_free(ctx, b[d]);
    dev_free(ctx, d_fbox[d]);
  }
}
template <class Real>
long long VCCoefs<Real>::packedFaces(int d) const {
  long long n = 0;
  for (int q = 0; q < lay->nbox; q++) {
    const DBox& x = lay->h_boxes[q];
    n += (long long)(x.hi[0] - x.lo[0] + 1 + (d == 0)) * (x.hi[1] - x.lo[1] + 1 + (d == 1)) *
         (x.hi[2] - x.lo[2] + 1 + (d == 2));
  }
  return n;
}
template <class Real>
long long VCCoefs<Real>::bytes() const {
  return (long long)sizeof(Real) * (2 * lay->slab + fslab[0] + fslab[1] + fslab[2]);
}
template <class Real>
void VCCoefs<Real>::setCells(const double* h) {
  std::vector<Real> s((size_t)lay->slab, Real(0));
  long long n = 0;
  for (int q = 0; q < lay->nbox; q++) {
    const DBox& x = lay->h_boxes[q];
    for (int k = x.lo[2]; k <= x.hi[2]; k++)
      for (int j = x.lo[1]; j <= x.hi[1]; j++)
        for (int i = x.lo[0]; i <= x.hi[0]; i++) s[lay->index(q, i, j, k)] = Real(h[n++]);
  }
  if (ctx->dry) return;
  hip_check(hipMemcpy(a.d, s.data(), sizeof(Real) * s.size(), hipMemcpyHostToDevice), "copy aCoef");
}
template <class Real>
void VCCoefs<Real>::setFaces(int d, const double* h) {
  if (d < 0 || d > 2) throw std::runtime_error("bCoef: direction must be 0, 1 or 2");
  std::vector<Real> s((size_t)fslab[d], Real(0));
  std::vector<char> seen((size_t)fslab[d], 0);
  long long n = 0;
  for (int q = 0; q < lay->nbox; q++) {
    const DBox& f = h_fbox[d][q];
    for (int k = f.lo[2]; k <= f.hi[2] + (d == 2); k++)
      for (int j = f.lo[1]; j <= f.hi[1] + (d == 1); j++)
        for (int i = f.lo[0]; i <= f.hi[0] + (d == 0); i++) {
          const long long x = f.off + (i - f.lo[0]) + (long long)(j - f.lo[1]) * f.sy + (long long)(k - f.lo[2]) * f.sz;
          if (x < 0 || x >= fslab[d]) throw std::runtime_error("bCoef: face index outside its array");
          const Real v = Real(h[n++]);
          if (seen[x] && std::memcmp(&s[x], &v, sizeof(Real)) != 0)
            throw std::runtime_error("bCoef: a face shared by two boxes has two values (direction " +
                                     std::to_string(d) + ")");
          s[x] = v;
          seen[x] = 1;
        }
  }
  if (ctx->dry) return;
  hip_check(hipMemcpy(b[d], s.data(), sizeof(Real) * s.size(), hipMemcpyHostToDevice), "copy bCoef");
}
template struct VCCoefs<double>;
template struct VCCoefs<float>;

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ long long cidx(const DBox& b, int i, int j, int k) {
  return b.off + (i - b.lo[0] + 1) + (long long)(j - b.lo[1] + 1) * b.sy + (long long)(k - b.lo[2] + 1) * b.sz;
}
__device__ __forceinline__ long long fidx(const DBox& f, int i, int j, int k) {
  return f.off + (i - f.lo[0]) + (long long)(j - f.lo[1]) * f.sy + (long long)(k - f.lo[2]) * f.sz;
}

struct Geom {
  int lx, ly;
};
Geom geom(const DeviceLayout& L) {
  Geom g{0, 0};
  while ((1 << g.lx) < L.TX) g.lx++;
  while ((1 << g.ly) < L.TY) g.ly++;
  return g;
}
__device__ __forceinline__ bool tile_cell(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles, Geom g,
                                          int& box, DBox& b, int& i, int& j, int& k) {
  const Tile t = tiles[blockIdx.x];
  box = t.box;
  b = boxes[t.box];
  const int tid = threadIdx.x;
  i = t.i0 + (tid & ((1 << g.lx) - 1));
  j = t.j0 + ((tid >> g.lx) & ((1 << g.ly) - 1));
  k = t.k0 + (tid >> (g.lx + g.ly));
  return i <= b.hi[0] && j <= b.hi[1] && k <= b.hi[2];
}

// the coefficient operands of a kernel
template <class T>
struct VCArgs {
  const T* a;
  const T* lam;
  const T* b0;
  const T* b1;
  const T* b2;
  const DBox* f0;
  const DBox* f1;
  const DBox* f2;
};
template <class T>
VCArgs<T> args(const VCCoefs<T>& C) {
  return VCArgs<T>{C.a.d, C.lambda.d, C.b[0], C.b[1], C.b[2], C.d_fbox[0], C.d_fbox[1], C.d_fbox[2]};
}

// the operator at cell c of box `box`: alpha*a*phi - beta*(sum of the six face
// terms in the CHF_DTERM order)*dxinv (VCAMRPoissonOpF.ChF:236-254)
template <class T>
__device__ __forceinline__ T vc_lofphi(const VCArgs<T>& v, const T* __restrict__ phi, const DBox& b, int box, int i,
                                       int j, int k, long long c, T alpha, T beta, T dxinv) {
  const DBox f0 = v.f0[box], f1 = v.f1[box], f2 = v.f2[box];
  const long long x0 = fidx(f0, i, j, k), x1 = fidx(f1, i, j, k), x2 = fidx(f2, i, j, k);
  const long long sy = b.sy, sz = b.sz;
  const T p = phi[c];
  const T s = v.b0[x0 + 1] * (phi[c + 1] - p) - v.b0[x0] * (p - phi[c - 1]) + v.b1[x1 + f1.sy] * (phi[c + sy] - p) -
              v.b1[x1] * (p - phi[c - sy]) + v.b2[x2 + f2.sz] * (phi[c + sz] - p) - v.b2[x2] * (p - phi[c - sz]);
  return alpha * v.a[c] * p - beta * s * dxinv;
}

template <class T>
__global__ void __launch_bounds__(256) k_vc_gsrb(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                                 Geom g, VCArgs<T> v, T* __restrict__ phi, const T* __restrict__ rhs,
                                                 int color, T dxinv, T alpha, T beta) {
  DBox b;
  int box, i, j, k;
  if (!tile_cell(boxes, tiles, g, box, b, i, j, k)) return;
  if ((i + j + k + color) & 1) return;
  const long long c = cidx(b, i, j, k);
  const T lofphi = vc_lofphi(v, phi, b, box, i, j, k, c, alpha, beta, dxinv);
  phi[c] = phi[c] - v.lam[c] * (lofphi - rhs[c]);
}
template <class T>
__global__ void __launch_bounds__(256) k_vc_apply(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                                  Geom g, VCArgs<T> v, T* __restrict__ lofphi,
                                                  const T* __restrict__ phi, T dxinv, T alpha, T beta) {
  DBox b;
  int box, i, j, k;
  if (!tile_cell(boxes, tiles, g, box, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  lofphi[c] = vc_lofphi(v, phi, b, box, i, j, k, c, alpha, beta, dxinv);
}
template <class T>
__global__ void __launch_bounds__(256) k_vc_residual(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                                     Geom g, VCArgs<T> v, T* res, const T* __restrict__ phi,
                                                     const T* rhs, T dxinv, T alpha, T beta) {
  DBox b;
  int box, i, j, k;
  if (!tile_cell(boxes, tiles, g, box, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  const T lofphi = vc_lofphi(v, phi, b, box, i, j, k, c, alpha, beta, dxinv);
  res[c] = rhs[c] - lofphi;
}
// one thread per coarse cell accumulates its eight children in the Fortran loop
// order of the fine region (res zeroed first: the sum starts from zero)
template <class T>
__global__ void __launch_bounds__(256) k_vc_restrict(const DBox* __restrict__ cboxes, const Tile* __restrict__ ctiles,
                                                     Geom g, const DBox* __restrict__ fboxes, VCArgs<T> v,
                                                     T* __restrict__ resc, const T* __restrict__ phi,
                                                     const T* __restrict__ rhs, T dxinv, T alpha, T beta) {
  DBox cb;
  int box, I, J, K;
  if (!tile_cell(cboxes, ctiles, g, box, cb, I, J, K)) return;
  const DBox fb = fboxes[box];
  const T denom = T(2 * 2 * 2);
  T acc = T(0.0);
#pragma unroll
  for (int kk = 0; kk < 2; kk++)
#pragma unroll
    for (int jj = 0; jj < 2; jj++)
#pragma unroll
      for (int ii = 0; ii < 2; ii++) {
        const int i = 2 * I + ii, j = 2 * J + jj, k = 2 * K + kk;
        const long long c = cidx(fb, i, j, k);
        const T lofphi = vc_lofphi(v, phi, fb, box, i, j, k, c, alpha, beta, dxinv);
        acc = acc + (rhs[c] - lofphi) / denom;
      }
  resc[cidx(cb, I, J, K)] = acc;
}
// lambda = 1 / (alpha*a + sum over d of scale*beta*(b_d(i+e_d) + b_d(i))), the three
// SUMFACES passes accumulated in direction order
template <class T>
__global__ void __launch_bounds__(256) k_vc_lambda(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                                   Geom g, VCArgs<T> v, T* __restrict__ lam, T scale, T alpha,
                                                   T beta) {
  DBox b;
  int box, i, j, k;
  if (!tile_cell(boxes, tiles, g, box, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  const DBox f0 = v.f0[box], f1 = v.f1[box], f2 = v.f2[box];
  const long long x0 = fidx(f0, i, j, k), x1 = fidx(f1, i, j, k), x2 = fidx(f2, i, j, k);
  T l = v.a[c] * alpha;
  T sumVal = v.b0[x0 + 1] + v.b0[x0];
  l = l + scale * beta * sumVal;
  sumVal = v.b1[x1 + f1.sy] + v.b1[x1];
  l = l + scale * beta * sumVal;
  sumVal = v.b2[x2 + f2.sz] + v.b2[x2];
  l = l + scale * beta * sumVal;
  lam[c] = T(1.0) / l;
}
template <class T>
__global__ void __launch_bounds__(256) k_vc_precond(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                                    Geom g, const T* __restrict__ lam, T* phi, const T* rhs) {
  DBox b;
  int box, i, j, k;
  if (!tile_cell(boxes, tiles, g, box, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  phi[c] = rhs[c] * lam[c];
}
// CoarseAverage (AverageF.ChF:9-162): the r^3 children, i fastest; ratios 2 and 4
// are unrolled sums times refScale, other ratios a sum from zero times refScale
template <class T>
__global__ void __launch_bounds__(256) k_vc_avg_cells(const DBox* __restrict__ cboxes, const Tile* __restrict__ ctiles,
                                                      Geom g, const DBox* __restrict__ fboxes, T* __restrict__ coarse,
                                                      const T* __restrict__ fine, int r, T refScale, int harmonic) {
  DBox cb;
  int box, I, J, K;
  if (!tile_cell(cboxes, ctiles, g, box, cb, I, J, K)) return;
  const DBox fb = fboxes[box];
  T sum = T(0.0);
  bool first = r <= 4;
  for (int kk = 0; kk < r; kk++)
    for (int jj = 0; jj < r; jj++)
      for (int ii = 0; ii < r; ii++) {
        const T v = fine[cidx(fb, I * r + ii, J * r + jj, K * r + kk)];
        const T f = harmonic ? T(1.0) / v : v;
        if (first) {
          sum = f;
          first = false;
        } else {
          sum = sum + f;
        }
      }
  // AVERAGEHARMONIC (AverageF.ChF:165-314): the reciprocal of refScale times the sum of 1/fine
  if (harmonic) coarse[cidx(cb, I, J, K)] = T(1.0) / ((r <= 4) ? refScale * sum : sum * refScale);
  else coarse[cidx(cb, I, J, K)] = (r <= 4) ? refScale * sum : sum * refScale;
}
// CoarseAverageFace (AverageFaceF.ChF AVERAGEFACE): each coarse face sums its r^2 fine
// faces (i fastest) from zero, times (1/r)^2.  A thread owns the low face of its
// cell in direction d, and the high face too when its cell is the last of the box.
// harmonic: AVERAGEFACEHARMONIC (AverageFaceF.ChF:84-128), 1 / (ref_scale * sum of 1/fine).
template <class T>
__device__ __forceinline__ T face_avg(const T* __restrict__ fine, const DBox& ff, int d, int I, int J, int K, int r,
                                      T ref_scale, int harmonic) {
  const int n0 = d == 0 ? 1 : r, n1 = d == 1 ? 1 : r, n2 = d == 2 ? 1 : r;
  T crseSum = T(0.0);
  for (int kk = 0; kk < n2; kk++)
    for (int jj = 0; jj < n1; jj++)
      for (int ii = 0; ii < n0; ii++) {
        const T v = fine[fidx(ff, I * r + ii, J * r + jj, K * r + kk)];
        crseSum = crseSum + (harmonic ? T(1.0) / v : v);
      }
  return harmonic ? T(1.0) / (ref_scale * crseSum) : ref_scale * crseSum;
}
template <class T>
__global__ void __launch_bounds__(256) k_vc_avg_faces(const DBox* __restrict__ cboxes, const Tile* __restrict__ ctiles,
                                                      Geom g, const DBox* __restrict__ cf, const DBox* __restrict__ ff,
                                                      T* __restrict__ coarse, const T* __restrict__ fine, int d, int r,
                                                      T ref_scale, int harmonic) {
  DBox cb;
  int box, I, J, K;
  if (!tile_cell(cboxes, ctiles, g, box, cb, I, J, K)) return;
  const DBox c = cf[box], f = ff[box];
  coarse[fidx(c, I, J, K)] = face_avg(fine, f, d, I, J, K, r, ref_scale, harmonic);
  const int idx[3] = {I, J, K};
  if (idx[d] == cb.hi[d]) {
    const int I1 = I + (d == 0), J1 = J + (d == 1), K1 = K + (d == 2);
    coarse[fidx(c, I1, J1, K1)] = face_avg(fine, f, d, I1, J1, K1, r, ref_scale, harmonic);
  }
}

// ---------------------------------------------------------------------------
// One whole GSRB iteration of a merged level in one launch, variable coefficients.
//
// The scheme of the constant-coefficient one-launch relaxation (ops::fused_column,
// whose equivalence with the reference's ghost refresh / red / ghost refresh /
// black sequence is derived at gsrb_fused in kernels.hip): a workgroup marches up
// a column of FX x FY cells; five LDS planes hold phi over the column grown by 2;
// step k loads plane k+2, updates the red cells of plane k+1 on the column grown
// by 1, the black cells of plane k on the column, and stores plane k-1.  Face
// ghosts outside the rectangle are formed in registers at the moment of the
// update, from the operands the ghost refresh would use.  The update is
// GSRBHELMHOLTZVC's, with lambda formed in registers by resetLambda's expression
// (SUMFACES: the operations and order of k_vc_lambda, hence the same value), so
// the lambda array is not read.  Per updated cell the kernel reads rhs, aCoef and
// the cell's six faces from global memory, two steps ahead of their use.
// ---------------------------------------------------------------------------
template <class T>
struct VCFused {
  const T* a;
  const T* b0;
  const T* b1;
  const T* b2;
  long long foff[3];  // face index of the rectangle's low corner, per direction
  int fsy[3], fsz[3];
};
template <class T>
struct VCCell {  // operands of one update
  T r, a, b0m, b0p, b1m, b1p, b2m, b2p;
};
// operands of cell (i, j, k), index clamped into the rectangle (a cell that is not
// updated still loads from a valid address, and its values are never used)
template <class T>
__device__ __forceinline__ void vc_fetch_cell(VCCell<T>& q, const ops::FusedArgs& a, const VCFused<T>& v,
                                              const T* __restrict__ rhs, int i, int j, int k) {
  const int di = ops::clampi(i, a.lo[0], a.hi[0]) - a.lo[0];
  const int dj = ops::clampi(j, a.lo[1], a.hi[1]) - a.lo[1];
  const int dk = ops::clampi(k, a.lo[2], a.hi[2]) - a.lo[2];
  const long long c = a.off + (di + 1) + (long long)(dj + 1) * a.sy + (long long)(dk + 1) * a.sz;
  const long long x0 = v.foff[0] + di + (long long)dj * v.fsy[0] + (long long)dk * v.fsz[0];
  const long long x1 = v.foff[1] + di + (long long)dj * v.fsy[1] + (long long)dk * v.fsz[1];
  const long long x2 = v.foff[2] + di + (long long)dj * v.fsy[2] + (long long)dk * v.fsz[2];
  q.r = rhs[c];
  q.a = v.a[c];
  q.b0m = v.b0[x0];
  q.b0p = v.b0[x0 + 1];
  q.b1m = v.b1[x1];
  q.b1p = v.b1[x1 + v.fsy[1]];
  q.b2m = v.b2[x2];
  q.b2p = v.b2[x2 + v.fsz[2]];
}
template <class T, int LW>
__device__ __forceinline__ T vc_fused_new(const ops::FusedArgs& a, const T* Sm, const T* S0, const T* Sp, int c,
                                          int i, int j, int k, const VCCell<T>& q, T dxinv, T scale, T alpha,
                                          T beta, T c1, T c2) {
  const T own = S0[c];
  const T vxp = S0[c + 1], vxm = S0[c - 1];
  const T vyp = S0[c + LW], vym = S0[c - LW];
  const T vzp = Sp[c], vzm = Sm[c];
  const T xp = (i == a.hi[0]) ? ops::face_ghost(a.kind[0][1], own, vxm, c1, c2) : vxp;
  const T xm = (i == a.lo[0]) ? ops::face_ghost(a.kind[0][0], own, vxp, c1, c2) : vxm;
  const T yp = (j == a.hi[1]) ? ops::face_ghost(a.kind[1][1], own, vym, c1, c2) : vyp;
  const T ym = (j == a.lo[1]) ? ops::face_ghost(a.kind[1][0], own, vyp, c1, c2) : vym;
  const T zp = (k == a.hi[2]) ? ops::face_ghost(a.kind[2][1], own, vzm, c1, c2) : vzp;
  const T zm = (k == a.lo[2]) ? ops::face_ghost(a.kind[2][0], own, vzp, c1, c2) : vzm;
  // lambda (k_vc_lambda's expression)
  T l = q.a * alpha;
  T sumVal = q.b0p + q.b0m;
  l = l + scale * beta * sumVal;
  sumVal = q.b1p + q.b1m;
  l = l + scale * beta * sumVal;
  sumVal = q.b2p + q.b2m;
  l = l + scale * beta * sumVal;
  const T lam = T(1.0) / l;
  // GSRBHELMHOLTZVC (vc_lofphi's expression)
  const T s = q.b0p * (xp - own) - q.b0m * (own - xm) + q.b1p * (yp - own) - q.b1m * (own - ym) +
              q.b2p * (zp - own) - q.b2m * (own - zm);
  const T lofphi = alpha * q.a * own - beta * s * dxinv;
  return own - lam * (lofphi - q.r);
}
template <class T, int FX, int FY, int NT>
__global__ void __launch_bounds__(NT) k_vc_gsrb_fused(ops::FusedArgs a, VCFused<T> v, T* __restrict__ dst,
                                                      const T* __restrict__ src, const T* __restrict__ rhs, T dxinv,
                                                      T scale, T alpha, T beta, T c1, T c2) {
  constexpr int LW = FX + 4, LH = FY + 4, LPLANE = LW * LH;
  constexpr int NLD = (LPLANE + NT - 1) / NT;
  constexpr int NRC = (FY + 2) * (FX / 2 + 1);
  constexpr int NRED = (NRC + NT - 1) / NT;
  constexpr int NBC = FX * FY / 2;
  constexpr int NBLK = (NBC + NT - 1) / NT;
  constexpr int NST = (FX * FY + NT - 1) / NT;
  static_assert(FX % 2 == 0, "column width must be even");
  __shared__ T P[ops::NSLOT][LPLANE];
  const int w = ops::xcd_contiguous_block();
  if (w >= a.nwork) return;
  const int tx = w % a.ntx, ty = (w / a.ntx) % a.nty, ch = w / (a.ntx * a.nty);
  const int x0 = a.lo[0] + tx * FX, y0 = a.lo[1] + ty * FY;
  const int k0 = a.lo[2] + ch * a.kc;
  const int k1 = min(k0 + a.kc, a.hi[2] + 1);
  const int tid = threadIdx.x;
  auto slot = [&](int p) { return (p - k0 + 2) % ops::NSLOT; };
  auto gidx = [&](int i, int j, int k) {
    return a.off + (i - a.lo[0] + 1) + (long long)(j - a.lo[1] + 1) * a.sy + (long long)(k - a.lo[2] + 1) * a.sz;
  };
  auto lidx = [&](int i, int j) { return (j - y0 + 2) * LW + (i - x0 + 2); };
  auto red_ij = [&](int q, int r, int& i, int& j) {
    j = y0 - 1 + q / (FX / 2 + 1);
    i = x0 - 1 + ((x0 - 1 + j + r) & 1) + 2 * (q % (FX / 2 + 1));
  };
  auto black_ij = [&](int q, int k, int& i, int& j) {
    j = y0 + q / (FX / 2);
    i = x0 + ((x0 + j + k + 1) & 1) + 2 * (q % (FX / 2));
  };
  struct Pre {
    T phi[NLD];
    VCCell<T> rr[NRED], rb[NBLK];
  };
  Pre A, B;
  auto fetch = [&](Pre& f, int p, int r, int k) {  // plane p; red cells of plane r; black cells of plane k
    const int pc = ops::clampi(p, a.lo[2], a.hi[2]);
#pragma unroll
    for (int u = 0; u < NLD; u++) {
      const int q = tid + NT * u;
      const int i = x0 - 2 + q % LW, j = y0 - 2 + q / LW;
      f.phi[u] = src[gidx(ops::clampi(i, a.lo[0], a.hi[0]), ops::clampi(j, a.lo[1], a.hi[1]), pc)];
    }
#pragma unroll
    for (int u = 0; u < NRED; u++) {
      int i, j;
      red_ij(tid + NT * u, r, i, j);
      vc_fetch_cell(f.rr[u], a, v, rhs, i, j, r);
    }
#pragma unroll
    for (int u = 0; u < NBLK; u++) {
      int i, j;
      black_ij(tid + NT * u, k, i, j);
      vc_fetch_cell(f.rb[u], a, v, rhs, i, j, k);
    }
  };
  auto put_phi = [&](const Pre& f, int p) {
    T* S = P[slot(p)];
#pragma unroll
    for (int u = 0; u < NLD; u++) {
      const int q = tid + NT * u;
      if (q < LPLANE) S[q] = f.phi[u];
    }
  };
  auto store = [&](int p) {
    const T* S = P[slot(p)];
    T val[NST];
#pragma unroll
    for (int u = 0; u < NST; u++) {
      const int q = min(tid + NT * u, FX * FY - 1);
      val[u] = S[(q / FX + 2) * LW + (q % FX + 2)];
    }
#pragma unroll
    for (int u = 0; u < NST; u++) {
      const int q = tid + NT * u;
      const int i = x0 + q % FX, j = y0 + q / FX;
      if (q < FX * FY && i <= a.hi[0] && j <= a.hi[1]) dst[gidx(i, j, p)] = val[u];
    }
  };
  // step k consumes the set f (plane k+2, red and black cells of plane k+1) and
  // then refills it for step k+2, so the loads have step k+1 to land.  A plane's
  // red and black operands share cache lines and are loaded together; the black
  // ones are used a step after the red ones, so step k keeps them in `carry` for
  // step k+1 and relaxes its own black cells (plane k) from the previous carry.
  VCCell<T> carry[NBLK];
  auto step = [&](Pre& f, int k) {
    put_phi(f, k + 2);
    ops::lds_barrier();
    if (k - 1 >= k0) store(k - 1);
    const int r = k + 1;
    if (r >= a.lo[2] && r <= a.hi[2]) {
#pragma unroll
      for (int u = 0; u < NRED; u++) {
        const int q = tid + NT * u;
        int i, j;
        red_ij(q, r, i, j);
        if (q < NRC && i >= a.lo[0] && i <= a.hi[0] && j >= a.lo[1] && j <= a.hi[1]) {
          const int c = lidx(i, j);
          P[slot(r)][c] = vc_fused_new<T, LW>(a, P[slot(r - 1)], P[slot(r)], P[slot(r + 1)], c, i, j, r, f.rr[u],
                                              dxinv, scale, alpha, beta, c1, c2);
        }
      }
    }
    ops::lds_barrier();
    if (k >= k0) {
#pragma unroll
      for (int u = 0; u < NBLK; u++) {
        const int q = tid + NT * u;
        int i, j;
        black_ij(q, k, i, j);
        if (q < NBC && i <= a.hi[0] && j <= a.hi[1]) {
          const int c = lidx(i, j);
          P[slot(k)][c] = vc_fused_new<T, LW>(a, P[slot(k - 1)], P[slot(k)], P[slot(k + 1)], c, i, j, k, carry[u],
                                              dxinv, scale, alpha, beta, c1, c2);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < NBLK; u++) carry[u] = f.rb[u];
    fetch(f, k + 4, k + 3, k + 3);
  };
  fetch(A, k0 - 2, k0 - 3, k0 - 3);
  put_phi(A, k0 - 2);
  fetch(A, k0 - 1, k0 - 3, k0 - 3);
  put_phi(A, k0 - 1);
  fetch(A, k0, k0 - 1, k0 - 1);
  fetch(B, k0 + 1, k0, k0);
#pragma unroll
  for (int u = 0; u < NBLK; u++) carry[u] = VCCell<T>{};  // black operands of plane k0-2: never used
  for (int k = k0 - 2; k < k1; k += 2) {
    step(A, k);
    if (k + 1 < k1) step(B, k + 1);
  }
  ops::lds_barrier();
  store(k1 - 1);
}
// tile shapes of the one-launch VC relaxation {FX, FY, NT}
constexpr int kVCShapes[][3] = {{64, 8, 256}, {32, 8, 256}, {64, 4, 256}, {64, 16, 512}};
constexpr int kVCShapeCount = 4;

#define VC_LAUNCH(L, kind, bytes, units, kernel, grid, ...)                                \
  do {                                                                                     \
    if ((grid) > 0 && !(L).ctx->dry) {                                                     \
      hipEvent_t ev_ = profile_begin((L).ctx, kind);                                       \
      hipLaunchKernelGGL(kernel, dim3(grid), dim3(256), 0, (L).ctx->stream, __VA_ARGS__);  \
      hip_check(hipGetLastError(), #kernel);                                               \
      profile_end((L).ctx, ev_, kind, bytes, units);                                       \
    }                                                                                      \
  } while (0)

template <class T>
void need_coefs(const DeviceLayout& L, const VCCoefs<T>& C, const char* what) {
  if (C.lay != &L) throw std::runtime_error(std::string(what) + ": coefficients belong to another level");
}
// the fine and coarse layouts pair box for box, the coarse boxes the fine ones coarsened by r
void check_ratio(const DeviceLayout& Lf, const DeviceLayout& Lc, int r, const char* what) {
  if (Lf.nbox != Lc.nbox || Lf.merged != Lc.merged || Lf.h_boxes.size() != Lc.h_boxes.size())
    throw std::runtime_error(std::string(what) + ": layouts do not pair box for box");
  for (size_t b = 0; b < Lf.h_boxes.size(); b++)
    for (int q = 0; q < 3; q++) {
      const int lo = Lf.h_boxes[b].lo[q], hi = Lf.h_boxes[b].hi[q];
      if (lo % r != 0 || (hi + 1) % r != 0 || lo / r != Lc.h_boxes[b].lo[q] || (hi + 1) / r - 1 != Lc.h_boxes[b].hi[q])
        throw std::runtime_error(std::string(what) + ": coarse box is not the fine box coarsened");
    }
}

}  // namespace

namespace vc {

// algorithmic bytes of one red-black iteration: phi read and written, rhs, aCoef
// and the three face arrays read once (7 values per cell: the high face of a cell
// is the next cell's low face, and lambda is a function of aCoef and bCoef); a
// colour launch is half an iteration
template <class T>
static double gsrb_bytes(const DeviceLayout& L) {
  return 3.5 * sizeof(T) * L.valid_cells;
}

template <class T>
void gsrb(const DeviceLayout& L, const VCCoefs<T>& C, T* phi, const T* rhs, int color, double dx, double alpha,
          double beta) {
  need_coefs(L, C, "vc::gsrb");
  const double dxinv = 1.0 / (dx * dx);
  VC_LAUNCH(L, KC_GSRB, gsrb_bytes<T>(L), 0.5 * L.valid_cells, k_vc_gsrb<T>, L.ntiles, L.d_boxes, L.d_tiles, geom(L),
            args(C), phi, rhs, color, T(dxinv), T(alpha), T(beta));
}
template <class T>
void gsrb_fused(const DeviceLayout& L, const ops::FaceRules& fr, const VCCoefs<T>& C, T* dst, const T* src,
                const T* rhs, double dx, double alpha, double beta, int shape) {
  need_coefs(L, C, "vc::gsrb_fused");
  if (!L.merged) throw std::runtime_error("vc::gsrb_fused: level is not stored merged");
  if (dst == src || dst == rhs) throw std::runtime_error("vc::gsrb_fused: dst must be a separate field");
  const DBox& R = L.h_boxes[L.nbox];
  // automatic: 64-wide columns where the level is that wide, 4 rows in FP64 (fewer
  // registers: more waves in flight), 8 in FP32; 32 x 8 on narrower levels
  if (shape < 0) shape = (R.hi[0] - R.lo[0] + 1 >= 64) ? (sizeof(T) == 8 ? 2 : 0) : 1;
  if (shape >= kVCShapeCount) throw std::runtime_error("vc::gsrb_fused: bad tile shape");
  const int FX = kVCShapes[shape][0], FY = kVCShapes[shape][1];
  ops::FusedArgs a;
  for (int q = 0; q < 3; q++) {
    a.lo[q] = R.lo[q];
    a.hi[q] = R.hi[q];
  }
  for (int d = 0; d < 3; d++)
    for (int s = 0; s < 2; s++) {
      if (fr.kind[d][s] < 0 || fr.kind[d][s] > 2) throw std::runtime_error("vc::gsrb_fused: bad face rule");
      a.kind[d][s] = fr.kind[d][s];
    }
  a.off = R.off;
  a.sy = R.sy;
  a.sz = R.sz;
  a.zsrc = 0;
  a.ntx = (R.hi[0] - R.lo[0] + 1 + FX - 1) / FX;
  a.nty = (R.hi[1] - R.lo[1] + 1 + FY - 1) / FY;
  const int nz = R.hi[2] - R.lo[2] + 1;
  int kc = 32;
  while (kc > 4 && (long long)a.ntx * a.nty * ((nz + kc - 1) / kc) < 2048) kc >>= 1;
  a.kc = kc;
  const long long nwork = (long long)a.ntx * a.nty * ((nz + kc - 1) / kc);
  if (nwork > INT32_MAX - 8) throw std::runtime_error("vc::gsrb_fused: level too large");
  a.nwork = (int)nwork;
  VCFused<T> v;
  v.a = C.a.d;
  v.b0 = C.b[0];
  v.b1 = C.b[1];
  v.b2 = C.b[2];
  for (int d = 0; d < 3; d++) {
    const DBox& f = C.h_fbox[d][L.nbox];
    v.foff[d] = f.off;
    v.fsy[d] = f.sy;
    v.fsz[d] = f.sz;
  }
  const int grid = (a.nwork + 7) / 8 * 8;
  const double dxinv = 1.0 / (dx * dx), scale = 1.0 / (dx * dx);
  const double bytes = 2.0 * gsrb_bytes<T>(L), units = (double)L.valid_cells;
  if (L.ctx->dry) return;
  hipEvent_t ev = profile_begin(L.ctx, KC_GSRB);
#define VC_FUSED_CASE(n, X, Y, NT)                                                                             \
  case n:                                                                                                      \
    hipLaunchKernelGGL((k_vc_gsrb_fused<T, X, Y, NT>), dim3(grid), dim3(NT), 0, L.ctx->stream, a, v, dst, src, rhs, \
                       T(dxinv), T(scale), T(alpha), T(beta), T(fr.c1), T(fr.c2));                              \
    break;
  switch (shape) {
    VC_FUSED_CASE(0, 64, 8, 256)
    VC_FUSED_CASE(1, 32, 8, 256)
    VC_FUSED_CASE(2, 64, 4, 256)
    VC_FUSED_CASE(3, 64, 16, 512)
  }
#undef VC_FUSED_CASE
  hip_check(hipGetLastError(), "k_vc_gsrb_fused");
  profile_end(L.ctx, ev, KC_GSRB, bytes, units);
}
template <class T>
void apply(const DeviceLayout& L, const VCCoefs<T>& C, T* lofphi, const T* phi, double dx, double alpha,
           double beta) {
  need_coefs(L, C, "vc::apply");
  const double dxinv = 1.0 / (dx * dx);
  VC_LAUNCH(L, KC_APPLY, 6.0 * sizeof(T) * L.valid_cells, 0.0, k_vc_apply<T>, L.ntiles, L.d_boxes, L.d_tiles, geom(L),
            args(C), lofphi, phi, T(dxinv), T(alpha), T(beta));
}
template <class T>
void residual(const DeviceLayout& L, const VCCoefs<T>& C, T* res, const T* phi, const T* rhs, double dx,
              double alpha, double beta) {
  need_coefs(L, C, "vc::residual");
  const double dxinv = 1.0 / (dx * dx);
  VC_LAUNCH(L, KC_RESIDUAL, 7.0 * sizeof(T) * L.valid_cells, 0.0, k_vc_residual<T>, L.ntiles, L.d_boxes, L.d_tiles,
            geom(L), args(C), res, phi, rhs, T(dxinv), T(alpha), T(beta));
}
template <class T>
void restrict_res(const DeviceLayout& Lf, const DeviceLayout& Lc, const VCCoefs<T>& C, T* resc, const T* phi,
                  const T* rhs, double dx, double alpha, double beta) {
  need_coefs(Lf, C, "vc::restrict_res");
  check_ratio(Lf, Lc, 2, "vc::restrict_res");
  const double dxinv = 1.0 / (dx * dx);
  VC_LAUNCH(Lc, KC_RESTRICT, (6.0 + 1.0 / 8.0) * sizeof(T) * Lf.valid_cells, 0.0, k_vc_restrict<T>, Lc.ntiles,
            Lc.d_boxes, Lc.d_tiles, geom(Lc), Lf.d_boxes, args(C), resc, phi, rhs, T(dxinv), T(alpha), T(beta));
}
template <class T>
void lambda(VCCoefs<T>& C, double dx, double alpha, double beta) {
  const DeviceLayout& L = *C.lay;
  const double scale = 1.0 / (dx * dx);
  VC_LAUNCH(L, KC_BLAS, 0.0, 0.0, k_vc_lambda<T>, L.ntiles, L.d_boxes, L.d_tiles, geom(L), args(C), C.lambda.d,
            T(scale), T(alpha), T(beta));
}
template <class T>
void precond_init(const DeviceLayout& L, const VCCoefs<T>& C, T* phi, const T* rhs) {
  need_coefs(L, C, "vc::precond_init");
  VC_LAUNCH(L, KC_BLAS, 0.0, 0.0, k_vc_precond<T>, L.ntiles, L.d_boxes, L.d_tiles, geom(L), C.lambda.d, phi, rhs);
}
template <class T>
void average_coefs(const VCCoefs<T>& F, VCCoefs<T>& C, int r, int harmonic) {
  const DeviceLayout& Lf = *F.lay;
  const DeviceLayout& Lc = *C.lay;
  if (r < 2 || (r & (r - 1))) throw std::runtime_error("vc::average_coefs: ratio must be a power of two");
  check_ratio(Lf, Lc, r, "vc::average_coefs");
  const double refScale = 1.0 / (double)(r * r * r);
  VC_LAUNCH(Lc, KC_AVERAGE, 0.0, 0.0, k_vc_avg_cells<T>, Lc.ntiles, Lc.d_boxes, Lc.d_tiles, geom(Lc), Lf.d_boxes,
            C.a.d, F.a.d, r, T(refScale), harmonic);
  const double ref_scale = (1.0 / r) * (1.0 / r);
  for (int d = 0; d < 3; d++)
    VC_LAUNCH(Lc, KC_AVERAGE, 0.0, 0.0, k_vc_avg_faces<T>, Lc.ntiles, Lc.d_boxes, Lc.d_tiles, geom(Lc), C.d_fbox[d],
              F.d_fbox[d], C.b[d], F.b[d], d, r, T(ref_scale), harmonic);
}

template <class T>
__global__ void __launch_bounds__(256) k_vc_scale_a(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                                    Geom g, T* x, const T* __restrict__ a, int divide) {
  DBox b;
  int box, i, j, k;
  if (!tile_cell(boxes, tiles, g, box, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  x[c] = divide ? x[c] / a[c] : x[c] * a[c];
}
template <class T>
void scale_by_a(const DeviceLayout& L, const VCCoefs<T>& C, T* x, bool divide) {
  need_coefs(L, C, "vc::scale_by_a");
  VC_LAUNCH(L, KC_BLAS, 0.0, 0.0, k_vc_scale_a<T>, L.ntiles, L.d_boxes, L.d_tiles, geom(L), x, C.a.d, divide ? 1 : 0);
}
template <class T>
void scale_by_lambda(const DeviceLayout& L, const VCCoefs<T>& C, T* x) {
  need_coefs(L, C, "vc::scale_by_lambda");
  VC_LAUNCH(L, KC_BLAS, 0.0, 0.0, k_vc_scale_a<T>, L.ntiles, L.d_boxes, L.d_tiles, geom(L), x, C.lambda.d, 0);
}

template <class T>
struct VCB3 {
  const T* b[3];
};
template <class T>
__global__ void __launch_bounds__(256) k_vc_reflux(const RefluxCell* __restrict__ cells, int n,
                                                   const RefluxCoarse* __restrict__ rc, const RefluxFine* __restrict__ rf,
                                                   const VCFaceIdx* __restrict__ fc, const VCFaceIdx* __restrict__ ff,
                                                   T* u, const T* __restrict__ phic, const T* __restrict__ phif,
                                                   VCB3<T> bc, VCB3<T> bf, T beta_dx_c, T beta_ref_dx_c, T minus_s) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const RefluxCell c = cells[r];
  T val = u[c.u];
  if (c.nc > 0) {
    T reg = T(0.0);
    for (int q = c.c0; q < c.c0 + c.nc; q++) {
      const RefluxCoarse x = rc[q];
      const VCFaceIdx f = fc[q];
      T F = (phic[x.a] - phic[x.b]) * beta_dx_c;
      F = -bc.b[f.dir][f.x[0]] * F;
      reg += F * T(x.sign);
    }
    val = val + reg * minus_s;
  }
  T freg = T(0.0);
  for (int q = c.f0; q < c.f0 + c.nf; q++) {
    const RefluxFine x = rf[q];
    const VCFaceIdx f = ff[q];
    if (x.first) freg = T(0.0);  // a chain of records per register with ratio 4
    for (int t = 0; t < 4; t++) {
      T F = (phif[x.a[t]] - phif[x.b[t]]) * beta_ref_dx_c;
      F = -bf.b[f.dir][f.x[t]] * F;
      freg = freg + T(x.scale) * F;
    }
    if (x.last) val = val + freg * minus_s;
  }
  u[c.u] = val;
}
template <class T>
void reflux(Context* ctx, const RefluxCell* cells, int ncells, const RefluxCoarse* rc, const RefluxFine* rf,
            const VCFaceIdx* fc, const VCFaceIdx* ff, T* u, const T* phic, const T* phif, const VCCoefs<T>& C,
            const VCCoefs<T>& F, double beta_dx_c, double beta_ref_dx_c, double minus_inv_dx) {
  if (ncells <= 0 || ctx->dry) return;
  VCB3<T> bc{{C.b[0], C.b[1], C.b[2]}}, bf{{F.b[0], F.b[1], F.b[2]}};
  hipEvent_t ev = profile_begin(ctx, KC_REFLUX);
  hipLaunchKernelGGL(k_vc_reflux<T>, dim3((ncells + 255) / 256), dim3(256), 0, ctx->stream, cells, ncells, rc, rf, fc,
                     ff, u, phic, phif, bc, bf, T(beta_dx_c), T(beta_ref_dx_c), T(minus_inv_dx));
  hip_check(hipGetLastError(), "k_vc_reflux");
  profile_end(ctx, ev, KC_REFLUX, 0.0);
}

__global__ void __launch_bounds__(256) k_vc_narrow(long long n, float* __restrict__ dst, const double* __restrict__ src) {
  for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += (long long)gridDim.x * 256) dst[q] = (float)src[q];
}
void narrow(const VCCoefs<double>& src, VCCoefs<float>& dst) {
  const DeviceLayout& L = *dst.lay;
  if (src.lay != dst.lay) throw std::runtime_error("vc::narrow: coefficients of two levels");
  auto grid = [](long long n) { return (int)std::min<long long>((n + 255) / 256, 1 << 16); };
  VC_LAUNCH(L, KC_BLAS, 0.0, 0.0, k_vc_narrow, grid(L.slab), L.slab, dst.a.d, src.a.d);
  for (int d = 0; d < 3; d++)
    VC_LAUNCH(L, KC_BLAS, 0.0, 0.0, k_vc_narrow, grid(src.fslab[d]), src.fslab[d], dst.b[d], src.b[d]);
}

#define VC_INST(T)                                                                                                  \
  template void gsrb<T>(const DeviceLayout&, const VCCoefs<T>&, T*, const T*, int, double, double, double);        \
  template void gsrb_fused<T>(const DeviceLayout&, const ops::FaceRules&, const VCCoefs<T>&, T*, const T*,          \
                              const T*, double, double, double, int);                                               \
  template void apply<T>(const DeviceLayout&, const VCCoefs<T>&, T*, const T*, double, double, double);            \
  template void residual<T>(const DeviceLayout&, const VCCoefs<T>&, T*, const T*, const T*, double, double, double); \
  template void restrict_res<T>(const DeviceLayout&, const DeviceLayout&, const VCCoefs<T>&, T*, const T*,          \
                                const T*, double, double, double);                                                  \
  template void lambda<T>(VCCoefs<T>&, double, double, double);                                                     \
  template void precond_init<T>(const DeviceLayout&, const VCCoefs<T>&, T*, const T*);                              \
  template void average_coefs<T>(const VCCoefs<T>&, VCCoefs<T>&, int, int);                                             \
  template void scale_by_a<T>(const DeviceLayout&, const VCCoefs<T>&, T*, bool);                                    \
  template void scale_by_lambda<T>(const DeviceLayout&, const VCCoefs<T>&, T*);                                     \
  template void reflux<T>(Context*, const RefluxCell*, int, const RefluxCoarse*, const RefluxFine*, const VCFaceIdx*, \
                          const VCFaceIdx*, T*, const T*, const T*, const VCCoefs<T>&, const VCCoefs<T>&, double,     \
                          double, double);
VC_INST(double)
VC_INST(float)

}  // namespace vc
}  // namespace amrmg
