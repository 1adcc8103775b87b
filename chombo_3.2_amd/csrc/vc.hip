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
                                                      const T* __restrict__ fine, int r, T refScale) {
  DBox cb;
  int box, I, J, K;
  if (!tile_cell(cboxes, ctiles, g, box, cb, I, J, K)) return;
  const DBox fb = fboxes[box];
  T sum = T(0.0);
  bool first = r <= 4;
  for (int kk = 0; kk < r; kk++)
    for (int jj = 0; jj < r; jj++)
      for (int ii = 0; ii < r; ii++) {
        const T f = fine[cidx(fb, I * r + ii, J * r + jj, K * r + kk)];
        if (first) {
          sum = f;
          first = false;
        } else {
          sum = sum + f;
        }
      }
  coarse[cidx(cb, I, J, K)] = (r <= 4) ? refScale * sum : sum * refScale;
}
// CoarseAverageFace (AverageFaceF.ChF AVERAGEFACE): each coarse face sums its r^2 fine
// faces (i fastest) from zero, times (1/r)^2.  A thread owns the low face of its
// cell in direction d, and the high face too when its cell is the last of the box.
template <class T>
__device__ __forceinline__ T face_avg(const T* __restrict__ fine, const DBox& ff, int d, int I, int J, int K, int r,
                                      T ref_scale) {
  const int n0 = d == 0 ? 1 : r, n1 = d == 1 ? 1 : r, n2 = d == 2 ? 1 : r;
  T crseSum = T(0.0);
  for (int kk = 0; kk < n2; kk++)
    for (int jj = 0; jj < n1; jj++)
      for (int ii = 0; ii < n0; ii++) crseSum = crseSum + fine[fidx(ff, I * r + ii, J * r + jj, K * r + kk)];
  return ref_scale * crseSum;
}
template <class T>
__global__ void __launch_bounds__(256) k_vc_avg_faces(const DBox* __restrict__ cboxes, const Tile* __restrict__ ctiles,
                                                      Geom g, const DBox* __restrict__ cf, const DBox* __restrict__ ff,
                                                      T* __restrict__ coarse, const T* __restrict__ fine, int d, int r,
                                                      T ref_scale) {
  DBox cb;
  int box, I, J, K;
  if (!tile_cell(cboxes, ctiles, g, box, cb, I, J, K)) return;
  const DBox c = cf[box], f = ff[box];
  coarse[fidx(c, I, J, K)] = face_avg(fine, f, d, I, J, K, r, ref_scale);
  const int idx[3] = {I, J, K};
  if (idx[d] == cb.hi[d]) {
    const int I1 = I + (d == 0), J1 = J + (d == 1), K1 = K + (d == 2);
    coarse[fidx(c, I1, J1, K1)] = face_avg(fine, f, d, I1, J1, K1, r, ref_scale);
  }
}

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

// bytes per cell moved by one launch (FP64: phi, rhs, a, lambda, three face arrays
// read, phi written; the second face of each pair is the next cell's first)
template <class T>
static double gsrb_bytes(const DeviceLayout& L) {
  return 8.0 * sizeof(T) * L.valid_cells;
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
void average_coefs(const VCCoefs<T>& F, VCCoefs<T>& C, int r) {
  const DeviceLayout& Lf = *F.lay;
  const DeviceLayout& Lc = *C.lay;
  if (r < 2 || (r & (r - 1))) throw std::runtime_error("vc::average_coefs: ratio must be a power of two");
  check_ratio(Lf, Lc, r, "vc::average_coefs");
  const double refScale = 1.0 / (double)(r * r * r);
  VC_LAUNCH(Lc, KC_AVERAGE, 0.0, 0.0, k_vc_avg_cells<T>, Lc.ntiles, Lc.d_boxes, Lc.d_tiles, geom(Lc), Lf.d_boxes,
            C.a.d, F.a.d, r, T(refScale));
  const double ref_scale = (1.0 / r) * (1.0 / r);
  for (int d = 0; d < 3; d++)
    VC_LAUNCH(Lc, KC_AVERAGE, 0.0, 0.0, k_vc_avg_faces<T>, Lc.ntiles, Lc.d_boxes, Lc.d_tiles, geom(Lc), C.d_fbox[d],
              F.d_fbox[d], C.b[d], F.b[d], d, r, T(ref_scale));
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
  template void apply<T>(const DeviceLayout&, const VCCoefs<T>&, T*, const T*, double, double, double);            \
  template void residual<T>(const DeviceLayout&, const VCCoefs<T>&, T*, const T*, const T*, double, double, double); \
  template void restrict_res<T>(const DeviceLayout&, const DeviceLayout&, const VCCoefs<T>&, T*, const T*,          \
                                const T*, double, double, double);                                                  \
  template void lambda<T>(VCCoefs<T>&, double, double, double);                                                     \
  template void precond_init<T>(const DeviceLayout&, const VCCoefs<T>&, T*, const T*);                              \
  template void average_coefs<T>(const VCCoefs<T>&, VCCoefs<T>&, int);
VC_INST(double)
VC_INST(float)

}  // namespace vc
}  // namespace amrmg
