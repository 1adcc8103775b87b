// Batched solver (see batch.hpp): kernels that run one V-cycle stage of every
// problem in one launch, the per-problem bottom BiCGStab as one workgroup per
// problem, and the host driver of the lockstep outer iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <type_traits>

#include "batch.hpp"
#include "zmarch.hpp"

#include <cstring>

namespace amrmg {

namespace {

__device__ __forceinline__ long long ridx(const DBox& b, int i, int j, int k) {
  return b.off + (i - b.lo[0] + 1) + (long long)(j - b.lo[1] + 1) * b.sy + (long long)(k - b.lo[2] + 1) * b.sz;
}

// ConstBCFunction ghost value (BCFunc.cpp:58-158) with the expressions of
// k_fill_ghosts: Dirichlet g = 2v - near, Neumann g = near (+ isign*dx*v)
template <class T>
__device__ __forceinline__ T bc_ghost(const BatchCoef& cf, int dir, int side, T nearv, int homog) {
  const int isign = side ? 1 : -1;
  if (cf.kind[dir][side] == 1) {
    T inhomogVal = T(0.0);
    if (!homog) inhomogVal = T(cf.val[dir][side]);
    return T(2) * inhomogVal - nearv;
  }
  T v = nearv;
  if (!homog) v += T(double(isign) * cf.dx) * T(cf.val[dir][side]);
  return v;
}

// the six neighbours of cell c of rectangle R, with the face ghosts of the
// rectangle formed from the cell itself (the fill-then-read of the reference)
template <class T>
struct Six {
  T xp, xm, yp, ym, zp, zm;
};
template <class T>
__device__ __forceinline__ Six<T> six(const T* phi, const DBox& R, const BatchCoef& cf, long long c, int i, int j,
                                      int k, int homog) {
  const T p = phi[c];
  const long long sy = R.sy, sz = R.sz;
  Six<T> s;
  s.xp = (i == R.hi[0]) ? bc_ghost<T>(cf, 0, 1, p, homog) : phi[c + 1];
  s.xm = (i == R.lo[0]) ? bc_ghost<T>(cf, 0, 0, p, homog) : phi[c - 1];
  s.yp = (j == R.hi[1]) ? bc_ghost<T>(cf, 1, 1, p, homog) : phi[c + sy];
  s.ym = (j == R.lo[1]) ? bc_ghost<T>(cf, 1, 0, p, homog) : phi[c - sy];
  s.zp = (k == R.hi[2]) ? bc_ghost<T>(cf, 2, 1, p, homog) : phi[c + sz];
  s.zm = (k == R.lo[2]) ? bc_ghost<T>(cf, 2, 0, p, homog) : phi[c - sz];
  return s;
}

// cell of this thread in its tile (64 x 4 x 1 layout, TX = 1 << lx, TY = 1 << ly)
__device__ __forceinline__ bool btile_cell(const Tile& t, const DBox& R, int lx, int ly, int& i, int& j, int& k) {
  const int tid = threadIdx.x;
  i = t.i0 + (tid & ((1 << lx) - 1));
  j = t.j0 + ((tid >> lx) & ((1 << ly) - 1));
  k = t.k0 + (tid >> (lx + ly));
  return i <= R.hi[0] && j <= R.hi[1] && k <= R.hi[2];
}

// ---------------------------------------------------------------------------
// one-launch relaxation of every problem present at a depth
// ---------------------------------------------------------------------------
template <class T, bool LAP, int SHAPE, int FX = ops::kFusedShapeTable[SHAPE][0],
          int FY = ops::kFusedShapeTable[SHAPE][1], int NT = ops::kFusedShapeTable[SHAPE][2]>
__global__ void __launch_bounds__(NT)
    k_batch_fused(const ops::FusedArgs* __restrict__ items, const int* __restrict__ wstart, int nitems,
                  const BatchCoef* __restrict__ coef, const int* __restrict__ active, int zsrc, T* __restrict__ dst,
                  const T* __restrict__ src, const T* __restrict__ rhs) {
  const int wg = ops::xcd_contiguous_block();
  if (wg >= wstart[nitems]) return;
  int lo = 0, hi = nitems - 1;  // the problem whose work range holds wg
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (wstart[mid] <= wg) lo = mid;
    else hi = mid - 1;
  }
  const int p = lo;
  if (!active[p] || coef[p].lap != (LAP ? 1 : 0)) return;
  const BatchCoef cf = coef[p];
  ops::FusedArgs a = items[p];
  a.zsrc = zsrc;
  if constexpr (ops::kFusedShapeTable[SHAPE][3] == 2) {
    static_assert(sizeof(T) == 4, "x-quad tiles are FP32 only");
    ops::zmarch4_column<T, LAP, FX, FY, NT>(a, wg - wstart[p], dst, src, rhs, T(cf.dxinv), T(cf.lambda),
                                            T(cf.alpha), T(cf.beta), T(0.0), T(0.0));
  } else if constexpr (ops::kFusedShapeTable[SHAPE][3] == 1)
    ops::zmarch_column<T, LAP, FX, FY, NT>(a, wg - wstart[p], dst, src, rhs, T(cf.dxinv), T(cf.lambda),
                                           T(cf.alpha), T(cf.beta), T(0.0), T(0.0));
  else
    ops::fused_column<T, LAP, FX, FY, NT>(a, wg - wstart[p], dst, src, rhs, T(cf.dxinv), T(cf.lambda),
                                          T(cf.alpha), T(cf.beta), T(0.0), T(0.0));
}

// RESTRICTRES (AMRPoissonOpF.ChF:394-438) after the homogeneous BC refresh: one
// thread per coarse cell, its eight children in the Fortran loop order
template <class T>
__global__ void __launch_bounds__(256)
    k_batch_restrict(const Tile* __restrict__ ctiles, int lx, int ly, const DBox* __restrict__ crect,
                     const DBox* __restrict__ frect, const BatchCoef* __restrict__ fcoef,
                     const int* __restrict__ active, T* __restrict__ resc, const T* __restrict__ phi,
                     const T* __restrict__ rhs) {
  const Tile t = ctiles[blockIdx.x];
  if (!active[t.box]) return;
  const DBox C = crect[t.box], F = frect[t.box];
  const BatchCoef cf = fcoef[t.box];
  int I, J, K;
  if (!btile_cell(t, C, lx, ly, I, J, K)) return;
  const T alpha = T(cf.alpha), beta = T(cf.beta), dxinv = T(cf.dxinv);
  const T denom = T(8.0);
  T acc = T(0.0);
#pragma unroll
  for (int kk = 0; kk < 2; kk++)
#pragma unroll
    for (int jj = 0; jj < 2; jj++)
#pragma unroll
      for (int ii = 0; ii < 2; ii++) {
        const int i = 2 * I + ii, j = 2 * J + jj, k = 2 * K + kk;
        const long long c = ridx(F, i, j, k);
        const T p = phi[c];
        const Six<T> s = six<T>(phi, F, cf, c, i, j, k, 1);
        const T lofphi = alpha * p + beta * (s.xp + s.xm + s.yp + s.ym + s.zp + s.zm - p * T(2.0) * T(3.0)) * dxinv;
        acc = acc + (rhs[c] - lofphi) / denom;
      }
  resc[ridx(C, I, J, K)] = acc;
}

// PROLONG (AMRPoissonOpF.ChF:453-478)
template <class T>
__global__ void __launch_bounds__(256)
    k_batch_prolong(const Tile* __restrict__ ftiles, int lx, int ly, const DBox* __restrict__ frect,
                    const DBox* __restrict__ crect, const int* __restrict__ active, T* __restrict__ phi,
                    const T* __restrict__ coarse) {
  const Tile t = ftiles[blockIdx.x];
  if (!active[t.box]) return;
  const DBox F = frect[t.box], C = crect[t.box];
  int i, j, k;
  if (!btile_cell(t, F, lx, ly, i, j, k)) return;
  const long long c = ridx(F, i, j, k);
  phi[c] = phi[c] + coarse[ridx(C, i >> 1, j >> 1, k >> 1)];
}

// OPERATORLAPRES (AMRPoissonOpF.ChF:341-377) with the inhomogeneous BC refresh, the
// per-problem max-norm of the result, and the result stored in the V-cycle's
// precision (computeAMRResidual of a single level, AMRMultiGrid.H:683-736),
// marching in z over a column tile of 64 x 4 cells and kz planes: the z neighbours
// come from the previous and next planes already in registers; workgroups of
// neighbouring tiles run on one XCD.
template <class TO>
__global__ void __launch_bounds__(256)
    k_batch_residual_norm_z(const Tile* __restrict__ ctiles, int ntiles, int kz, const DBox* __restrict__ rect,
                            const BatchCoef* __restrict__ coef, const int* __restrict__ active,
                            TO* __restrict__ res, const double* __restrict__ phi, const double* __restrict__ rhs,
                            double* norms) {
  const int wg = ops::xcd_contiguous_block();
  if (wg >= ntiles) return;
  const Tile t = ctiles[wg];
  if (!active[t.box]) return;
  const DBox R = rect[t.box];
  const BatchCoef cf = coef[t.box];
  const int i = t.i0 + (threadIdx.x & 63), j = t.j0 + (threadIdx.x >> 6);
  const bool in = i <= R.hi[0] && j <= R.hi[1];
  const int ic = min(i, R.hi[0]), jc = min(j, R.hi[1]);
  const int kend = min(t.k0 + kz - 1, R.hi[2]);
  const long long sy = R.sy, sz = R.sz;
  long long c = ridx(R, ic, jc, t.k0);
  double pm = phi[c - sz], p = phi[c];
  double v = 0.0;
  for (int k = t.k0; k <= kend; k++, c += sz) {
    const double pz = phi[c + sz];  // plane k + 1 <= hi + 1 (the ghost plane)
    double xp = phi[c + 1], xm = phi[c - 1], yp = phi[c + sy], ym = phi[c - sy];
    const double rh = rhs[c];
    if (ic == R.hi[0]) xp = bc_ghost<double>(cf, 0, 1, p, 0);
    if (ic == R.lo[0]) xm = bc_ghost<double>(cf, 0, 0, p, 0);
    if (jc == R.hi[1]) yp = bc_ghost<double>(cf, 1, 1, p, 0);
    if (jc == R.lo[1]) ym = bc_ghost<double>(cf, 1, 0, p, 0);
    const double zp = (k == R.hi[2]) ? bc_ghost<double>(cf, 2, 1, p, 0) : pz;
    const double zm = (k == R.lo[2]) ? bc_ghost<double>(cf, 2, 0, p, 0) : pm;
    const double lap = (xp + xm + yp + ym + zp + zm - 6.0 * p) * cf.dxinv;
    const double r = -(cf.alpha * p) - cf.beta * lap + rh;
    if (in) {
      res[c] = (TO)r;
      v = fmax(v, fabs(r));
    }
    pm = p;
    p = pz;
  }
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  __shared__ double sm[4];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double m = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
    if (m > 0.0) atomicMax((unsigned long long*)(norms + t.box), (unsigned long long)__double_as_longlong(m));
  }
}

// phi += correction (incr(phi, uberCorrection, 1.0) of solveNoInitResid; a mixed
// solve widens the FP32 correction first, as incr_f2d into a zero FP64 field), over a
// column tile of 64 x 4 cells and kz planes
template <class T>
__global__ void __launch_bounds__(256)
    k_batch_update_z(const Tile* __restrict__ ctiles, int ntiles, int kz, const DBox* __restrict__ rect,
                     const int* __restrict__ active, double* __restrict__ phi, const T* __restrict__ corr) {
  const int wg = ops::xcd_contiguous_block();
  if (wg >= ntiles) return;
  const Tile t = ctiles[wg];
  if (!active[t.box]) return;
  const DBox R = rect[t.box];
  const int i = t.i0 + (threadIdx.x & 63), j = t.j0 + (threadIdx.x >> 6);
  if (i > R.hi[0] || j > R.hi[1]) return;
  const int kend = min(t.k0 + kz - 1, R.hi[2]);
  long long c = ridx(R, i, j, t.k0);
  const int n = kend - t.k0 + 1;
  constexpr int U = 4;
  for (int k0 = 0; k0 < n; k0 += U, c += U * (long long)R.sz) {
    double pv[U];
    T cv[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const long long cc = c + (long long)min(u, n - 1 - k0) * R.sz;
      pv[u] = phi[cc];
      cv[u] = corr[cc];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k0 + u >= n) break;
      const long long cc = c + (long long)u * R.sz;
      if constexpr (std::is_same<T, double>::value) {
        phi[cc] = pv[u] + cv[u] * 1.0;
      } else {
        const double x = 0.0 + (double)cv[u];  // uberCorrection (zero) += widened correction
        phi[cc] = pv[u] + x * 1.0;
      }
    }
  }
}

__global__ void k_batch_zero_norms(double* norms, const int* active, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n && active[p]) norms[p] = 0.0;
}

// ---------------------------------------------------------------------------
// bottom of every problem: relax(m_bottom) from zero, then BiCGStabSolver::solve
// (BiCGStabSolver.H:199-519) in FP64 on the deepest level, one workgroup per problem.
// Level operations mirror AMRPoissonOp on that level (homogeneous BC) and the
// reductions keep the reference orders (dot: per box serially in Fortran order,
// boxes in layout order; norm 2: FArrayBox::sumPow column sums per box).
// ---------------------------------------------------------------------------
struct BottomArgs {
  int imax;
  double reps, hang, small;
  int numRestarts;
  int nbottom;                 // relax iterations before the Krylov solve
  int relaxFirst;              // 1: zero the correction and relax nbottom times first (the
                               // whole bottom of MultiGrid::cycle); 0: start from the correction
  int ldsDoubles;              // dynamic LDS of the launch, in doubles: a problem whose ten
                               // Krylov vectors fit keeps them there instead of in scratch
};
// dynamic LDS offered to a single solver's bottom Krylov vectors (the static LDS of
// the kernel is at most 70 KB, the workgroup may hold 160 KB)
constexpr int kBottomLdsDoubles = 8192;

// A walk over the cells of a box of extents (ea, eb, *), a fastest, in steps of
// `step` cells: the position (a, b, c) and its element offset are advanced with
// adds and compares only (no division per cell).
struct Walk {
  int a, b, c, off;
  int da, db, dc, doff, ea, eb, ca, cb;
  __device__ void init(int n, int step, int ea_, int eb_, int base, int s0, int s1, int s2) {
    ea = ea_;
    eb = eb_;
    int r = n / ea;
    a = n - r * ea;
    b = r % eb;
    c = r / eb;
    r = step / ea;
    da = step - r * ea;
    db = r % eb;
    dc = r / eb;
    off = base + a * s0 + b * s1 + c * s2;
    doff = da * s0 + db * s1 + dc * s2;
    ca = s1 - ea * s0;
    cb = s2 - eb * s1;
  }
  __device__ void next() {
    a += da;
    b += db;
    c += dc;
    off += doff;
    if (a >= ea) { a -= ea; b++; off += ca; }
    if (b >= eb) { b -= eb; c++; off += cb; }
  }
};

// The bottom level of one problem, swept by one workgroup.  Fields are addressed
// by offsets from the base of the level's ghosted rectangle R (R.off = 0).  A sweep
// gives each thread U cells per pass and issues all their loads (at always valid
// offsets) before any store, so the loads of a pass are in flight together; the
// cells of a pass are distinct, and cells of one GSRB colour read only the other
// colour, so the results are those of a cell-by-cell loop.
struct BottomLevel {
  BatchCoef cf;
  DBox R;
  int tid, nt;
  int ncell;
  int nx, ny, nz;
  int off0;  // offset of cell lo
  template <class V>
  struct Nb {
    V p, xp, xm, yp, ym, zp, zm, r;
  };
  // the six neighbours with the homogeneous BC ghosts of the rectangle formed
  // from the cell itself (the fill-then-read of the reference)
  template <class V>
  __device__ void bc(Nb<V>& s, int i, int j, int k) const {
    if (i == R.hi[0]) s.xp = bc_ghost<V>(cf, 0, 1, s.p, 1);
    if (i == R.lo[0]) s.xm = bc_ghost<V>(cf, 0, 0, s.p, 1);
    if (j == R.hi[1]) s.yp = bc_ghost<V>(cf, 1, 1, s.p, 1);
    if (j == R.lo[1]) s.ym = bc_ghost<V>(cf, 1, 0, s.p, 1);
    if (k == R.hi[2]) s.zp = bc_ghost<V>(cf, 2, 1, s.p, 1);
    if (k == R.lo[2]) s.zm = bc_ghost<V>(cf, 2, 0, s.p, 1);
  }
  template <class V>
  __device__ Nb<V> load_nb(const V* phi, const V* rhs, int c) const {
    Nb<V> s;
    s.p = phi[c];
    s.xp = phi[c + 1];
    s.xm = phi[c - 1];
    s.yp = phi[c + R.sy];
    s.ym = phi[c - R.sy];
    s.zp = phi[c + R.sz];
    s.zm = phi[c - R.sz];
    s.r = rhs ? rhs[c] : V(0);
    return s;
  }
  // one GSRB colour (GSRBLAPLACIAN / GSRBHELMHOLTZ), cells of the colour walked
  // over half-rows of (nx + 1) / 2 candidates; with src, phi is not read: every
  // value of phi is taken as src * mult (the preconditioner's phi = rhs * mult
  // folded into its first colour), and the cells of the other colour are written
  // with that value
  template <int U, class V>
  __device__ void gsrb(V* phi, const V* rhs, int color, const V* src = nullptr, V mult = V(0)) const {
    const V dxinv = V(cf.dxinv), lambda = V(cf.lambda), alpha = V(cf.alpha), beta = V(cf.beta);
    const int nxh = (nx + 1) / 2, tot = nxh * ny * nz;
    const int par0 = R.lo[0] + R.lo[1] + R.lo[2] + color;
    Walk w;
    w.init(tid, nt, nxh, ny, off0, 2, R.sy, R.sz);
    for (int q0 = tid; q0 < tot; q0 += U * nt) {
      Nb<V> s[U];
      int c[U], ii[U], jj[U], kk[U];
      bool ok[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int par = (par0 + w.b + w.c) & 1;
        ii[u] = R.lo[0] + par + 2 * w.a;
        jj[u] = R.lo[1] + w.b;
        kk[u] = R.lo[2] + w.c;
        ok[u] = w.c < nz && ii[u] <= R.hi[0];
        c[u] = ok[u] ? w.off + par : off0;
        s[u] = load_nb<V>(src ? src : phi, rhs, c[u]);
        w.next();
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (!ok[u]) continue;
        Nb<V> v = s[u];
        if (src) {
          v.p = v.p * mult; v.xp = v.xp * mult; v.xm = v.xm * mult; v.yp = v.yp * mult;
          v.ym = v.ym * mult; v.zp = v.zp * mult; v.zm = v.zm * mult;
        }
        bc(v, ii[u], jj[u], kk[u]);
        const V s6 = v.xp + v.xm + v.yp + v.ym + v.zp + v.zm;
        V out;
        if (cf.lap) {
          const V lphi = s6 * dxinv;
          out = lambda * (v.r - lphi);
        } else {
          const V lphi = (s6 - V(6.0) * v.p) * dxinv;
          const V helmop = alpha * v.p + beta * lphi;
          out = v.p + lambda * (helmop - v.r);
        }
        phi[c[u]] = out;
      }
    }
    if (src) {
      // the other colour: phi = src * mult
      const int par1 = par0 + 1;
      w.init(tid, nt, nxh, ny, off0, 2, R.sy, R.sz);
      for (int q0 = tid; q0 < tot; q0 += U * nt) {
        V s[U];
        int c[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
          const int par = (par1 + w.b + w.c) & 1;
          ok[u] = w.c < nz && par + 2 * w.a < nx;
          c[u] = ok[u] ? w.off + par : off0;
          s[u] = src[c[u]];
          w.next();
        }
#pragma unroll
        for (int u = 0; u < U; u++)
          if (ok[u]) phi[c[u]] = s[u] * mult;
      }
    }
    __syncthreads();
  }
  template <int U, class V>
  __device__ void relax(V* phi, const V* rhs, int n) const {
    for (int it = 0; it < n; it++) {
      gsrb<U>(phi, rhs, 0);
      gsrb<U>(phi, rhs, 1);
    }
  }
  // every valid cell: ld(c) (loads only) for U cells, then st(c, i, j, k, loaded)
  template <int U, class Ld, class St>
  __device__ void sweep(Ld ld, St st) const {
    using V = decltype(ld(0));
    Walk w;
    w.init(tid, nt, nx, ny, off0, 1, R.sy, R.sz);
    for (int q0 = tid; q0 < ncell; q0 += U * nt) {
      V v[U];
      int c[U], ii[U], jj[U], kk[U];
      bool ok[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        ok[u] = w.c < nz;
        c[u] = ok[u] ? w.off : off0;
        ii[u] = R.lo[0] + w.a;
        jj[u] = R.lo[1] + w.b;
        kk[u] = R.lo[2] + w.c;
        v[u] = ld(c[u]);
        w.next();
      }
#pragma unroll
      for (int u = 0; u < U; u++)
        if (ok[u]) st(c[u], ii[u], jj[u], kk[u], v[u]);
    }
    __syncthreads();
  }
};

// One chain of a reduction: len terms at the offsets of a walk over extents
// (ea, eb, *) with strides (s0, s1, s2) from base.
struct Chain {
  int len, base, ea, eb, s0, s1, s2;
};

// Serial sums of nch chains (nch <= blockDim.x): chain c's terms (x[o] * y[o] at
// its offsets o, or x[o] * x[o] with SQUARE) are added in order (with firstAssign
// the first term is assigned, otherwise added to 0.0), exactly as a serial loop
// adds them.  The terms of a stage are formed by the whole workgroup into LDS (the
// blockDim.x / nch threads of chain c walk it, kFillSteps terms each per stage, all
// loads issued together; term n of the stage at c * cst + n, cst odd so that the
// chains start in different banks); then the owner of each chain (thread c) adds
// its part of the stage in order.  Leaves chain c's sum in out[c].  The stage
// holds kFillSteps * blockDim.x + nch terms.
constexpr int kFillSteps = 14;
template <bool SQUARE, class Geo, class Ptrs>
__device__ void chain_sums(double* stage, double* out, int nch, int maxlen, bool firstAssign, Geo geo, Ptrs ptrs) {
  constexpr int K = 16, S = kFillSteps;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int per = nt / nch;  // walkers per chain
  const int seg = S * per;   // terms per chain and stage
  const int cst = seg | 1;
  const int myc = tid % nch;
  const bool filler = tid < per * nch;
  const Chain g = geo(myc);
  const double *px, *py;
  ptrs(myc, px, py);
  Walk w;
  w.init(tid / nch, per, g.ea, g.eb, g.base, g.s0, g.s1, g.s2);
  int n = tid / nch;  // term index of the walk
  const int ownlen = tid < nch ? geo(tid).len : 0;
  double acc = 0.0;
  for (int c0 = 0; c0 < maxlen; c0 += seg) {
    if (filler) {
      double x[S], y[S];
      int slot[S];
#pragma unroll
      for (int u = 0; u < S; u++) {
        const int o = n < g.len ? w.off : g.base;
        x[u] = px[o];
        if (!SQUARE) y[u] = py[o];
        slot[u] = myc * cst + (n - c0);  // terms past the chain's end land in slots nobody reads
        w.next();
        n += per;
      }
#pragma unroll
      for (int u = 0; u < S; u++) stage[slot[u]] = SQUARE ? x[u] * x[u] : x[u] * y[u];
    }
    __syncthreads();
    if (tid < nch) {
      const double* st = stage + tid * cst;
      const int lim = min(seg, ownlen - c0);
      int m = 0;
      if (firstAssign && c0 == 0 && lim > 0) acc = st[m++];
      for (; m + 2 * K <= lim; m += 2 * K) {
        double a[K], b[K];
#pragma unroll
        for (int u = 0; u < K; u++) a[u] = st[m + u];
#pragma unroll
        for (int u = 0; u < K; u++) b[u] = st[m + K + u];
#pragma unroll
        for (int u = 0; u < K; u++) acc += a[u];
#pragma unroll
        for (int u = 0; u < K; u++) acc += b[u];
      }
      for (; m < lim; m++) acc += st[m];
    }
    __syncthreads();
  }
  if (tid < nch) out[tid] = acc;
  __syncthreads();
}

// the bottom level's boxes (offsets from the rectangle's base), the first
// kBoxCache of them in LDS
constexpr int kBoxCache = 64;
struct BoxSet {
  const DBox* g;    // all boxes (global)
  const DBox* s;    // the first kBoxCache of them (LDS)
  int n;
  short2* colChain; // norm chains: (box, column) per chain (LDS)
  int ncol;         // number of norm chains
  int maxCells, maxCol;
  __device__ DBox get(int q) const { return q < kBoxCache ? s[q] : g[q]; }
};

// FArrayBox::dotProduct per box (ForAllThisCBNNXC: Fortran order, serial from
// 0.0), the box results added in layout order from 0.0.  NPAIR dot products
// (a[m], b[m]) are formed in one pass, each with its own chains.
template <int NPAIR>
__device__ void block_dots(double* stage, double* sh, const BoxSet& B, const double* const (&a)[NPAIR],
                           const double* const (&b)[NPAIR], double (&res)[NPAIR]) {
  chain_sums<false>(
      stage, sh, NPAIR * B.n, B.maxCells, false,
      [&](int c) {
        const DBox bx = B.get(c % B.n);
        const int nx = bx.hi[0] - bx.lo[0] + 1, ny = bx.hi[1] - bx.lo[1] + 1, nz = bx.hi[2] - bx.lo[2] + 1;
        return Chain{nx * ny * nz, (int)bx.off + 1 + bx.sy + bx.sz, nx, ny, 1, bx.sy, bx.sz};
      },
      [&](int c, const double*& x, const double*& y) {
        const int m = c / B.n;
        x = a[0];
        y = b[0];
#pragma unroll
        for (int q = 1; q < NPAIR; q++)
          if (m == q) { x = a[q]; y = b[q]; }
      });
  __shared__ double sres[NPAIR];
  if (threadIdx.x == 0) {
    for (int m = 0; m < NPAIR; m++) {
      double val = 0.0;
      for (int q = 0; q < B.n; q++) val += sh[m * B.n + q];
      sres[m] = val;
    }
  }
  __syncthreads();
  for (int m = 0; m < NPAIR; m++) res[m] = sres[m];
  __syncthreads();
}

// FArrayBox::sumPow(2) per box: column sums over (k, j) (the first square assigned),
// the columns added in i order; boxes added in order from 0.0; then sqrt
// (norm(.., 2)).  Each (box, column) is one chain.
__device__ double block_norm2_boxes(double* stage, double* sh, const BoxSet& B, const double* x) {
  chain_sums<true>(
      stage, sh, B.ncol, B.maxCol, true,
      [&](int c) {
        const short2 qc = B.colChain[c];
        const DBox bx = B.get(qc.x);
        const int ny = bx.hi[1] - bx.lo[1] + 1, nz = bx.hi[2] - bx.lo[2] + 1;
        return Chain{ny * nz, (int)bx.off + 1 + qc.y + bx.sy + bx.sz, ny, nz, bx.sy, bx.sz, 0};
      },
      [&](int, const double*& px, const double*& py) {
        px = x;
        py = x;
      });
  __shared__ double normAcc;
  if (threadIdx.x == 0) {
    double acc = 0.0;
    int c = 0;
    for (int q = 0; q < B.n; q++) {
      const DBox bx = B.get(q);
      const int nx = bx.hi[0] - bx.lo[0] + 1;
      double sum = sh[c];
      for (int i = 1; i < nx; i++) sum += sh[c + i];
      c += nx;
      acc += sum;
    }
    normAcc = acc;
  }
  __syncthreads();
  const double n = normAcc;
  __syncthreads();
  return std::sqrt(n);
}

constexpr int kColumnPlanes = 8;     // planes of a column tile
constexpr int kBottomThreads = 512;  // one workgroup per bottom solve
constexpr int kStage = kFillSteps * kBottomThreads + kBottomThreads;
constexpr int kUS = 4, kUE = 8;      // cells per thread and pass: stencil / elementwise sweeps
template <class T, bool TREE>
__global__ void __launch_bounds__(kBottomThreads)
    k_batch_bottom(const BatchBottom* __restrict__ bots, const BatchCoef* __restrict__ bcoef,
                   const DBox* __restrict__ bboxes, T* const* __restrict__ corr, const T* const* __restrict__ res,
                   double* __restrict__ scratch, const long long* __restrict__ soff, const double* __restrict__ metric,
                   const int* __restrict__ active, int* __restrict__ stats, BottomArgs ba) {
  // cells per thread and pass of the stencil / elementwise sweeps (fewer cells per
  // pass, for more problems per CU, measured slower: the sweeps are bound by the
  // loads each thread keeps in flight)
  constexpr int US = kUS, UE = kUE;
  const int p = blockIdx.x;
  if (!active[p]) return;
  __shared__ double sh[kBottomThreads];
  __shared__ double stage[kStage];
  __shared__ DBox sbox[kBoxCache];
  __shared__ short2 colChain[kBottomThreads];
  __shared__ int smax[2];
  const BatchBottom bt = bots[p];
  const int tid = threadIdx.x;
  // fields are addressed from the base of the level's ghosted rectangle
  BottomLevel L;
  L.cf = bcoef[p];
  L.R = bt.rect;
  L.R.off = 0;
  L.tid = tid;
  L.nt = blockDim.x;
  L.nx = L.R.hi[0] - L.R.lo[0] + 1;
  L.ny = L.R.hi[1] - L.R.lo[1] + 1;
  L.nz = L.R.hi[2] - L.R.lo[2] + 1;
  L.ncell = L.nx * L.ny * L.nz;
  L.off0 = 1 + L.R.sy + L.R.sz;
  BoxSet B;
  B.g = bboxes + bt.box0;
  B.s = sbox;
  B.n = bt.nbox;
  B.colChain = colChain;
  for (int q = tid; q < min(B.n, kBoxCache); q += blockDim.x) sbox[q] = B.g[q];
  if (tid == 0) {
    int c = 0, mc = 0, mcol = 0;
    for (int q = 0; q < B.n; q++) {
      const DBox bx = B.g[q];
      const int nx = bx.hi[0] - bx.lo[0] + 1, nyz = (bx.hi[1] - bx.lo[1] + 1) * (bx.hi[2] - bx.lo[2] + 1);
      for (int i = 0; i < nx; i++) colChain[c++] = make_short2((short)q, (short)i);
      mc = max(mc, nx * nyz);
      mcol = max(mcol, nyz);
    }
    smax[0] = mc;
    smax[1] = mcol;
  }
  __syncthreads();
  B.maxCells = smax[0];
  B.maxCol = smax[1];
  B.ncol = 0;
  for (int q = 0; q < B.n; q++) B.ncol += B.get(q).hi[0] - B.get(q).lo[0] + 1;
  T* e32 = corr[p] + bt.rect.off;
  const T* r32 = res[p] + bt.rect.off;
  if (ba.relaxFirst) {
    // MultiGrid::cycle at the bottom: the correction is zero on entry
    L.sweep<UE>([&](int) { return 0; }, [&](int c, int, int, int, int) { e32[c] = T(0.0); });
    L.relax<US>(e32, r32, ba.nbottom);
    if (L.ncell == 1) return;  // m_bottomCells == 1: relax only
  }
  const long long vs = 15 + (long long)L.R.sz * (L.nz + 2);  // elements per scratch vector
  // the ten vectors live in LDS when they fit there (every cell a vector is read at
  // is written first in the solve, and its ghost cells are never used: the face
  // rules replace them), else in the problem's global scratch
  extern __shared__ double ldsVec[];
  double* base = (10 * vs <= ba.ldsDoubles) ? ldsVec + 15 : scratch + soff[p] + 15;  // rectangle base of the first
  double *x = base, *b = base + vs, *r = base + 2 * vs, *r_tilde = base + 3 * vs, *e = base + 4 * vs,
         *pp = base + 5 * vs, *p_tilde = base + 6 * vs, *s_tilde = base + 7 * vs, *t = base + 8 * vs,
         *v = base + 9 * vs;
  struct D2 {
    double a, b;
  };
  // PromotedBottomSolver: widen the correction and the residual
  L.sweep<UE>(
      [&](int c) { return D2{(double)e32[c], (double)r32[c]}; },
      [&](int c, int, int, int, D2 w) {
        x[c] = w.a;
        b[c] = w.b;
      });
  const BatchCoef cf = L.cf;
  // residual / applyOp of AMRPoissonOp on this level, homogeneous BC
  auto residual = [&](double* out, const double* phi, const double* rhs) {
    L.sweep<US>(
        [&](int c) { return L.load_nb<double>(phi, rhs, c); },
        [&](int c, int i, int j, int k, BottomLevel::Nb<double> s) {
          L.bc(s, i, j, k);
          const double lap = (s.xp + s.xm + s.yp + s.ym + s.zp + s.zm - 6.0 * s.p) * cf.dxinv;
          out[c] = -(cf.alpha * s.p) - cf.beta * lap + s.r;
        });
  };
  // zero(out) followed by applyOp(out, phi): applyOp writes every valid cell
  auto applyOp = [&](double* out, const double* phi) {
    L.sweep<US>(
        [&](int c) { return L.load_nb<double>(phi, nullptr, c); },
        [&](int c, int i, int j, int k, BottomLevel::Nb<double> s) {
          L.bc(s, i, j, k);
          const double lap = (s.xp + s.xm + s.yp + s.ym + s.zp + s.zm - 6.0 * s.p) * cf.dxinv;
          out[c] = cf.alpha * s.p + cf.beta * lap;
        });
  };
  // preCond (AMRPoissonOp.cpp:304-329): phi = rhs * mult, then relax(phi, rhs, 2);
  // the scaling is folded into the first colour
  const double mult = 1.0 / (cf.alpha - 2.0 * 3 * cf.beta / (cf.dx * cf.dx));
  auto preCond = [&](double* cor, const double* rv) {
    L.gsrb<US>(cor, rv, 0, rv, mult);
    L.gsrb<US>(cor, rv, 1);
    L.relax<US>(cor, rv, 1);
  };
  // lhs += xv * s
  auto incr = [&](double* lhs, const double* xv, double s) {
    L.sweep<UE>(
        [&](int c) { return D2{lhs[c], xv[c]}; },
        [&](int c, int, int, int, D2 w) { lhs[c] = w.a + w.b * s; });
  };
  // lhs1 += x1 * s1 and lhs2 += x2 * s2 in one sweep
  auto incr2 = [&](double* l1, const double* x1, double s1, double* l2, const double* x2, double s2) {
    struct D4 {
      double a, b, c, d;
    };
    L.sweep<US>(
        [&](int c) { return D4{l1[c], x1[c], l2[c], x2[c]}; },
        [&](int c, int, int, int, D4 w) {
          l1[c] = w.a + w.b * s1;
          l2[c] = w.c + w.d * s2;
        });
  };
  auto assign = [&](double* lhs, const double* rv) {
    L.sweep<UE>(
        [&](int c) { return rv[c]; }, [&](int c, int, int, int, double w) { lhs[c] = w; });
  };
  // setToZero of several vectors in one sweep
  auto zero3 = [&](double* z1, double* z2, double* z3) {
    L.sweep<UE>(
        [&](int) { return 0; },
        [&](int c, int, int, int, int) {
          z1[c] = 0.0;
          z2[c] = 0.0;
          z3[c] = 0.0;
        });
  };
  auto dot = [&](const double* a1, const double* b1) {
    const double* const aa[1] = {a1};
    const double* const bb[1] = {b1};
    double rr[1];
    block_dots<1>(stage, sh, B, aa, bb, rr);
    return rr[0];
  };
  auto norm = [&](const double* a) { return block_norm2_boxes(stage, sh, B, a); };
  // Tree order (TREE): each thread adds the terms of its cells in its walk order,
  // then the kBottomThreads partial sums are added by a fixed pairwise tree.  The
  // result depends only on the problem's own bottom level (the workgroup size is
  // fixed), so a problem solves the same alone and in any batch; it differs from the
  // reference order by rounding only.
  auto tree_sum = [&](double acc) {
    sh[tid] = acc;
    __syncthreads();
    for (int st = kBottomThreads / 2; st > 0; st >>= 1) {
      if (tid < st) sh[tid] = sh[tid] + sh[tid + st];
      __syncthreads();
    }
    const double total = sh[0];
    __syncthreads();
    return total;
  };
  auto tree_dot2 = [&](const double* a1, const double* b1, const double* a2, const double* b2, double& d1,
                       double& d2) {
    double acc1 = 0.0, acc2 = 0.0;
    struct D4 {
      double a, b, c, d;
    };
    L.sweep<US>(
        [&](int c) { return D4{a1[c], b1[c], a2[c], b2[c]}; },
        [&](int, int, int, int, D4 w) {
          acc1 += w.a * w.b;
          acc2 += w.c * w.d;
        });
    d1 = tree_sum(acc1);
    d2 = tree_sum(acc2);
  };
  auto tree_dot = [&](const double* a1, const double* b1) {
    double acc = 0.0;
    L.sweep<UE>([&](int c) { return D2{a1[c], b1[c]}; }, [&](int, int, int, int, D2 w) { acc += w.a * w.b; });
    return tree_sum(acc);
  };
  auto tree_norm = [&](const double* a) {
    double acc = 0.0;
    L.sweep<UE>([&](int c) { return a[c]; }, [&](int, int, int, int, double w) { acc += w * w; });
    return std::sqrt(tree_sum(acc));
  };
  auto rdot = [&](const double* a1, const double* b1) {
    if constexpr (TREE) return tree_dot(a1, b1);
    else return dot(a1, b1);
  };
  auto rnorm = [&](const double* a) {
    if constexpr (TREE) return tree_norm(a);
    else return norm(a);
  };

  const double eps = metric[2 * p + 1];
  const double convergenceMetric = metric[2 * p];
  int recount = 0;
  residual(r, x, b);  // setToZero(r) then residual: every valid cell is written
  assign(r_tilde, r);
  zero3(e, p_tilde, s_tilde);
  int iter = 0;
  double rho[4] = {0, 0, 0, 0}, nrm[2];
  nrm[0] = rnorm(r);
  double initial_norm = nrm[0];
  const double initial_residual = nrm[0];
  nrm[1] = nrm[0];
  double alpha[2] = {0, 0}, beta[2] = {0, 0}, omega[2] = {0, 0};
  bool init = true;
  int restarts = 0;
  bool done = false;
  if (convergenceMetric > 0) initial_norm = convergenceMetric;
  while (!done && (iter < ba.imax && nrm[0] > eps * nrm[1]) && (nrm[1] > 0)) {
    iter++;
    nrm[1] = nrm[0]; alpha[1] = alpha[0]; beta[1] = beta[0]; omega[1] = omega[0];
    rho[3] = rho[2]; rho[2] = rho[1];
    rho[1] = rdot(r_tilde, r);
    if (rho[1] == 0.0) {  // exit status 2
      incr(x, e, 1.0);
      done = true;
      break;
    }
    if (init) {
      assign(pp, r);
      init = false;
    } else {
      // scale(p, beta); incr(p, v, -beta*omega); incr(p, r, 1.0) in one sweep
      beta[1] = (rho[1] / rho[2]) * (alpha[1] / omega[1]);
      const double bw = -beta[1] * omega[1], bt1 = beta[1];
      struct D3 {
        double p, v, r;
      };
      L.sweep<US>(
          [&](int c) { return D3{pp[c], v[c], r[c]}; },
          [&](int c, int, int, int, D3 w) {
            const double s1 = w.p * bt1;
            const double s2 = s1 + w.v * bw;
            pp[c] = s2 + w.r * 1.0;
          });
    }
    preCond(p_tilde, pp);
    applyOp(v, p_tilde);
    const double m = rdot(r_tilde, v);
    alpha[0] = rho[1] / m;
    if (fabs(m) > ba.small * fabs(rho[1])) {
      incr2(r, v, -alpha[0], e, p_tilde, alpha[0]);
      nrm[0] = rnorm(r);
    } else {
      L.sweep<UE>([&](int) { return 0; }, [&](int c, int, int, int, int) { r[c] = 0.0; });
      nrm[0] = 0.0;
    }
    if (nrm[0] > eps * initial_norm && nrm[0] > ba.reps * initial_residual) {
      preCond(s_tilde, r);
      applyOp(t, s_tilde);
      double tt[2];
      if constexpr (TREE) {
        tree_dot2(t, r, t, t, tt[0], tt[1]);
      } else {
        const double* const aa[2] = {t, t};
        const double* const bb[2] = {r, t};
        block_dots<2>(stage, sh, B, aa, bb, tt);
      }
      omega[0] = tt[0] / tt[1];
      incr2(e, s_tilde, omega[0], r, t, -omega[0]);
      nrm[0] = rnorm(r);
    }
    if (nrm[0] <= eps * initial_norm || nrm[0] <= ba.reps * initial_residual) break;  // exit status 1
    if (omega[0] == 0.0 || nrm[0] > (1 - ba.hang) * nrm[1]) {
      if (recount == 0) {
        recount = 1;
      } else {
        recount = 0;
        incr(x, e, 1.0);
        if (restarts == ba.numRestarts) {  // exit status 3
          done = true;
          break;
        }
        residual(r, x, b);
        nrm[0] = rnorm(r);
        rho[1] = 0.0; rho[2] = 0.0; rho[3] = 0.0;
        alpha[0] = 0; beta[0] = 0; omega[0] = 0;
        assign(r_tilde, r);
        L.sweep<UE>([&](int) { return 0; }, [&](int c, int, int, int, int) { e[c] = 0.0; });
        restarts++;
        init = true;
      }
    }
  }
  if (!done) incr(x, e, 1.0);
  if (tid == 0) stats[p] += iter;
  // round the correction back into the V-cycle's precision
  L.sweep<UE>([&](int c) { return x[c]; }, [&](int c, int, int, int, double w) { e32[c] = (T)w; });
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void BatchSolver::alloc(void** p, size_t bytes) {
  dev_alloc(m_ctx, p, bytes, "hipMalloc batch");
  if (*p) hip_check(hipMemsetAsync(*p, 0, bytes, m_ctx->stream), "hipMemset batch");
  m_bytes += (long long)bytes;
}

template <class V>
static V* upload(Context* c, const std::vector<V>& h, long long* bytes) {
  if (h.empty()) return nullptr;
  void* p = nullptr;
  dev_alloc(c, &p, sizeof(V) * h.size(), "hipMalloc batch table");
  dev_h2d(c, p, h.data(), sizeof(V) * h.size(), "copy batch table");
  *bytes += (long long)(sizeof(V) * h.size());
  return (V*)p;
}

static int mg_depths(const DBL& g) {
  int n = 1;
  for (int depth = 1;; depth++) {  // DeviceHierarchy::define / MGnewOp's coarsening rule
    if (!g.coarsenable((1 << depth) * 2)) break;
    n++;
  }
  return n;
}

long long batchBottomCells(const BatchProblemSpec& P) {
  const int nd = mg_depths(P.boxes);
  return coarsen(P.domain, 1 << (nd - 1)).npts();
}

void BatchSolver::setContext(Context* c) {
  if (c == m_ctx) return;
  if (c->device != m_ctx->device) throw std::runtime_error("batch: context on another device");
  dropGraph();
  (void)hipStreamSynchronize(m_ctx->stream);
  m_ctx = c;
}

BatchSolver::BatchSolver(Context* ctx, std::vector<BatchProblemSpec> problems, int precision)
    : m_ctx(ctx), m_prec(precision), m_np((int)problems.size()) {
  if (precision != 0 && precision != 1) throw std::runtime_error("batch: precision must be 0 or 1");
  if (m_np <= 0) throw std::runtime_error("batch: no problems");
  std::vector<int> nd(m_np);
  for (int q = 0; q < m_np; q++) {
    const BatchProblemSpec& P = problems[q];
    if (P.boxes.size() == 0) throw std::runtime_error("batch: a problem has no boxes");
    if (!P.boxes.sorted()) throw std::runtime_error("batch: boxes must be sorted (Box::operator<)");
    if (P.boxes.numCells() != P.domain.npts()) throw std::runtime_error("batch: boxes must tile the domain");
    for (auto& b : P.boxes.boxes)
      if (!P.domain.contains(b.lo) || !P.domain.contains(b.hi)) throw std::runtime_error("batch: box outside domain");
    nd[q] = mg_depths(P.boxes);
  }
  // slots: problems by depth count, descending (stable)
  std::vector<int> order(m_np);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nd[a] > nd[b]; });
  m_slot.assign(m_np, 0);
  for (int s = 0; s < m_np; s++) {
    m_prob.push_back(problems[order[s]]);
    m_ndepth.push_back(nd[order[s]]);
    m_slot[order[s]] = s;
  }
  const int D = m_ndepth[0];
  m_depth.resize(D);
  for (int d = 0; d < D; d++) {
    Depth& L = m_depth[d];
    while (L.np < m_np && m_ndepth[L.np] > d) L.np++;
    long long off = 0;
    L.tileStart.assign(1, 0);
    L.wStart.assign(1, 0);
    L.cellStart.assign(1, 0);
    std::vector<Tile> tiles, ctiles;
    for (int p = 0; p < L.np; p++) {
      const BatchProblemSpec& P = m_prob[p];
      const Box r = coarsen(P.domain, 1 << d);
      DBox x;
      for (int q = 0; q < 3; q++) { x.lo[q] = r.lo[q]; x.hi[q] = r.hi[q]; }
      x.sy = (r.len(0) + 2 + 15) / 16 * 16;
      const long long sz = (long long)x.sy * (r.len(1) + 2);
      if (sz > INT32_MAX) throw std::runtime_error("batch: problem plane too large");
      x.sz = (int)sz;
      x.off = off + 15;  // first valid cell of every row 128-byte aligned
      off += ((15 + sz * (r.len(2) + 2)) + 15) / 16 * 16;
      L.h_rect.push_back(x);
      BatchCoef c{};
      c.dx = P.dx * (double)(1 << d);
      c.dxinv = 1.0 / (c.dx * c.dx);
      double sum_b = 0.0;
      for (int q = 0; q < 3; q++) sum_b = sum_b + 2.0 * c.dxinv;
      c.lap = (P.alpha == 0.0 && P.beta == 1.0) ? 1 : 0;
      c.lambda = c.lap ? -1.0 / sum_b : -1.0 / (P.alpha - P.beta * sum_b);
      c.alpha = P.alpha;
      c.beta = P.beta;
      for (int dir = 0; dir < 3; dir++)
        for (int side = 0; side < 2; side++) {
          if (P.bc.type[side][dir] > 1)
            throw std::runtime_error("batch: quadratic Dirichlet and periodic faces are not supported "
                                     "(use the single solver)");
          c.kind[dir][side] = P.bc.type[side][dir];
          c.val[dir][side] = P.bc.val[side][dir];
        }
      L.h_coef.push_back(c);
      for (int k = r.lo[2]; k <= r.hi[2]; k++)
        for (int j = r.lo[1]; j <= r.hi[1]; j += 4)
          for (int i = r.lo[0]; i <= r.hi[0]; i += 64) tiles.push_back(Tile{p, i, j, k});
      L.tileStart.push_back((int)tiles.size());
      if (d == 0) {
        if (L.ctileStart.empty()) L.ctileStart.push_back(0);
        for (int k = r.lo[2]; k <= r.hi[2]; k += kColumnPlanes)
          for (int j = r.lo[1]; j <= r.hi[1]; j += 4)
            for (int i = r.lo[0]; i <= r.hi[0]; i += 64) ctiles.push_back(Tile{p, i, j, k});
        L.ctileStart.push_back((int)ctiles.size());
      }
      L.cellStart.push_back(L.cellStart.back() + r.npts());
    }
    L.slab = off;
    L.cellsPresent = L.cellStart.back();
    L.d_rect = upload(m_ctx, L.h_rect, &m_bytes);
    L.d_coef = upload(m_ctx, L.h_coef, &m_bytes);
    L.d_tiles = upload(m_ctx, tiles, &m_bytes);
    L.d_ctiles = upload(m_ctx, ctiles, &m_bytes);
    // tile addresses stay inside each problem's array (threads past the rectangle return)
    for (const std::vector<Tile>* ts : {&tiles, &ctiles})
    for (const Tile& t : *ts) {
      const DBox& x = L.h_rect[t.box];
      if (t.i0 < x.lo[0] || t.i0 > x.hi[0] || t.j0 < x.lo[1] || t.j0 > x.hi[1] || t.k0 < x.lo[2] || t.k0 > x.hi[2])
        throw std::runtime_error("batch: tile outside its problem");
    }
    const size_t es = m_prec ? sizeof(float) : sizeof(double);
    alloc(&L.corr, es * L.slab);
    alloc(&L.res, es * L.slab);
    alloc(&L.tmp, es * L.slab);
  }
  alloc((void**)&m_phi, sizeof(double) * m_depth[0].slab);
  alloc((void**)&m_rhs, sizeof(double) * m_depth[0].slab);
  alloc((void**)&m_norm, sizeof(double) * m_np);
  alloc((void**)&d_active, sizeof(int) * m_np);
  alloc((void**)&d_metric, sizeof(double) * 2 * m_np);
  alloc((void**)&d_bstats, sizeof(int) * m_np);
  if (!m_ctx->dry) hip_check(hipHostMalloc((void**)&h_norm, sizeof(double) * m_np), "hipHostMalloc norms");
  h_active.assign(m_np, 0);
  // bottoms
  std::vector<BatchCoef> bcoef;
  std::vector<void*> corrp, resp;
  long long soff = 0;
  for (int p = 0; p < m_np; p++) {
    const int db = m_ndepth[p] - 1;
    const Depth& L = m_depth[db];
    BatchBottom b{};
    b.rect = L.h_rect[p];
    b.depth = db;
    b.cells = L.cellStart[p + 1] - L.cellStart[p];
    b.box0 = (int)h_bbox.size();
    const DBL g = m_prob[p].boxes.coarsened(1 << db);
    b.nbox = g.size();
    // the bottom's reductions give each chain (box of a dot product, column of a
    // norm) one thread: two dot products share a pass
    if (2 * b.nbox > kBottomThreads) throw std::runtime_error("batch: too many bottom boxes");
    long long widths = 0;
    for (auto& bx : g.boxes) widths += bx.len(0);
    if (widths > kBottomThreads) throw std::runtime_error("batch: the bottom boxes' widths add up to too many columns");
    for (auto& bx : g.boxes) {  // addressed in the problem's scratch vectors (rectangle at offset 15)
      DBox x = b.rect;
      for (int q = 0; q < 3; q++) { x.lo[q] = bx.lo[q]; x.hi[q] = bx.hi[q]; }
      x.off = (bx.lo[0] - b.rect.lo[0]) + (long long)(bx.lo[1] - b.rect.lo[1]) * b.rect.sy +
              (long long)(bx.lo[2] - b.rect.lo[2]) * b.rect.sz;
      h_bbox.push_back(x);
    }
    h_bottom.push_back(b);
    bcoef.push_back(L.h_coef[p]);
    const size_t es = m_prec ? sizeof(float) : sizeof(double);
    corrp.push_back(L.corr);
    resp.push_back(L.res);
    (void)es;
    // scratch: 10 FP64 vectors of the same ghosted shape, each addressed with the
    // bottom rectangle's own offsets
    const long long vs = 15 + (long long)b.rect.sz * (b.rect.hi[2] - b.rect.lo[2] + 3);
    h_bscratchOff.push_back(soff);
    soff += 10 * vs;
  }
  d_bottom = upload(m_ctx, h_bottom, &m_bytes);
  d_bbox = upload(m_ctx, h_bbox, &m_bytes);
  m_bcoef = upload(m_ctx, bcoef, &m_bytes);
  d_corrPtr = (void**)upload(m_ctx, corrp, &m_bytes);
  d_resPtr = (void**)upload(m_ctx, resp, &m_bytes);
  d_bscratchOff = upload(m_ctx, h_bscratchOff, &m_bytes);
  alloc((void**)&d_bscratch, sizeof(double) * soff);
  for (int d = 0; d < D; d++) setRelaxShape(d, 0, 0);
  m_iter.assign(m_np, 0);
  m_exit.assign(m_np, -1);
  m_hist.assign(m_np, {});
  if (!m_ctx->dry) m_ctx->sync();
}

BatchSolver::~BatchSolver() {
  if (m_ctx->dry) return;
  dropGraph();
  (void)hipStreamSynchronize(m_ctx->stream);
  auto f = [&](void* p) { if (p) (void)hipFree(p); };
  for (auto& L : m_depth) {
    f(L.d_rect); f(L.d_coef); f(L.d_tiles); f(L.d_ctiles); f(L.d_items); f(L.d_wStart); f(L.corr); f(L.res); f(L.tmp);
  }
  f(m_phi); f(m_rhs); f(m_norm); f(d_active); f(d_metric); f(d_bstats); f(d_bottom); f(d_bbox); f(m_bcoef);
  f(d_corrPtr); f(d_resPtr); f(d_bscratchOff); f(d_bscratch);
  if (h_norm) (void)hipHostFree(h_norm);
}

long long BatchSolver::cells(int q) const { return m_prob[m_slot[q]].domain.npts(); }

// whether tile shape `shape` applies to every problem of depth d (the relaxation
// runs in FP32 for a mixed batch)
bool BatchSolver::shapeApplies(int d, int shape) const {
  const Depth& L = m_depth[d];
  for (int p = 0; p < L.np; p++)
    if (!ops::fused_shape_ok(shape, m_prec ? 4 : 8, L.h_rect[p].hi[0] - L.h_rect[p].lo[0] + 1)) return false;
  return true;
}

void BatchSolver::setRelaxShape(int d, int shape, int kc) {
  // the work tables are built and checked before anything is released, so a
  // rejected shape leaves the depth as it was
  if (d < 0 || d >= (int)m_depth.size()) throw std::runtime_error("batch: bad depth");
  if (shape < 0 || shape >= ops::kFusedShapeCount) throw std::runtime_error("batch: bad tile shape");
  if (!shapeApplies(d, shape))
    throw std::runtime_error("batch: x-quad tile shapes need mixed precision and x extents that are multiples of 4");
  Depth& L = m_depth[d];
  const int FX = ops::kFusedShapeTable[shape][0], FY = ops::kFusedShapeTable[shape][1];
  int k = kc;
  if (k <= 0) {  // long enough chunks to amortise the prologue, enough work items to fill the chip
    k = 32;
    while (k > 4) {
      long long n = 0;
      for (int p = 0; p < L.np; p++) {
        const DBox& R = L.h_rect[p];
        n += (long long)((R.hi[0] - R.lo[0] + FX) / FX) * ((R.hi[1] - R.lo[1] + FY) / FY) *
             ((R.hi[2] - R.lo[2] + k) / k);
      }
      if (n >= 2048) break;
      k >>= 1;
    }
  }
  std::vector<ops::FusedArgs> items;
  std::vector<int> wStart(1, 0);
  for (int p = 0; p < L.np; p++) {
    const DBox& R = L.h_rect[p];
    ops::FusedArgs a{};
    for (int q = 0; q < 3; q++) { a.lo[q] = R.lo[q]; a.hi[q] = R.hi[q]; }
    a.off = R.off;
    a.sy = R.sy;
    a.sz = R.sz;
    a.ntx = (R.hi[0] - R.lo[0] + FX) / FX;
    a.nty = (R.hi[1] - R.lo[1] + FY) / FY;
    a.kc = k;
    const long long nw = (long long)a.ntx * a.nty * ((R.hi[2] - R.lo[2] + k) / k);
    if (nw + wStart.back() > INT32_MAX - 8) throw std::runtime_error("batch: too many work items");
    a.nwork = (int)nw;
    for (int dir = 0; dir < 3; dir++)
      for (int side = 0; side < 2; side++) a.kind[dir][side] = L.h_coef[p].kind[dir][side];
    a.zsrc = 0;
    ops::check_pair_rows(a);
    if (ops::kFusedShapeTable[shape][3] == 2) ops::check_quad_rows(a);
    items.push_back(a);
    wStart.push_back(wStart.back() + a.nwork);
  }
  dropGraph();  // the captured cycle holds the old work tables and shape
  if (L.d_items) { dev_free(m_ctx, L.d_items); L.d_items = nullptr; }
  if (L.d_wStart) { dev_free(m_ctx, L.d_wStart); L.d_wStart = nullptr; }
  L.shape = shape;
  L.kc = kc;
  L.h_items = std::move(items);
  L.wStart = std::move(wStart);
  L.d_items = upload(m_ctx, L.h_items, &m_bytes);
  L.d_wStart = upload(m_ctx, L.wStart, &m_bytes);
}

void BatchSolver::setSolverParameters(int pre, int post, int bottom, int numMG, int iterMax, int imin, double eps,
                                      double hang, double normThresh) {
  if (numMG != 1) throw std::runtime_error("batch: V-cycles only (numMG = 1)");
  dropGraph();  // the captured cycle holds the relaxation counts
  m_pre = pre; m_post = post; m_nbottom = bottom; m_numMG = numMG;
  m_iterMax = iterMax; m_imin = imin; m_eps = eps; m_hang = hang; m_normThresh = normThresh;
}

void BatchSolver::setRhs(int q, const double* h) {
  const int p = m_slot.at(q);
  const Depth& L = m_depth[0];
  const DBox& R = L.h_rect[p];
  const BatchProblemSpec& P = m_prob[p];
  // valid-packed: boxes in layout order, Fortran order within a box
  const long long n = 15 + (long long)R.sz * (R.hi[2] - R.lo[2] + 3);  // the problem's array
  std::vector<double> buf((size_t)n, 0.0);
  const long long base = R.off - 15;
  long long o = 0;
  for (auto& b : P.boxes.boxes)
    for (int k = b.lo[2]; k <= b.hi[2]; k++)
      for (int j = b.lo[1]; j <= b.hi[1]; j++)
        for (int i = b.lo[0]; i <= b.hi[0]; i++) {
          const long long c = R.off + (i - R.lo[0] + 1) + (long long)(j - R.lo[1] + 1) * R.sy +
                              (long long)(k - R.lo[2] + 1) * R.sz - base;
          buf[(size_t)c] = h[o++];
        }
  hip_check(hipMemcpy(m_rhs + base, buf.data(), sizeof(double) * n, hipMemcpyHostToDevice), "batch rhs upload");
}

void BatchSolver::getPhi(int q, double* h) const {
  const int p = m_slot.at(q);
  const Depth& L = m_depth[0];
  const DBox& R = L.h_rect[p];
  const BatchProblemSpec& P = m_prob[p];
  const long long base = R.off - 15;
  const long long n = 15 + (long long)R.sz * (R.hi[2] - R.lo[2] + 3);
  std::vector<double> buf((size_t)n);
  m_ctx->sync();
  hip_check(hipMemcpy(buf.data(), m_phi + base, sizeof(double) * n, hipMemcpyDeviceToHost), "batch phi download");
  long long o = 0;
  for (auto& b : P.boxes.boxes)
    for (int k = b.lo[2]; k <= b.hi[2]; k++)
      for (int j = b.lo[1]; j <= b.hi[1]; j++)
        for (int i = b.lo[0]; i <= b.hi[0]; i++) {
          const long long c = R.off + (i - R.lo[0] + 1) + (long long)(j - R.lo[1] + 1) * R.sy +
                              (long long)(k - R.lo[2] + 1) * R.sz - base;
          h[o++] = buf[(size_t)c];
        }
}

void BatchSolver::uploadActive() {
  hip_check(hipMemcpyAsync(d_active, h_active.data(), sizeof(int) * m_np, hipMemcpyHostToDevice, m_ctx->stream),
            "batch active flags");
}

#define BATCH_LAUNCH(kernel, grid, block, ...)                                                   \
  do {                                                                                           \
    if ((grid) > 0) {                                                                            \
      hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, m_ctx->stream, __VA_ARGS__);        \
      hip_check(hipGetLastError(), #kernel);                                                     \
    }                                                                                            \
  } while (0)

template <class T, bool LAP>
static void launch_batch_fused(Context* ctx, int shape, int grid, const ops::FusedArgs* items, const int* wstart,
                               int n, const BatchCoef* coef, const int* active, int zsrc, T* dst, const T* src,
                               const T* rhs) {
#define BF_CASE(q)                                                                                             \
  case q:                                                                                                        \
    hipLaunchKernelGGL((k_batch_fused<T, LAP, q>), dim3(grid), dim3(ops::kFusedShapeTable[q][2]), 0, ctx->stream, \
                       items, wstart, n, coef, active, zsrc, dst, src, rhs);                                     \
    break;
  static_assert(ops::kFusedShapeCount == 26, "batch: one case per tile shape");
  switch (shape) {
    BF_CASE(0) BF_CASE(1) BF_CASE(2) BF_CASE(3) BF_CASE(4) BF_CASE(5) BF_CASE(6) BF_CASE(7) BF_CASE(8) BF_CASE(9)
    BF_CASE(10) BF_CASE(11) BF_CASE(12) BF_CASE(13) BF_CASE(14) BF_CASE(15) BF_CASE(16) BF_CASE(17) BF_CASE(18)
    default:
      if constexpr (sizeof(T) == 4) {  // x-quad shapes (FP32)
        switch (shape) {
          BF_CASE(19) BF_CASE(20) BF_CASE(21) BF_CASE(22) BF_CASE(23) BF_CASE(24) BF_CASE(25)
          default: throw std::runtime_error("batch: bad tile shape");
        }
      } else {
        throw std::runtime_error("batch: bad tile shape for FP64");
      }
  }
#undef BF_CASE
  hip_check(hipGetLastError(), "k_batch_fused");
}

// relaxations of depth d for its first np problems; the first from a zero
// correction if fromZero.  Iterations alternate corr -> tmp -> corr (no buffer
// swap: the problems of the depth that are not relaxed keep their data in corr);
// an odd count ends with a copy back.
template <class T>
__global__ void __launch_bounds__(256)
    k_batch_copy(const Tile* __restrict__ tiles, int lx, int ly, const DBox* __restrict__ rect,
                 const int* __restrict__ active, T* __restrict__ dst, const T* __restrict__ src) {
  const Tile t = tiles[blockIdx.x];
  if (!active[t.box]) return;
  const DBox R = rect[t.box];
  int i, j, k;
  if (!btile_cell(t, R, lx, ly, i, j, k)) return;
  const long long c = ridx(R, i, j, k);
  dst[c] = src[c];
}
template <class T>
void BatchSolver::relaxDepth(int d, int np, int iters, bool fromZero) {
  Depth& L = m_depth[d];
  if (np <= 0 || iters <= 0) return;
  const int nw = L.wStart[np];
  const int grid = (nw + 7) / 8 * 8;
  bool anyLap = false, anyHelm = false;
  for (int p = 0; p < np; p++) (L.h_coef[p].lap ? anyLap : anyHelm) = true;
  double cells = 0;
  for (int p = 0; p < np; p++) cells += h_active[p] ? (double)(L.cellStart[p + 1] - L.cellStart[p]) : 0.0;
  for (int it = 0; it < iters; it++) {
    const int zsrc = (fromZero && it == 0) ? 1 : 0;
    T* src = (T*)((it & 1) ? L.tmp : L.corr);
    T* dst = (T*)((it & 1) ? L.corr : L.tmp);
    ops::check_pair_pointer(src, "batch relaxation");
    ops::check_pair_pointer(dst, "batch relaxation");
    hipEvent_t ev = profile_begin(m_ctx, KC_GSRB);
    if (anyLap)
      launch_batch_fused<T, true>(m_ctx, L.shape, grid, L.d_items, L.d_wStart, np, L.d_coef, d_active, zsrc, dst, src,
                                  (const T*)L.res);
    if (anyHelm)
      launch_batch_fused<T, false>(m_ctx, L.shape, grid, L.d_items, L.d_wStart, np, L.d_coef, d_active, zsrc, dst,
                                   src, (const T*)L.res);
    // algorithmic bytes: read phi (not when it is known zero), read rhs, write phi
    profile_end(m_ctx, ev, KC_GSRB, (zsrc ? 16.0 : 24.0) * sizeof(T) / 8.0 * cells, cells);
  }
  if (iters & 1)
    BATCH_LAUNCH(k_batch_copy<T>, L.tileStart[np], 256, L.d_tiles, 6, 2, L.d_rect, d_active, (T*)L.corr,
                 (const T*)L.tmp);
}

// MultiGrid::cycle in lockstep over the problems: descent (relax, restrict) at
// each depth for the problems that go deeper, every problem's bottom in one
// launch, then ascent (prolong, relax)
template <class T>
void BatchSolver::vcycle() {
  const int D = (int)m_depth.size();
  for (int d = 0; d + 1 < D; d++) {
    const int np = m_depth[d + 1].np;  // problems with a depth below d
    relaxDepth<T>(d, np, m_pre, true);
    Depth& F = m_depth[d];
    Depth& C = m_depth[d + 1];
    BATCH_LAUNCH(k_batch_restrict<T>, C.tileStart[np], 256, C.d_tiles, 6, 2, C.d_rect, F.d_rect, F.d_coef, d_active,
                 (T*)C.res, (const T*)F.corr, (const T*)F.res);
  }
  // Krylov vectors in global scratch: a batch's bottom time is set by its largest
  // bottoms, which do not fit in LDS (measured no faster with the small ones there)
  BottomArgs ba{80, 1e-12, 1e-8, 1e-30, 5, m_nbottom, 1, 0};
  if (m_treeOrder)
    BATCH_LAUNCH((k_batch_bottom<T, true>), m_np, kBottomThreads, d_bottom, m_bcoef, d_bbox, (T* const*)d_corrPtr,
                 (const T* const*)d_resPtr, d_bscratch, d_bscratchOff, d_metric, d_active, d_bstats, ba);
  else
    BATCH_LAUNCH((k_batch_bottom<T, false>), m_np, kBottomThreads, d_bottom, m_bcoef, d_bbox, (T* const*)d_corrPtr,
                 (const T* const*)d_resPtr, d_bscratch, d_bscratchOff, d_metric, d_active, d_bstats, ba);
  for (int d = D - 2; d >= 0; d--) {
    const int np = m_depth[d + 1].np;
    Depth& F = m_depth[d];
    Depth& C = m_depth[d + 1];
    BATCH_LAUNCH(k_batch_prolong<T>, F.tileStart[np], 256, F.d_tiles, 6, 2, F.d_rect, C.d_rect, d_active, (T*)F.corr,
                 (const T*)C.corr);
    relaxDepth<T>(d, np, m_post, false);
  }
}

// cell-relaxations of one V-cycle (pre and post relaxation at every depth above
// the bottom, for the active problems that go below it)
double BatchSolver::cycleUnits() const {
  double units = 0;
  for (size_t d = 0; d + 1 < m_depth.size(); d++) {
    const Depth& L = m_depth[d];
    const int np = m_depth[d + 1].np;
    double cells = 0;
    for (int p = 0; p < np; p++) cells += h_active[p] ? (double)(L.cellStart[p + 1] - L.cellStart[p]) : 0.0;
    units += (m_pre + m_post) * cells;
  }
  return units;
}

// One V-cycle.  With the graph enabled (and the profiler off), the cycle's launches
// are captured once into a HIP graph and every later cycle replays it: the launch
// sequence, grids and arguments are the same for every cycle of every solve (the
// problems that have left the lockstep are masked on the device), so the replay
// runs exactly the captured work.  The cycle holds no memset or copy.
template <class T>
void BatchSolver::cycle() {
  if (!m_useGraph || m_ctx->prof) {
    vcycle<T>();
    return;
  }
  if (!m_graph) {
    hipGraph_t g = nullptr;
    hip_check(hipStreamBeginCapture(m_ctx->stream, hipStreamCaptureModeThreadLocal), "batch: begin capture");
    try {
      vcycle<T>();
    } catch (...) {
      (void)hipStreamEndCapture(m_ctx->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    hip_check(hipStreamEndCapture(m_ctx->stream, &g), "batch: end capture");
    const hipError_t e = hipGraphInstantiate(&m_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    hip_check(e, "batch: graph instantiate");
  }
  hip_check(hipGraphLaunch(m_graph, m_ctx->stream), "batch: graph launch");
  m_graphLaunches++;
}

void BatchSolver::dropGraph() {
  if (m_graph) {
    (void)hipStreamSynchronize(m_ctx->stream);
    (void)hipGraphExecDestroy(m_graph);
    m_graph = nullptr;
  }
}

void BatchSolver::setGraph(bool on) {
  if (!on) dropGraph();
  m_useGraph = on;
}

void BatchSolver::setTreeOrder(bool on) {
  if (on != m_treeOrder) dropGraph();  // the captured cycle holds the bottom's arguments
  m_treeOrder = on;
}

template <class T>
void BatchSolver::residualNorm(bool homog) {
  (void)homog;
  Depth& L = m_depth[0];
  BATCH_LAUNCH(k_batch_zero_norms, (m_np + 255) / 256, 256, m_norm, d_active, m_np);
  double cells = 0;
  for (int p = 0; p < m_np; p++) cells += h_active[p] ? (double)(L.cellStart[p + 1] - L.cellStart[p]) : 0.0;
  hipEvent_t ev = profile_begin(m_ctx, KC_RESIDUAL);
  const int nct = L.ctileStart[m_np];
  BATCH_LAUNCH(k_batch_residual_norm_z<T>, (nct + 7) / 8 * 8, 256, L.d_ctiles, nct, kColumnPlanes, L.d_rect,
               L.d_coef, d_active, (T*)L.res, (const double*)m_phi, (const double*)m_rhs, m_norm);
  profile_end(m_ctx, ev, KC_RESIDUAL, (16.0 + sizeof(T)) * cells, cells);
  hip_check(hipMemcpyAsync(h_norm, m_norm, sizeof(double) * m_np, hipMemcpyDeviceToHost, m_ctx->stream),
            "batch norms");
  m_ctx->sync();
}

template <class T>
void BatchSolver::update() {
  Depth& L = m_depth[0];
  const int nct = L.ctileStart[m_np];
  BATCH_LAUNCH(k_batch_update_z<T>, (nct + 7) / 8 * 8, 256, L.d_ctiles, nct, kColumnPlanes, L.d_rect, d_active,
               m_phi, (const T*)L.corr);
}

// solveNoInitResid (AMRMultiGrid.H:935-1124) for every problem at once: each
// problem keeps its own history and convergence tests and leaves the lockstep
// when they stop it
template <class T>
void BatchSolver::runSolve() {
  std::fill(h_active.begin(), h_active.end(), 1);
  uploadActive();
  hip_check(hipMemsetAsync(m_phi, 0, sizeof(double) * m_depth[0].slab, m_ctx->stream), "batch phi zero");
  hip_check(hipMemsetAsync(d_bstats, 0, sizeof(int) * m_np, m_ctx->stream), "batch stats");
  m_cellRelax = 0;
  residualNorm<T>(false);
  std::vector<double> r0(m_np), rnorm(m_np), normLast(m_np);
  std::vector<double> metric(2 * m_np);
  for (int p = 0; p < m_np; p++) {
    r0[p] = rnorm[p] = h_norm[p];
    normLast[p] = 2 * r0[p];
    m_iter[p] = 0;
    m_hist[p].assign(1, r0[p]);
    metric[2 * p] = r0[p];
    metric[2 * p + 1] = 1.0 * m_eps;  // m_bottomSolverEpsCushion * m_eps
  }
  hip_check(hipMemcpyAsync(d_metric, metric.data(), sizeof(double) * 2 * m_np, hipMemcpyHostToDevice, m_ctx->stream),
            "batch metrics");
  auto keepGoing = [&](int p, bool& goNorm, bool& goRedu, bool& goIter, bool& goHang) {
    goNorm = rnorm[p] > m_normThresh;
    goRedu = rnorm[p] > m_eps * r0[p];
    goIter = m_iter[p] < m_iterMax;
    goHang = m_iter[p] < m_imin || rnorm[p] < (1 - m_hang) * normLast[p];
    const bool goMin = m_iter[p] < m_iterMin;
    return goMin || (goIter && goRedu && goHang && goNorm);
  };
  int nactive = 0;
  for (int p = 0; p < m_np; p++) {
    bool a, b, c, d;
    h_active[p] = keepGoing(p, a, b, c, d) ? 1 : 0;
    if (!h_active[p]) m_exit[p] = int(!b) + int(!c) * 2 + int(!d) * 4 + int(!a) * 8;
    nactive += h_active[p];
  }
  while (nactive > 0) {
    uploadActive();
    for (int p = 0; p < m_np; p++)
      if (h_active[p]) normLast[p] = rnorm[p];
    cycle<T>();
    m_cellRelax += cycleUnits();
    update<T>();
    residualNorm<T>(false);
    nactive = 0;
    for (int p = 0; p < m_np; p++) {
      if (!h_active[p]) continue;
      rnorm[p] = h_norm[p];
      m_iter[p]++;
      m_hist[p].push_back(rnorm[p]);
      bool goNorm, goRedu, goIter, goHang;
      if (keepGoing(p, goNorm, goRedu, goIter, goHang)) {
        nactive++;
      } else {
        h_active[p] = 0;
        m_exit[p] = int(!goRedu) + int(!goIter) * 2 + int(!goHang) * 4 + int(!goNorm) * 8;
      }
    }
  }
  std::vector<int> st(m_np);
  hip_check(hipMemcpy(st.data(), d_bstats, sizeof(int) * m_np, hipMemcpyDeviceToHost), "batch stats");
  m_bottomIters = 0;
  for (int v : st) m_bottomIters += v;
}

void BatchSolver::solve() {
  if (m_prec) runSolve<float>();
  else runSolve<double>();
}

std::vector<double> BatchSolver::tuneRelax() {
  dropGraph();
  std::vector<double> best(m_depth.size(), 0.0);
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  std::fill(h_active.begin(), h_active.end(), 1);
  uploadActive();
  const int kcs[4] = {0, 8, 16, 32};
  for (size_t d = 0; d < m_depth.size(); d++) {
    const int np = m_depth[d].np;
    double bt = 1e300;
    int bs = 0, bk = 0;
    for (int sh = 0; sh < ops::kFusedShapeCount; sh++)
      for (int q = 0; q < 4; q++) {
        if (!shapeApplies((int)d, sh)) continue;
        setRelaxShape((int)d, sh, kcs[q]);
        auto run = [&]() {
          if (m_prec) relaxDepth<float>((int)d, np, 1, false);
          else relaxDepth<double>((int)d, np, 1, false);
        };
        run();
        hip_check(hipEventRecord(e0, m_ctx->stream), "record");
        for (int r = 0; r < 3; r++) run();
        hip_check(hipEventRecord(e1, m_ctx->stream), "record");
        hip_check(hipEventSynchronize(e1), "sync");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        if (ms / 3 < bt) { bt = ms / 3; bs = sh; bk = kcs[q]; }
      }
    setRelaxShape((int)d, bs, bk);
    best[d] = bt;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  m_cellRelax = 0;
  return best;
}

// ---------------------------------------------------------------------------
// DeviceBottomSolver: the bottom BiCGStab of one hierarchy's multigrid as one
// workgroup on the device (the batched bottom kernel on a one-entry table)
// ---------------------------------------------------------------------------
template <class T>
DeviceBottomSolver<T>::~DeviceBottomSolver() {
  release();
}
template <class T>
void DeviceBottomSolver<T>::release() {
  if (m_ctx->dry) return;
  if (m_ctx->stream) (void)hipStreamSynchronize(m_ctx->stream);
  void* ps[] = {d_bottom, d_bcoef, d_bbox, d_ptrs, d_scratch, d_soff, d_metric, d_active, d_stats};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  if (h_stage) (void)hipHostFree(h_stage);
  d_bottom = nullptr; d_bcoef = nullptr; d_bbox = nullptr; d_ptrs = nullptr; d_scratch = nullptr;
  d_soff = nullptr; d_metric = nullptr; d_active = nullptr; d_stats = nullptr; h_stage = nullptr;
}
// the operator's coefficients as the bottom kernel reads them (face rules are set by define)
template <class T>
static BatchCoef bottom_coef(const AMRPoissonOp<T>* ap) {
  BatchCoef c{};
  c.dx = ap->m_dx;
  c.dxinv = 1.0 / (c.dx * c.dx);
  double sum_b = 0.0;
  for (int q = 0; q < 3; q++) sum_b = sum_b + 2.0 * c.dxinv;
  c.lap = (ap->m_alpha == 0.0 && ap->m_beta == 1.0) ? 1 : 0;
  c.lambda = c.lap ? -1.0 / sum_b : -1.0 / (ap->m_alpha - ap->m_beta * sum_b);
  c.alpha = ap->m_alpha;
  c.beta = ap->m_beta;
  return c;
}
template <class T>
void DeviceBottomSolver<T>::define(LinearOp<LevelData<T>>* op, bool homogeneous) {
  m_fallback.define(op, homogeneous);
  auto* ap = dynamic_cast<AMRPoissonOp<T>*>(op);
  // variable coefficients or Jacobi relaxation: the host-driven solve
  const bool vc = ap && (ap->m_vc || ap->m_relaxMode != 1);
  if (m_device && op == m_boundOp && homogeneous && !vc) return;  // MultiGrid re-binds its bottom at every solve
  m_device = false;
  m_boundOp = op;
  if (!homogeneous || !ap || vc || !ap->m_lay->merged || m_ctx->dry) return;
  release();
  const DeviceLayout& L = *ap->m_lay;
  m_lay = &L;
  BatchBottom b{};
  b.rect = L.h_boxes[L.nbox];
  b.cells = L.valid_cells;
  b.box0 = 0;
  b.nbox = L.nbox;
  // one thread per reduction chain (two dot products share a pass; a norm has a
  // chain per box column): otherwise the host-driven solve
  if (2 * b.nbox > kBottomThreads) return;
  long long widths = 0;
  for (int q = 0; q < L.nbox; q++) widths += L.h_boxes[q].hi[0] - L.h_boxes[q].lo[0] + 1;
  if (widths > kBottomThreads) return;
  std::vector<DBox> boxes;
  for (int q = 0; q < L.nbox; q++) {
    const DBox& x0 = L.h_boxes[q];
    DBox x = x0;  // addressed in the scratch vectors (rectangle at offset 15)
    x.off = (x0.lo[0] - b.rect.lo[0]) + (long long)(x0.lo[1] - b.rect.lo[1]) * b.rect.sy +
            (long long)(x0.lo[2] - b.rect.lo[2]) * b.rect.sz;
    boxes.push_back(x);
  }
  BatchCoef c = bottom_coef(ap);
  // the bottom of a merged multigrid level is a whole (coarsened) domain: every
  // face is a domain face with the operator's BC type
  for (int dir = 0; dir < 3; dir++)
    for (int side = 0; side < 2; side++) {
      const int face = side ? b.rect.hi[dir] : b.rect.lo[dir];
      const int dface = side ? ap->m_domain.hi[dir] : ap->m_domain.lo[dir];
      if (face != dface) return;  // a level with coarse-fine faces keeps the host-driven solver
      if (ap->m_bc.type[side][dir] > 1) return;  // quadratic Dirichlet faces: the host-driven solver
      c.kind[dir][side] = ap->m_bc.type[side][dir];
      c.val[dir][side] = ap->m_bc.val[side][dir];
    }
  const long long vs = 15 + (long long)b.rect.sz * (b.rect.hi[2] - b.rect.lo[2] + 3);
  m_vs = vs;
  m_ap = ap;
  m_coef = c;
  long long bytes = 0;
  d_bottom = upload(m_ctx, std::vector<BatchBottom>{b}, &bytes);
  d_bcoef = upload(m_ctx, std::vector<BatchCoef>{c}, &bytes);
  d_bbox = upload(m_ctx, boxes, &bytes);
  d_soff = upload(m_ctx, std::vector<long long>{0}, &bytes);
  d_active = upload(m_ctx, std::vector<int>{1}, &bytes);
  d_stats = upload(m_ctx, std::vector<int>{0}, &bytes);
  hip_check(hipMalloc(&d_scratch, sizeof(double) * 10 * vs), "hipMalloc bottom scratch");
  hip_check(hipMalloc(&d_metric, sizeof(double) * 2), "hipMalloc bottom metric");
  hip_check(hipMalloc(&d_ptrs, sizeof(void*) * 2), "hipMalloc bottom pointers");
  hip_check(hipHostMalloc(&h_stage, 64 + sizeof(BatchCoef)), "hipHostMalloc bottom staging");
  m_device = true;
  setConvergenceMetrics(m_metric, m_eps);
}
template <class T>
void DeviceBottomSolver<T>::setConvergenceMetrics(double metric, double tol) {
  m_metric = metric;
  m_eps = tol;
  m_fallback.setConvergenceMetrics(metric, tol);
  if (!m_device) return;
  double* h = (double*)h_stage;
  (void)hipStreamSynchronize(m_ctx->stream);  // the staging buffer may still feed an earlier copy
  h[0] = metric;
  h[1] = tol;
  hip_check(hipMemcpyAsync(d_metric, h, sizeof(double) * 2, hipMemcpyHostToDevice, m_ctx->stream), "bottom metric");
}
template <class T>
void DeviceBottomSolver<T>::solve(LevelData<T>& phi, const LevelData<T>& rhs) {
  if (m_device && (static_cast<const AMRPoissonOp<T>*>(m_ap)->m_vc ||
                   static_cast<const AMRPoissonOp<T>*>(m_ap)->m_relaxMode != 1))
    throw std::runtime_error("DeviceBottomSolver: operator changed after define");
  if (!m_device || phi.nc > 1) {  // fields of several components: the host-driven solve
    m_fallback.solve(phi, rhs);
    return;
  }
  if (phi.lay != m_lay || rhs.lay != m_lay) throw std::runtime_error("DeviceBottomSolver: layout mismatch");
  // the fields' buffers move (the relaxation swaps them), so their addresses go with
  // each launch; the operator's coefficients may have been reset since define
  void** h = (void**)((char*)h_stage + 32);
  (void)hipStreamSynchronize(m_ctx->stream);
  h[0] = phi.d;
  h[1] = (void*)rhs.d;
  hip_check(hipMemcpyAsync(d_ptrs, h, sizeof(void*) * 2, hipMemcpyHostToDevice, m_ctx->stream), "bottom pointers");
  const auto* ap = static_cast<const AMRPoissonOp<T>*>(m_ap);
  if (ap->m_alpha != m_coef.alpha || ap->m_beta != m_coef.beta || ap->m_dx != m_coef.dx) {
    BatchCoef c = bottom_coef(ap);
    for (int dir = 0; dir < 3; dir++)
      for (int side = 0; side < 2; side++) {
        c.kind[dir][side] = m_coef.kind[dir][side];
        c.val[dir][side] = m_coef.val[dir][side];
      }
    m_coef = c;
    std::memcpy((char*)h_stage + 64, &m_coef, sizeof(BatchCoef));
    hip_check(hipMemcpyAsync(d_bcoef, (char*)h_stage + 64, sizeof(BatchCoef), hipMemcpyHostToDevice, m_ctx->stream),
              "bottom coefficients");
  }
  // the Krylov vectors in LDS when the level is small enough
  const int lds = (10 * m_vs <= kBottomLdsDoubles) ? (int)(10 * m_vs) : 0;
  BottomArgs ba{m_fallback.m_imax, m_fallback.m_reps, m_fallback.m_hang, m_fallback.m_small,
                m_fallback.m_numRestarts, 0, 0, lds};
  hipLaunchKernelGGL((k_batch_bottom<T, false>), dim3(1), dim3(kBottomThreads), sizeof(double) * lds, m_ctx->stream,
                     (const BatchBottom*)d_bottom,
                     (const BatchCoef*)d_bcoef, (const DBox*)d_bbox, (T* const*)d_ptrs,
                     (const T* const*)((void**)d_ptrs + 1), d_scratch, (const long long*)d_soff,
                     (const double*)d_metric, (const int*)d_active, (int*)d_stats, ba);
  hip_check(hipGetLastError(), "k_batch_bottom");
}
template <class T>
long long DeviceBottomSolver<T>::totalIterations() {
  if (!m_device) return m_fallback.m_total_iters;
  int v = 0;
  hip_check(hipMemcpyAsync(h_stage, d_stats, sizeof(int), hipMemcpyDeviceToHost, m_ctx->stream), "bottom stats");
  m_ctx->sync();
  std::memcpy(&v, h_stage, sizeof(int));
  return v;
}
template class DeviceBottomSolver<double>;
template class DeviceBottomSolver<float>;

}  // namespace amrmg
