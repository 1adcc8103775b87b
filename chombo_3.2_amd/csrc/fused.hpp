// The one-launch GSRB relaxation of a merged level (device side), shared by the
// single-hierarchy launcher (kernels.hip) and the batched solver (batch.hip).
// The derivation of its equivalence with the reference's four-launch sequence
// (ghost refresh, red, ghost refresh, black) is given at gsrb_fused in kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace amrmg {
namespace ops {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
constexpr int NSLOT = 5;  // LDS plane buffers: planes k-1 .. k+2 in use plus the one being filled
// Two x-adjacent values moved as one 8- or 16-byte access (loads, LDS, stores).
template <class T>
struct alignas(2 * sizeof(T)) Pair {
  T x, y;
};
struct FusedArgs {
  int lo[3], hi[3];   // the level's rectangle
  long long off;      // slab index of (lo-1, lo-1, lo-1)
  int sy, sz;
  int ntx, nty, kc, nwork;
  int kind[3][2];     // face rule: 0 Neumann, 1 Dirichlet, 2 homogeneous quadratic CF
  int zsrc;           // 1: the old values are all zero (src is not read)
};
template <class T>
__device__ __forceinline__ T face_ghost(int kind, T nearv, T farv, T c1, T c2) {
  // all three rules are formed and one is selected, so no LDS read sits in a branch
  const T gd = T(2) * T(0.0) - nearv;  // PK_DIRICHLET with the homogeneous value
  const T gc = c1 * nearv + c2 * farv;  // PK_CFHOMO
  return kind == 1 ? gd : (kind == 2 ? gc : nearv);  // else PK_NEUMANN, homogeneous
}
// new value of cell (i,j,k); S0 is its plane, Sm / Sp the planes below / above,
// c its LDS index, LW the LDS row length
template <class T, bool LAP, int LW>
__device__ __forceinline__ T fused_new(const FusedArgs& a, const T* Sm, const T* S0, const T* Sp, int c, int i, int j,
                                       int k, T r, T dxinv, T lambda, T alpha, T beta, T c1, T c2) {
  // the seven reads are unconditional (all inside the LDS planes); a face ghost's
  // inner operand is the opposite neighbour
  const T own = S0[c];
  const T vxp = S0[c + 1], vxm = S0[c - 1];
  const T vyp = S0[c + LW], vym = S0[c - LW];
  const T vzp = Sp[c], vzm = Sm[c];
  const T xp = (i == a.hi[0]) ? face_ghost(a.kind[0][1], own, vxm, c1, c2) : vxp;
  const T xm = (i == a.lo[0]) ? face_ghost(a.kind[0][0], own, vxp, c1, c2) : vxm;
  const T yp = (j == a.hi[1]) ? face_ghost(a.kind[1][1], own, vym, c1, c2) : vyp;
  const T ym = (j == a.lo[1]) ? face_ghost(a.kind[1][0], own, vyp, c1, c2) : vym;
  const T zp = (k == a.hi[2]) ? face_ghost(a.kind[2][1], own, vzm, c1, c2) : vzp;
  const T zm = (k == a.lo[2]) ? face_ghost(a.kind[2][0], own, vzp, c1, c2) : vzm;
  const T s6 = xp + xm + yp + ym + zp + zm;
  if (LAP) {
    const T lphi = s6 * dxinv;
    return lambda * (r - lphi);
  } else {
    const T lphi = (s6 - T(6.0) * own) * dxinv;
    const T helmop = alpha * own + beta * lphi;
    return own + lambda * (helmop - r);
  }
}
// Host-side checks of what the pair accesses of fused_column assume.
inline void check_pair_rows(const FusedArgs& a) {
  if ((a.off + 1) % 2 != 0 || a.sy % 2 != 0 || a.sz % 2 != 0)
    throw std::runtime_error("one-launch relaxation: rows must start on an even slab index");
}
template <class T>
inline void check_pair_pointer(const T* p, const char* what) {
  if (reinterpret_cast<uintptr_t>(p) % (2 * sizeof(T)) != 0)
    throw std::runtime_error(std::string(what) + ": field not aligned to a pair of values");
}
// Workgroup barrier for LDS traffic only: waits for this wave's LDS operations,
// not for its global loads, so the next planes' loads stay in flight across it.
// Global memory needs no ordering here (src and rhs are read-only, each dst cell
// is written by one thread).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// One workgroup's share of a one-launch relaxation: work item w (a column of
// FX x FY cells, FX even, over one z-chunk) of the rectangle described by a.
//
// Old values move in pairs of x-adjacent cells (i, i+1) with i - lo[0] even: one
// 8- or 16-byte load, LDS write, LDS read and store per pair.  The launchers check
// that this is aligned: the slab index of (lo[0], j, k) is even for every row
// (off odd, sy and sz even) and the field pointers are aligned to a pair.  A pair
// load clamps its first cell into [lo[0], last pair start], so it stays inside the
// row's storage (the pair that holds hi[0] may end on the ghost cell hi[0]+1).
template <class T, bool LAP, int FX, int FY, int NT>
__device__ __forceinline__ void fused_column(const FusedArgs& a, int w, T* __restrict__ dst,
                                             const T* __restrict__ src, const T* __restrict__ rhs, T dxinv,
                                             T lambda, T alpha, T beta, T c1, T c2) {
  constexpr int LW = FX + 4, LH = FY + 4, LPLANE = LW * LH;  // LDS plane: the column grown by 2
  constexpr int LW2 = LW / 2, NPAIR = LW2 * LH;                // pairs per LDS row, per plane
  constexpr int NLD = (NPAIR + NT - 1) / NT;                    // plane pairs per thread
  constexpr int SW2 = FX / 2, NSP = SW2 * FY;                   // pairs per stored row, per plane
  constexpr int NRC = (FY + 2) * (FX / 2 + 1);                // red cells of the column grown by 1
  constexpr int NRED = (NRC + NT - 1) / NT;
  constexpr int NBC = FX * FY / 2;                            // black cells of the column
  constexpr int NBLK = (NBC + NT - 1) / NT;
  constexpr int NST = (NSP + NT - 1) / NT;
  static_assert(FX % 2 == 0, "column width must be even");
  __shared__ __align__(16) T P[NSLOT][LPLANE];
  using P2 = Pair<T>;
  const int pmax = a.lo[0] + 2 * ((a.hi[0] - a.lo[0]) / 2);  // start of the pair that holds hi[0]
  const int tx = w % a.ntx, ty = (w / a.ntx) % a.nty, ch = w / (a.ntx * a.nty);
  const int x0 = a.lo[0] + tx * FX, y0 = a.lo[1] + ty * FY;
  const int k0 = a.lo[2] + ch * a.kc;
  const int k1 = min(k0 + a.kc, a.hi[2] + 1);
  const int tid = threadIdx.x;
  auto slot = [&](int p) { return (p - k0 + 2) % NSLOT; };
  auto gidx = [&](int i, int j, int k) {
    return a.off + (i - a.lo[0] + 1) + (long long)(j - a.lo[1] + 1) * a.sy + (long long)(k - a.lo[2] + 1) * a.sz;
  };
  auto lidx = [&](int i, int j) { return (j - y0 + 2) * LW + (i - x0 + 2); };
  // red cell q of plane r (on the column grown by 1, (i + j + r) even); black cell q
  // of plane k (on the column, (i + j + k) odd)
  auto red_ij = [&](int q, int r, int& i, int& j) {
    j = y0 - 1 + q / (FX / 2 + 1);
    i = x0 - 1 + ((x0 - 1 + j + r) & 1) + 2 * (q % (FX / 2 + 1));
  };
  auto black_ij = [&](int q, int k, int& i, int& j) {
    j = y0 + q / (FX / 2);
    i = x0 + ((x0 + j + k + 1) & 1) + 2 * (q % (FX / 2));
  };
  // software pipeline, two steps deep: the old values of plane k+4 and the rhs of
  // the red and the black cells of plane k+3 are loaded into registers during
  // step k, into one of two register sets used alternately (no register moves
  // between the sets, so each step waits only for the loads issued two steps
  // before).  A plane's red and black rhs values share cache lines and are loaded
  // together; the black ones are used a step after the red ones, so step k keeps
  // them in `carry` for step k+1.
  // Every load is unconditional from an address clamped into the rectangle, and
  // its value is used as loaded: a value at a position outside the rectangle (or
  // of a cell that is not updated) is never selected into a result.  A select at
  // load time would make the wave wait for the load there, ending the pipeline.
  struct Pre {
    P2 phi[NLD];
    T rr[NRED], rb[NBLK];
  };
  Pre A, B;
  auto fetch = [&](Pre& f, int p, int r, int k) {  // plane p; red cells of plane r; black cells of plane k
    const int pc = clampi(p, a.lo[2], a.hi[2]);
    if (a.zsrc) {  // uniform branch: relaxation of a field known to be zero
#pragma unroll
      for (int u = 0; u < NLD; u++) f.phi[u] = P2{T(0.0), T(0.0)};
    } else {
#pragma unroll
      for (int u = 0; u < NLD; u++) {
        const int q = tid + NT * u;
        const int i = x0 - 2 + 2 * (q % LW2), j = y0 - 2 + q / LW2;
        f.phi[u] = *reinterpret_cast<const P2*>(src + gidx(clampi(i, a.lo[0], pmax), clampi(j, a.lo[1], a.hi[1]), pc));
      }
    }
    const int rc = clampi(r, a.lo[2], a.hi[2]);
#pragma unroll
    for (int u = 0; u < NRED; u++) {
      int i, j;
      red_ij(tid + NT * u, r, i, j);
      f.rr[u] = rhs[gidx(clampi(i, a.lo[0], a.hi[0]), clampi(j, a.lo[1], a.hi[1]), rc)];
    }
    const int kc = clampi(k, a.lo[2], a.hi[2]);
#pragma unroll
    for (int u = 0; u < NBLK; u++) {
      int i, j;
      black_ij(tid + NT * u, k, i, j);
      f.rb[u] = rhs[gidx(clampi(i, a.lo[0], a.hi[0]), clampi(j, a.lo[1], a.hi[1]), kc)];
    }
  };
  auto put_phi = [&](const Pre& f, int p) {
    P2* S = reinterpret_cast<P2*>(P[slot(p)]);  // pair q of the plane: LDS values 2q, 2q+1
#pragma unroll
    for (int u = 0; u < NLD; u++) {
      const int q = tid + NT * u;
      if (q < NPAIR) S[q] = f.phi[u];
    }
  };
  auto store = [&](int p) {
    const P2* S = reinterpret_cast<const P2*>(P[slot(p)]);
    P2 v[NST];
#pragma unroll
    for (int u = 0; u < NST; u++) {
      const int q = min(tid + NT * u, NSP - 1);
      v[u] = S[(q / SW2 + 2) * LW2 + (q % SW2 + 1)];
    }
#pragma unroll
    for (int u = 0; u < NST; u++) {
      const int q = tid + NT * u;
      const int i = x0 + 2 * (q % SW2), j = y0 + q / SW2;
      if (q < NSP && i <= a.hi[0] && j <= a.hi[1]) {
        T* d = dst + gidx(i, j, p);
        if (i < a.hi[0]) *reinterpret_cast<P2*>(d) = v[u];
        else *d = v[u].x;  // the pair's second cell is the ghost cell hi[0]+1
      }
    }
  };
  // black rhs values of the plane after the current one (see the pipeline above)
  T carry[NBLK];
  // step k: plane k+2 into LDS, red on plane k+1, black on plane k, store plane k-1
  auto step = [&](Pre& f, int k) {
    put_phi(f, k + 2);
    T rr[NRED], rb[NBLK], rbn[NBLK];
#pragma unroll
    for (int u = 0; u < NRED; u++) rr[u] = f.rr[u];
#pragma unroll
    for (int u = 0; u < NBLK; u++) {
      rb[u] = carry[u];
      rbn[u] = f.rb[u];
    }
    lds_barrier();
    fetch(f, k + 4, k + 3, k + 3);
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
          P[slot(r)][c] = fused_new<T, LAP, LW>(a, P[slot(r - 1)], P[slot(r)], P[slot(r + 1)], c, i, j, r, rr[u],
                                                dxinv, lambda, alpha, beta, c1, c2);
        }
      }
    }
    lds_barrier();
    if (k >= k0) {
#pragma unroll
      for (int u = 0; u < NBLK; u++) {
        const int q = tid + NT * u;
        int i, j;
        black_ij(q, k, i, j);
        if (q < NBC && i <= a.hi[0] && j <= a.hi[1]) {
          const int c = lidx(i, j);
          P[slot(k)][c] = fused_new<T, LAP, LW>(a, P[slot(k - 1)], P[slot(k)], P[slot(k + 1)], c, i, j, k, rb[u],
                                                dxinv, lambda, alpha, beta, c1, c2);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < NBLK; u++) carry[u] = rbn[u];
  };
  fetch(A, k0 - 2, k0 - 3, k0 - 3);
  put_phi(A, k0 - 2);
  fetch(A, k0 - 1, k0 - 3, k0 - 3);
  put_phi(A, k0 - 1);
  fetch(A, k0, k0 - 1, k0 - 1);
  fetch(B, k0 + 1, k0, k0);
#pragma unroll
  for (int u = 0; u < NBLK; u++) carry[u] = T(0.0);  // black rhs of plane k0-2: never used
  for (int k = k0 - 2; k < k1; k += 2) {
    step(A, k);
    if (k + 1 < k1) step(B, k + 1);
  }
  lds_barrier();
  store(k1 - 1);
}
// Tile shapes of the one-launch relaxation: {FX, FY, NT, scheme}, column FX x FY,
// NT threads; scheme 0 is fused_column (five LDS planes), scheme 1 is
// zmarch_column (zmarch.hpp: z neighbours in registers, (FX/2 + 2) x (FY + 4)
// threads own the column grown by 2 in x pairs), scheme 2 is zmarch4_column (the
// same with x quads, (FX/4 + 1) x (FY + 4) threads; FP32 and x extents that are a
// multiple of 4 only).  All give bitwise the same result; the fastest depends on
// the level's size.
// Shapes 10-12 are sized so that the red cells of the column grown by one,
// (FY + 2)(FX / 2 + 1), just fit in one pass of the NT threads.
constexpr int kFusedShapeTable[][4] = {
    {64, 8, 256, 0},   {32, 8, 256, 0},    {64, 4, 256, 0},    {32, 16, 256, 0},  {128, 4, 256, 0},
    {64, 16, 256, 0},  {32, 16, 512, 0},   {64, 16, 512, 0},   {64, 32, 1024, 0}, {128, 8, 1024, 0},
    {64, 13, 512, 0},  {32, 13, 256, 0},   {64, 29, 1024, 0},  {64, 26, 1024, 1}, {32, 24, 512, 1},
    {64, 11, 512, 1},  {32, 10, 256, 1},   {128, 11, 1024, 1}, {16, 21, 256, 1},  {64, 11, 256, 2},
    {128, 11, 512, 2}, {64, 26, 512, 2},   {32, 24, 256, 2},   {128, 26, 1024, 2}, {64, 56, 1024, 2},
    {16, 47, 256, 2}};
// Whether a shape applies to a rectangle of x extent nx in elem_bytes-wide values
// (scheme 2: FP32, nx a multiple of 4).
inline bool fused_shape_ok(int shape, int elem_bytes, int nx) {
  return kFusedShapeTable[shape][3] != 2 || (elem_bytes == 4 && nx % 4 == 0);
}
// x-quad rows: quads start at lo[0] - 2 (one value precedes the first row) and the
// one holding hi[0] ends at hi[0] + 2 (rows padded by at least 4 values)
inline void check_quad_rows(const FusedArgs& a) {
  const int nx = a.hi[0] - a.lo[0] + 1;
  if (nx % 4 != 0 || a.off < 1 || a.sy < nx + 4)
    throw std::runtime_error("one-launch relaxation: x-quad tiles need an x extent that is a multiple of 4 and padded rows");
}

// consecutive work items on one XCD (workgroups go round-robin over the 8 XCDs,
// grid a multiple of 8), so neighbouring columns share that XCD's L2 for their
// overlapping halos
__device__ __forceinline__ int xcd_contiguous_block() {
  return (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
}

}  // namespace ops
}  // namespace amrmg
