// The one-launch GSRB relaxation of a merged level with the z neighbours held in
// registers (device side), shared by the single-hierarchy launcher (kernels.hip)
// and the batched solver (batch.hip).  It computes exactly what fused_column
// (fused.hpp) computes -- one red pass and one black pass with the ghost refresh
// before each formed in registers -- with a different blocking:
//
// - Thread (px, py) of a PX x PY tile owns the x-adjacent pair of cells
//   (x0 - 2 + 2 px, x0 - 1 + 2 px) of row y0 - 2 + py in every plane of the
//   z-chunk.  The owned pairs cover the column FX x FY (FX = 2 PX - 4,
//   FY = PY - 4) grown by 2 in x and y: the outer ring is loaded only, the next
//   ring gets red updates only (the neighbouring workgroups recompute it too),
//   the column gets both colours and is stored.
// - Of the seven operands of an update, the own value, the other cell of the
//   pair and both z neighbours are this thread's registers; only the far x
//   neighbour and the two y neighbours are read from LDS.  A cell's red update
//   at plane k+1 and its black update at plane k are the same column, so each
//   step updates one cell of the pair twice.
// - Three LDS plane buffers, one barrier per z-step.  Step k:
//     barrier; write plane k+2's old values into buffer (k+2) mod 3 (its last
//     reader, the black pass of plane k-1, finished before the barrier);
//     red pass of plane k+1 (black neighbours: old values of buffer (k+1) mod 3,
//     written in step k-1), writing the new red values into that buffer;
//     black pass of plane k (red neighbours: new values of buffer k mod 3,
//     written in step k-1); store plane k.
//   Within a step the red pass reads only black positions and writes only red
//   ones, so no second barrier is needed.
// - A pair is moved as one 8- or 16-byte load, LDS write and store.  In
//   registers each plane is kept as (red value, black value), reordered once
//   when it arrives, so the updates select nothing.
// - Old values of plane k+2+DEPTH and the rhs of plane k+1+DEPTH are loaded
//   during step k (DEPTH steps ahead of their use); each load is unconditional
//   from an address clamped into the rectangle, and a value at a position
//   outside it is never used.
#pragma once
#include <type_traits>

#include "fused.hpp"

namespace amrmg {
namespace ops {

// x pair of (red, black) values of one plane, or a pair in cell order
template <class T>
struct RB {
  T r, b;
};

// sum of the six neighbours and the update, in the reference's operand order
// (xp + xm + yp + ym + zp + zm; the first sum is commutative, so the x pair
// enters as (in-pair neighbour + far neighbour) without knowing which is which)
template <class T, bool LAP>
__device__ __forceinline__ T zm_new(T sx, T yp, T ym, T zp, T zm, T own, T r, T dxinv, T lambda, T alpha, T beta) {
  const T s6 = sx + yp + ym + zp + zm;
  if (LAP) {
    const T lphi = s6 * dxinv;
    return lambda * (r - lphi);
  } else {
    const T lphi = (s6 - T(6.0) * own) * dxinv;
    const T helmop = alpha * own + beta * lphi;
    return own + lambda * (helmop - r);
  }
}

// the three LDS planes of a z-march tile (one allocation per tile shape, shared
// by both instances of zmarch_body)
template <class T, int FX, int FY>
__device__ __forceinline__ T (*zmarch_lds())[(FX + 4) * (FY + 4)] {
  __shared__ __align__(16) T S[3][(FX + 4) * (FY + 4)];
  return S;
}

// ZSRC: the old values are all zero (src is not read).
// The code in the z march is kept free of per-lane branches, which cost more
// scalar instructions than the stencil costs vector ones:
// - every global load is unconditional and its value is always used, so the
//   compiler's wait counts keep the loads of later planes in flight;
// - every lane computes both updates of a step, from LDS offsets fixed per
//   thread (clamped into the plane for lanes outside the updated region).  A
//   lane outside that region computes values nothing stored depends on: a red
//   value is read by the black updates of its x / y / z neighbours and of its
//   own column, a black value only by the store, and the black updates of the
//   column (the stored cells) read red values of the column grown by 1 only
//   (face cells replace an outside neighbour by the face rule);
// - the face rules are formed only in tiles that touch a face of the rectangle
//   in x or y (a branch that is uniform over the workgroup), and z faces only on
//   the first and last plane (uniform as well).
template <class T, bool LAP, int FX, int FY, int NT, int DEPTH, bool ZSRC>
__device__ __forceinline__ void zmarch_body(const FusedArgs& a, int w, T* __restrict__ dst,
                                            const T* __restrict__ src, const T* __restrict__ rhs, T dxinv,
                                            T lambda, T alpha, T beta, T c1, T c2) {
  constexpr int PX = FX / 2 + 2, PY = FY + 4, NP = PX * PY;
  constexpr int LW = 2 * PX, LPL = LW * PY;  // LDS plane: the column grown by 2
  static_assert(FX % 2 == 0 && NP <= NT, "z-march tile does not fit the workgroup");
  static_assert(DEPTH >= 2 && DEPTH <= 4, "loads run 2 to 4 steps ahead");
  T(*S)[LPL] = zmarch_lds<T, FX, FY>();
  using P2 = Pair<T>;

  // threads beyond the tile take the position of thread 0: they load and write
  // into LDS what thread 0 does (the same values) and update and store nothing
  const bool act = threadIdx.x < NP;
  const int tid = act ? (int)threadIdx.x : 0;
  const int px = tid % PX, py = tid / PX;
  const int tx = w % a.ntx, ty = (w / a.ntx) % a.nty, ch = w / (a.ntx * a.nty);
  const int x0 = a.lo[0] + tx * FX, y0 = a.lo[1] + ty * FY;
  const int k0 = a.lo[2] + ch * a.kc;
  const int k1 = min(k0 + a.kc, a.hi[2] + 1);
  const int i0 = x0 - 2 + 2 * px, j = y0 - 2 + py;
  const int pmax = a.lo[0] + 2 * ((a.hi[0] - a.lo[0]) / 2);  // start of the pair that holds hi[0]
  const long long xy = a.off + (clampi(i0, a.lo[0], pmax) - a.lo[0] + 1) +
                       (long long)(clampi(j, a.lo[1], a.hi[1]) - a.lo[1] + 1) * a.sy;
  auto zoff = [&](int p) { return (long long)(clampi(p, a.lo[2], a.hi[2]) - a.lo[2] + 1) * a.sz; };
  const int rb = py * LW + 2 * px;  // LDS index of the pair's first cell

  // In step k the red cell of plane k+1 and the black cell of plane k are the
  // pair's cell of index s = (i0 + j + k + 1) & 1.  With P = (k - k0) & 1 (the
  // parity of the step, known at compile time in the unrolled march),
  // s = sA ^ P.  Per P: the LDS offsets of the cell and of its far x and its y
  // neighbours, which updates the thread makes, and its x / y faces.
  const int sA = (i0 + j + k0 + 1) & 1;
  int o_own[2], o_far[2], o_yp[2], o_ym[2];
  bool sodd[2];
#pragma unroll
  for (int P = 0; P < 2; P++) {
    const int s = sA ^ P;
    sodd[P] = s == 1;
    o_own[P] = rb + s;
    o_far[P] = clampi(rb + (s ? 2 : -1), 0, LPL - 1);
    o_yp[P] = min(rb + LW + s, LPL - 1);
    o_ym[P] = max(rb - LW + s, 0);
  }
  // does any cell this tile updates lie on an x or y face of the rectangle?
  const bool tile_face = x0 - 1 <= a.lo[0] || x0 + FX >= a.hi[0] || y0 - 1 <= a.lo[1] || y0 + FY >= a.hi[1];
  // the pair is stored when it lies in the column; on an odd-width rectangle the
  // pair that holds hi[0] ends on the ghost cell hi[0] + 1, and only its first
  // cell is stored (odd_tile: this tile holds such a pair)
  const bool st_ok = act && px >= 1 && px <= PX - 2 && i0 <= a.hi[0] && j >= y0 && j <= min(y0 + FY - 1, a.hi[1]);
  const bool odd_tile = ((a.hi[0] - a.lo[0]) & 1) == 0 && x0 + FX > a.hi[0];
  const bool st_pair = i0 + 1 <= a.hi[0];

  // registers: 6 plane slots, plane p in slot (p - (k0 - 2)) mod 6
  P2 ph[6];    // old values as loaded (cell order), in flight
  P2 rr[6];    // rhs as loaded (cell order), in flight
  RB<T> v[6];  // (red, black) values of a plane
  RB<T> f[6];  // (red, black) rhs of a plane
  auto load_phi = [&](P2& d, int p) {
    if (ZSRC) d = P2{T(0.0), T(0.0)};
    else d = *reinterpret_cast<const P2*>(src + xy + zoff(p));
  };
  auto load_rhs = [&](P2& d, int p) { d = *reinterpret_cast<const P2*>(rhs + xy + zoff(p)); };
  // reorder a loaded pair into (red, black): first = true when the red cell is
  // the pair's first
  auto to_rb = [&](const P2& d, bool first) { return first ? RB<T>{d.x, d.y} : RB<T>{d.y, d.x}; };

  // new value of the cell of step parity P at plane q: own value own, the other
  // cell of the pair oth, z neighbours zp / zm; LDS plane L holds its x and y
  // neighbours
  auto update = [&](auto ZFc, const T* L, int P, int q, T own, T oth, T zp, T zm, T r) {
    constexpr bool ZF = decltype(ZFc)::value;  // q is the first or last plane of the rectangle
    const T far = L[o_far[P]], yp = L[o_yp[P]], ym = L[o_ym[P]];
    T sx = oth + far, yp2 = yp, ym2 = ym;
    if (tile_face) {
      const int i = i0 + (sA ^ P);
      const T vxp = sodd[P] ? far : oth, vxm = sodd[P] ? oth : far;
      const T xp = (i == a.hi[0]) ? face_ghost(a.kind[0][1], own, vxm, c1, c2) : vxp;
      const T xm = (i == a.lo[0]) ? face_ghost(a.kind[0][0], own, vxp, c1, c2) : vxm;
      sx = xp + xm;
      yp2 = (j == a.hi[1]) ? face_ghost(a.kind[1][1], own, ym, c1, c2) : yp;
      ym2 = (j == a.lo[1]) ? face_ghost(a.kind[1][0], own, yp, c1, c2) : ym;
    }
    T zp2 = zp, zm2 = zm;
    if (ZF) {
      zp2 = (q == a.hi[2]) ? face_ghost(a.kind[2][1], own, zm, c1, c2) : zp;
      zm2 = (q == a.lo[2]) ? face_ghost(a.kind[2][0], own, zp, c1, c2) : zm;
    }
    return zm_new<T, LAP>(sx, yp2, ym2, zp2, zm2, own, r, dxinv, lambda, alpha, beta);
  };

  // plane q relative to the rectangle: q - lo[2] <= nz1 (unsigned) holds for the
  // planes of the rectangle, q - lo[2] - 1 < nz1 - 1 for those strictly inside
  // (the others take the z face rules)
  const unsigned nz1 = (unsigned)(a.hi[2] - a.lo[2]);
  auto in_rect = [&](int q) { return (unsigned)(q - a.lo[2]) <= nz1; };
  auto inside = [&](int q) { return (unsigned)(q - a.lo[2] - 1) < nz1 - 1u; };

  // step k = k0 - 2 + 6 m + U; plane k + d is in slot (U + d) mod 6 and LDS
  // buffer (U + d) mod 3.  FULL: the step has a black pass (k >= k0; the two
  // steps before k0 only compute the red cells of planes k0-1 and k0).
  auto step = [&](auto Uc, auto FULLc, int k) {
    constexpr int U = decltype(Uc)::value;
    constexpr bool FULL = decltype(FULLc)::value;
    constexpr int Sm1 = (U + 5) % 6, S0 = U % 6, S1 = (U + 1) % 6, S2 = (U + 2) % 6;
    constexpr int SPH = (U + 2 + DEPTH) % 6, SRH = (U + 1 + DEPTH) % 6;  // slots of this step's loads
    constexpr int P = U & 1;  // k - k0 = 6 m + U - 2
    lds_barrier();
    // plane k+2: old values into LDS and into registers as (red, black); the red
    // cell of plane k+2 has index 1 - s
    reinterpret_cast<P2*>(S[S2 % 3])[rb / 2] = ph[S2];
    v[S2] = to_rb(ph[S2], sodd[P]);
    f[S1] = to_rb(rr[S1], !sodd[P]);  // rhs of plane k+1 (red index s)
    load_phi(ph[SPH], k + 2 + DEPTH);
    load_rhs(rr[SRH], k + 1 + DEPTH);
    // red pass of plane k+1: the cell of index s; black neighbours are old
    const int r = k + 1;
    if (in_rect(r)) {
      T* L = S[S1 % 3];
      T nv;
      if (inside(r))
        nv = update(std::false_type{}, L, P, r, v[S1].r, v[S1].b, v[S2].b, v[S0].b, f[S1].r);
      else
        nv = update(std::true_type{}, L, P, r, v[S1].r, v[S1].b, v[S2].b, v[S0].b, f[S1].r);
      v[S1].r = nv;
      L[o_own[P]] = nv;
    }
    // black pass of plane k: the cell of index s; red neighbours are new
    if (FULL) {
      const T* L = S[S0 % 3];
      T nb;
      if (inside(k))
        nb = update(std::false_type{}, L, P, k, v[S0].b, v[S0].r, v[S1].r, v[Sm1].r, f[S0].b);
      else
        nb = update(std::true_type{}, L, P, k, v[S0].b, v[S0].r, v[S1].r, v[Sm1].r, f[S0].b);
      v[S0].b = nb;
      // store plane k in cell order: index s is the black cell
      const P2 o = sodd[P] ? P2{v[S0].r, v[S0].b} : P2{v[S0].b, v[S0].r};
      T* d = dst + xy + zoff(k);
      if (odd_tile) {
        if (st_ok) {
          if (st_pair) *reinterpret_cast<P2*>(d) = o;
          else *d = o.x;
        }
      } else if (st_ok) {
        *reinterpret_cast<P2*>(d) = o;
      }
    }
  };

  // prologue: old values of planes k0-2 .. k0+DEPTH-1 and the rhs of planes
  // k0-1 .. k0+DEPTH-2 (plane p in slot p - k0 + 2); plane k0-1 goes into LDS
  // buffer 1, planes k0-2 and k0-1 into registers.  The red cell of plane p has
  // index (i0 + j + p) & 1: 1 - sA at k0-2 and sA at k0-1.
#pragma unroll
  for (int d = 0; d < DEPTH + 2; d++) load_phi(ph[d], k0 - 2 + d);
#pragma unroll
  for (int d = 1; d <= DEPTH; d++) load_rhs(rr[d], k0 - 2 + d);
  v[0] = to_rb(ph[0], sA == 1);  // plane k0-2
  v[1] = to_rb(ph[1], sA == 0);  // plane k0-1
  reinterpret_cast<P2*>(S[1])[rb / 2] = ph[1];
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I4 = std::integral_constant<int, 4>;
  using I5 = std::integral_constant<int, 5>;
  using F = std::false_type;
  using Tr = std::true_type;
  step(I0{}, F{}, k0 - 2);
  step(I1{}, F{}, k0 - 1);
  for (int kb = k0 - 2; kb + 2 < k1; kb += 6) {
    step(I2{}, Tr{}, kb + 2);
    if (kb + 3 < k1) step(I3{}, Tr{}, kb + 3);
    if (kb + 4 < k1) step(I4{}, Tr{}, kb + 4);
    if (kb + 5 < k1) step(I5{}, Tr{}, kb + 5);
    if (kb + 6 < k1) step(I0{}, Tr{}, kb + 6);
    if (kb + 7 < k1) step(I1{}, Tr{}, kb + 7);
  }
}

template <class T, bool LAP, int FX, int FY, int NT, int DEPTH = 2>
__device__ __forceinline__ void zmarch_column(const FusedArgs& a, int w, T* __restrict__ dst,
                                              const T* __restrict__ src, const T* __restrict__ rhs, T dxinv,
                                              T lambda, T alpha, T beta, T c1, T c2) {
  if (a.zsrc)
    zmarch_body<T, LAP, FX, FY, NT, DEPTH, true>(a, w, dst, src, rhs, dxinv, lambda, alpha, beta, c1, c2);
  else
    zmarch_body<T, LAP, FX, FY, NT, DEPTH, false>(a, w, dst, src, rhs, dxinv, lambda, alpha, beta, c1, c2);
}


// ---------------------------------------------------------------------------
// The same march with x quads (FP32): thread (px, py) of a PX x PY tile owns the
// four x-adjacent cells i0 .. i0+3, i0 = x0 - 2 + 4 px, of row y0 - 2 + py; the
// quads again cover the column grown by 2 in x and y.  A quad moves as one 16-byte
// load and store of cell-order values and one 16-byte LDS write, so per cell the
// march issues half the loads, LDS operations and address arithmetic of the pair
// form, which is what limits FP32.
// - Within a row the cells of one parity class, positions {s, s+2} of the quad, are
//   updated together: in step k the red cells of plane k+1 and the black cells of
//   plane k are the positions {s, s+2}, s = (i0 + j + k + 1) & 1.
// - LDS planes hold each quad by position parity, (p0, p2, p1, p3): the y
//   neighbours of positions {s, s+2} are the pair at parity s of the quads above and
//   below (one 8-byte read each), the far x neighbour is p3 of the quad to the left
//   (s = 0) or p0 of the quad to the right (s = 1).
// - In registers each plane is kept as (red pair, black pair), each in x order.
// Requirements, checked by the launchers: the rectangle's x extent is a multiple of
// 4 (so column and face bounds fall between the quad's two pairs, and the quad
// holding hi[0] ends at hi[0] + 2, inside the row's padding), rows have at least 4
// padding values and one value precedes the first row (quads start at lo[0] - 2).
template <class T>
struct alignas(2 * sizeof(T)) Quad {
  T a, b, c, d;  // cells i0 .. i0+3
};
template <class T>
struct alignas(4 * sizeof(T)) QuadLds {
  T e0, e1, o0, o1;  // positions 0, 2, 1, 3
};
// (red, black) pairs of a plane, each pair in x order
template <class T>
struct RB4 {
  Pair<T> r, b;
};

template <class T, int FX, int FY>
__device__ __forceinline__ T (*zmarch4_lds())[4 * (FX / 4 + 1) * (FY + 4)] {
  __shared__ __align__(16) T S[3][4 * (FX / 4 + 1) * (FY + 4)];
  return S;
}

template <class T, bool LAP, int FX, int FY, int NT, int DEPTH, bool ZSRC>
__device__ __forceinline__ void zmarch4_body(const FusedArgs& a, int w, T* __restrict__ dst,
                                             const T* __restrict__ src, const T* __restrict__ rhs, T dxinv,
                                             T lambda, T alpha, T beta, T c1, T c2) {
  constexpr int PX = FX / 4 + 1, PY = FY + 4, NP = PX * PY;
  constexpr int LW = 4 * PX, LPL = LW * PY;
  static_assert(FX % 4 == 0 && NP <= NT, "z-march quad tile does not fit the workgroup");
  static_assert(DEPTH >= 2 && DEPTH <= 4, "loads run 2 to 4 steps ahead");
  T(*S)[LPL] = zmarch4_lds<T, FX, FY>();
  using P2 = Pair<T>;
  using Q4 = Quad<T>;

  const bool act = threadIdx.x < NP;
  const int tid = act ? (int)threadIdx.x : 0;
  const int px = tid % PX, py = tid / PX;
  const int tx = w % a.ntx, ty = (w / a.ntx) % a.nty, ch = w / (a.ntx * a.nty);
  const int x0 = a.lo[0] + tx * FX, y0 = a.lo[1] + ty * FY;
  const int k0 = a.lo[2] + ch * a.kc;
  const int k1 = min(k0 + a.kc, a.hi[2] + 1);
  const int i0 = x0 - 2 + 4 * px, j = y0 - 2 + py;
  // loads clamp the quad into [lo - 2, hi + 2] and the row into the rectangle
  const long long xy = a.off + (clampi(i0, a.lo[0] - 2, a.hi[0] - 1) - a.lo[0] + 1) +
                       (long long)(clampi(j, a.lo[1], a.hi[1]) - a.lo[1] + 1) * a.sy;
  auto zoff = [&](int p) { return (long long)(clampi(p, a.lo[2], a.hi[2]) - a.lo[2] + 1) * a.sz; };
  const int qb = py * LW + 4 * px;  // LDS index of the quad

  const int sA = (i0 + j + k0 + 1) & 1;
  int o_own[2], o_far[2], o_yp[2], o_ym[2];
  bool sodd[2];
#pragma unroll
  for (int P = 0; P < 2; P++) {
    const int s = sA ^ P;
    sodd[P] = s == 1;
    o_own[P] = qb + 2 * s;
    o_far[P] = clampi(s ? qb + 4 : qb - 1, 0, LPL - 1);
    o_yp[P] = min(qb + LW + 2 * s, LPL - 2);
    o_ym[P] = max(qb - LW + 2 * s, 0);
  }
  const bool tile_face = x0 - 1 <= a.lo[0] || x0 + FX >= a.hi[0] || y0 - 1 <= a.lo[1] || y0 + FY >= a.hi[1];
  // stores: the quad's first pair (i0, i0+1) and second pair (i0+2, i0+3) each lie
  // wholly inside or outside the column and the rectangle
  const int xe = min(x0 + FX - 1, a.hi[0]);
  const bool st_row = act && j >= y0 && j <= min(y0 + FY - 1, a.hi[1]);
  const bool st_p0 = st_row && i0 >= x0 && i0 + 1 <= xe;
  const bool st_p1 = st_row && i0 + 2 >= x0 && i0 + 3 <= xe;

  Q4 ph[6];     // old values as loaded (cell order), in flight
  Q4 rr[6];     // rhs as loaded (cell order), in flight
  RB4<T> v[6];  // (red, black) values of a plane
  RB4<T> f[6];  // (red, black) rhs of a plane
  auto load_phi = [&](Q4& d, int p) {
    if (ZSRC) d = Q4{T(0.0), T(0.0), T(0.0), T(0.0)};
    else d = *reinterpret_cast<const Q4*>(src + xy + zoff(p));
  };
  auto load_rhs = [&](Q4& d, int p) { d = *reinterpret_cast<const Q4*>(rhs + xy + zoff(p)); };
  // (red, black) from cell order; red_odd: the red cells are positions 1 and 3
  auto to_rb = [&](const Q4& q, bool red_odd) {
    const P2 ev{q.a, q.c}, od{q.b, q.d};
    return red_odd ? RB4<T>{od, ev} : RB4<T>{ev, od};
  };

  // new values of the two cells at positions {s, s+2} of plane q: own pair, the
  // other pair oth (positions {s^1, s^1+2}), z neighbour pairs zp / zm
  auto update = [&](auto ZFc, const T* L, int P, int q, P2 own, P2 oth, P2 zp, P2 zm, P2 r) {
    constexpr bool ZF = decltype(ZFc)::value;
    const bool so = sodd[P];
    const T far = L[o_far[P]];
    const P2 yp = *reinterpret_cast<const P2*>(L + o_yp[P]);
    const P2 ym = *reinterpret_cast<const P2*>(L + o_ym[P]);
    // x neighbours: s = 0: cell p0 (far, p1), cell p2 (p1, p3); s = 1: cell p1 (p0, p2),
    // cell p3 (p2, far); (xm, xp) per cell
    const T xm0 = so ? oth.x : far, xp0 = so ? oth.y : oth.x;
    const T xm1 = so ? oth.y : oth.x, xp1 = so ? far : oth.y;
    T sx0 = xp0 + xm0, sx1 = xp1 + xm1;
    T yp0 = yp.x, ym0 = ym.x, yp1 = yp.y, ym1 = ym.y;
    if (tile_face) {
      const int ia = i0 + (so ? 1 : 0), ib = ia + 2;
      const T p0 = (ia == a.hi[0]) ? face_ghost(a.kind[0][1], own.x, xm0, c1, c2) : xp0;
      const T m0 = (ia == a.lo[0]) ? face_ghost(a.kind[0][0], own.x, xp0, c1, c2) : xm0;
      const T p1 = (ib == a.hi[0]) ? face_ghost(a.kind[0][1], own.y, xm1, c1, c2) : xp1;
      const T m1 = (ib == a.lo[0]) ? face_ghost(a.kind[0][0], own.y, xp1, c1, c2) : xm1;
      sx0 = p0 + m0;
      sx1 = p1 + m1;
      if (j == a.hi[1]) {
        yp0 = face_ghost(a.kind[1][1], own.x, ym.x, c1, c2);
        yp1 = face_ghost(a.kind[1][1], own.y, ym.y, c1, c2);
      }
      if (j == a.lo[1]) {
        ym0 = face_ghost(a.kind[1][0], own.x, yp.x, c1, c2);
        ym1 = face_ghost(a.kind[1][0], own.y, yp.y, c1, c2);
      }
    }
    T zp0 = zp.x, zm0 = zm.x, zp1 = zp.y, zm1 = zm.y;
    if (ZF) {
      if (q == a.hi[2]) {
        zp0 = face_ghost(a.kind[2][1], own.x, zm.x, c1, c2);
        zp1 = face_ghost(a.kind[2][1], own.y, zm.y, c1, c2);
      }
      if (q == a.lo[2]) {
        zm0 = face_ghost(a.kind[2][0], own.x, zp.x, c1, c2);
        zm1 = face_ghost(a.kind[2][0], own.y, zp.y, c1, c2);
      }
    }
    return P2{zm_new<T, LAP>(sx0, yp0, ym0, zp0, zm0, own.x, r.x, dxinv, lambda, alpha, beta),
              zm_new<T, LAP>(sx1, yp1, ym1, zp1, zm1, own.y, r.y, dxinv, lambda, alpha, beta)};
  };

  const unsigned nz1 = (unsigned)(a.hi[2] - a.lo[2]);
  auto in_rect = [&](int q) { return (unsigned)(q - a.lo[2]) <= nz1; };
  auto inside = [&](int q) { return (unsigned)(q - a.lo[2] - 1) < nz1 - 1u; };

  auto step = [&](auto Uc, auto FULLc, int k) {
    constexpr int U = decltype(Uc)::value;
    constexpr bool FULL = decltype(FULLc)::value;
    constexpr int Sm1 = (U + 5) % 6, S0 = U % 6, S1 = (U + 1) % 6, S2 = (U + 2) % 6;
    constexpr int SPH = (U + 2 + DEPTH) % 6, SRH = (U + 1 + DEPTH) % 6;
    constexpr int P = U & 1;
    lds_barrier();
    // plane k+2: old values into LDS by position parity and into registers; its red
    // cells are positions {s^1, s^1+2}
    {
      const Q4& q = ph[S2];
      *reinterpret_cast<QuadLds<T>*>(&S[S2 % 3][qb]) = QuadLds<T>{q.a, q.c, q.b, q.d};
    }
    v[S2] = to_rb(ph[S2], !sodd[P]);
    f[S1] = to_rb(rr[S1], sodd[P]);  // rhs of plane k+1 (red positions {s, s+2})
    load_phi(ph[SPH], k + 2 + DEPTH);
    load_rhs(rr[SRH], k + 1 + DEPTH);
    // red pass of plane k+1: black neighbours are old
    const int r = k + 1;
    if (in_rect(r)) {
      T* L = S[S1 % 3];
      P2 nv;
      if (inside(r))
        nv = update(std::false_type{}, L, P, r, v[S1].r, v[S1].b, v[S2].b, v[S0].b, f[S1].r);
      else
        nv = update(std::true_type{}, L, P, r, v[S1].r, v[S1].b, v[S2].b, v[S0].b, f[S1].r);
      v[S1].r = nv;
      *reinterpret_cast<P2*>(L + o_own[P]) = nv;
    }
    // black pass of plane k: red neighbours are new
    if (FULL) {
      const T* L = S[S0 % 3];
      P2 nb;
      if (inside(k))
        nb = update(std::false_type{}, L, P, k, v[S0].b, v[S0].r, v[S1].r, v[Sm1].r, f[S0].b);
      else
        nb = update(std::true_type{}, L, P, k, v[S0].b, v[S0].r, v[S1].r, v[Sm1].r, f[S0].b);
      v[S0].b = nb;
      // cell order: the black cells are positions {s, s+2}
      const P2 R = v[S0].r;
      const Q4 o = sodd[P] ? Q4{R.x, nb.x, R.y, nb.y} : Q4{nb.x, R.x, nb.y, R.y};
      T* d = dst + xy + zoff(k);
      if (st_p0) *reinterpret_cast<P2*>(d) = P2{o.a, o.b};
      if (st_p1) *reinterpret_cast<P2*>(d + 2) = P2{o.c, o.d};
    }
  };

  // prologue as in zmarch_body: the red cells of plane p are positions
  // {(i0 + j + p) & 1, ...}: odd (1 - sA) at k0-2, sA at k0-1
#pragma unroll
  for (int d = 0; d < DEPTH + 2; d++) load_phi(ph[d], k0 - 2 + d);
#pragma unroll
  for (int d = 1; d <= DEPTH; d++) load_rhs(rr[d], k0 - 2 + d);
  v[0] = to_rb(ph[0], sA == 0);  // plane k0-2
  v[1] = to_rb(ph[1], sA == 1);  // plane k0-1
  {
    const Q4& q = ph[1];
    *reinterpret_cast<QuadLds<T>*>(&S[1][qb]) = QuadLds<T>{q.a, q.c, q.b, q.d};
  }
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I4 = std::integral_constant<int, 4>;
  using I5 = std::integral_constant<int, 5>;
  using F = std::false_type;
  using Tr = std::true_type;
  step(I0{}, F{}, k0 - 2);
  step(I1{}, F{}, k0 - 1);
  for (int kb = k0 - 2; kb + 2 < k1; kb += 6) {
    step(I2{}, Tr{}, kb + 2);
    if (kb + 3 < k1) step(I3{}, Tr{}, kb + 3);
    if (kb + 4 < k1) step(I4{}, Tr{}, kb + 4);
    if (kb + 5 < k1) step(I5{}, Tr{}, kb + 5);
    if (kb + 6 < k1) step(I0{}, Tr{}, kb + 6);
    if (kb + 7 < k1) step(I1{}, Tr{}, kb + 7);
  }
}

template <class T, bool LAP, int FX, int FY, int NT, int DEPTH = 2>
__device__ __forceinline__ void zmarch4_column(const FusedArgs& a, int w, T* __restrict__ dst,
                                               const T* __restrict__ src, const T* __restrict__ rhs, T dxinv,
                                               T lambda, T alpha, T beta, T c1, T c2) {
  if (a.zsrc)
    zmarch4_body<T, LAP, FX, FY, NT, DEPTH, true>(a, w, dst, src, rhs, dxinv, lambda, alpha, beta, c1, c2);
  else
    zmarch4_body<T, LAP, FX, FY, NT, DEPTH, false>(a, w, dst, src, rhs, dxinv, lambda, alpha, beta, c1, c2);
}

}  // namespace ops
}  // namespace amrmg
