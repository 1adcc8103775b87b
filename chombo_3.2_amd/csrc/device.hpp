// Device-resident data model of the AMR Poisson/Helmholtz path.
//
// A level (one DisjointBoxLayout) lives in HBM as a box table plus, per field,
// one slab holding every box's FAB (valid box grown by one ghost cell, Fortran
// order, i fastest).  All kernels address a cell through the box table:
//   slab[ box.off + (i-lo0+1) + (j-lo1+1)*sy + (k-lo2+1)*sz ].
// Every field of a level shares the same box table, so a kernel over several
// fields uses one index.  Work over the valid cells of all boxes is enumerated by
// a tile table (one 256-thread workgroup per tile), so every per-box Fortran call
// of the reference (FORT_* in AMRPoissonOp.cpp) becomes one launch per level.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "geometry.hpp"

namespace amrmg {

// ---------------- POD tables shared by host and device ----------------
struct DBox {
  int lo[3], hi[3];
  long long off;  // slab index of cell (lo-1, lo-1, lo-1)
  int sy, sz;     // strides in y and z (x stride is 1)
};

struct Tile {
  int box, i0, j0, k0;
};

// One rectangle of face-ghost cells and the rule that fills it.
enum PatchKind : int {
  PK_EXCHANGE = 0,   // copy from the valid cells of box `src` (Copier exchange)
  PK_DIRICHLET = 1,  // ConstBCFunction type 1: g = 2*v - near
  PK_NEUMANN = 2,    // ConstBCFunction type 0: g = near (+ isign*dx*v)
  PK_CFHOMO = 3,     // INTERPHOMO: g = c1*phi(1 in) + c2*phi(2 in)
  PK_CFHOMOLIN = 4,  // INTERPHOMOLINEAR: g = factor*phi(1 in)
  PK_DIRICHLET2 = 5, // DiriBC order 2 (quadraticInterp): g = (8/3)*v + (1/3)*phi(2 in) - 2*phi(1 in)
};
struct Patch {
  int kind, dst, src, dir, side;
  int lo[3], hi[3];
  int sh[3];  // PK_EXCHANGE: ghost cell c takes box src's cell c + sh (a periodic image; else 0)
  int voff;   // BC rules: -1 for the face constant val, else the offset of this patch's face
              // values (one per ghost cell, x fastest) in the level's value array
  double val;
};
struct PatchWork {  // a patch split into chunks of at most 256 cells
  int patch, start;
};

// Motion item of a Copier between two layouts (same index space)
struct Motion {
  int from, to;
  int lo[3], hi[3];
};
struct MotionWork {
  int item, start;
};

// Inhomogeneous coarse-fine interpolation (QuadCFInterp), coarse side: one record
// per (fine box, face, coarse cell under the face).  All offsets index the
// coarse level's slab.
enum DerivKind : int { DK_STD = 0, DK_GEN = 1, DK_ZERO = 2 };
struct CFCoarse {
  int pc;              // coarse cell
  int std;             // 1: all derivatives use the centred formulas
  int nbr[8];          // std: you1-, you1+, you2-, you2+, ur, ul, lr, ll
  int gen;             // index into the generic pool when !std
};
struct CFGen {
  int kind[5];         // first(you1), second(you1), first(you2), second(you2), mixed
  int n[5];
  int off[5][9];
  double w[5][9];
};
// fine side: one record per CF ghost cell
struct CFFine {
  long long g;         // ghost cell (fine slab)
  long long a1, a2;    // one / two cells inward
  int coarse;          // CFCoarse index
  int packed;          // FORT_QUADINTERP (1) or interpOnIVS C++ loop (0)
  double x1, x2;       // tangential offsets, computed on the host as the reference does
};

// Reflux (LevelFluxRegister + AMRPoissonOp::reflux): one record per coarse cell
// that the reference's registers touch, with its contributions in reference order.
struct RefluxCell {
  long long u;         // output cell in the coarse residual/operator slab
  int c0, nc;          // coarse flux contributions [c0, c0+nc)
  int f0, nf;          // fine register contributions [f0, f0+nf)
  int sy, sz;          // strides of the cell's box (to re-form its 7-point stencil)
};
struct RefluxCoarse {  // reg += (phi[a]-phi[b])*(beta/dx) * sign
  long long a, b;
  double sign;
};
struct RefluxFine {    // reg_f = sum_q scale*(phi_f[a_q]-phi_f[b_q])*(beta*ref/dx)
  long long a[4], b[4];
  double scale;
  int first, last;     // a register of ratio^2 fluxes spans ratio^2/4 records, first..last
};

// ---------------- host-side runtime ----------------
// Kernel classes timed by the event profiler (bench roofline measurement).
enum KernelClass : int {
  KC_GSRB = 0, KC_RESIDUAL, KC_APPLY, KC_RESTRICT, KC_PROLONG, KC_AVERAGE, KC_GHOST, KC_BLAS, KC_REDUCE,
  KC_CF, KC_REFLUX, KC_MOTION, KC_COUNT
};
struct Profiler;

struct Context {
  int device = 0;
  bool dry = false;                   // host-only: build and check tables, never touch a device
  Profiler* prof = nullptr;           // non-null while event profiling is on
  bool fused_relax = true;            // one-launch GSRB iterations on merged levels
  hipStream_t stream = nullptr;
  double* d_scalars = nullptr;        // small device scratch for reductions
  double* h_scalars = nullptr;        // pinned host mirror
  std::string last_error;
  explicit Context(int dev, bool dry_run = false);
  ~Context();
  void sync();
};

// device allocation / copy helpers that are no-ops for a dry (host-only) context
void dev_alloc(Context* c, void** p, size_t bytes, const char* what);
void dev_free(Context* c, void* p);
void dev_h2d(Context* c, void* dst, const void* src, size_t bytes, const char* what);

// Storage of one level's fields.
//  - per-box: every box owns a ghosted FAB (box grown by 1), FABs back to back,
//    as the reference's LevelData<FArrayBox> holds them;
//  - merged: when the boxes tile their bounding rectangle exactly, the level is
//    one ghosted array over that rectangle.  Every box keeps its own DBox, with
//    the merged array's strides, so a box's face ghost next to another box is
//    that box's valid cell (a Copier exchange is the identity) and the ghost
//    cells outside the rectangle are each the face ghost of exactly one box.
//    Rows are padded to 16 elements and the array starts so that every row's
//    first valid cell is 128-byte aligned.  h_boxes[nbox] is the rectangle
//    itself and the tiles cover it rather than each box.
struct DeviceLayout {
  DBL dbl;
  Box domain;
  std::vector<DBox> h_boxes;   // nbox boxes, then (merged) the rectangle
  DBox* d_boxes = nullptr;
  int nbox = 0;
  bool merged = false;
  Box rect;                    // bounding rectangle of the boxes
  long long slab = 0;          // elements per field
  Tile* d_tiles = nullptr;
  int ntiles = 0;
  int TX = 64, TY = 4, TZ = 1;
  // merged levels: column tiles of 64 x 4 cells and kColumnPlanes planes over the
  // rectangle, for the stencil kernels that march in z
  static constexpr int kColumnPlanes = 8;
  Tile* d_ctiles = nullptr;
  int nctiles = 0;
  long long valid_cells = 0;
  Context* ctx = nullptr;
  DeviceLayout(Context* c, const DBL& d, const Box& dom, bool merge = true);
  ~DeviceLayout();
  DeviceLayout(const DeviceLayout&) = delete;
  DeviceLayout& operator=(const DeviceLayout&) = delete;
  long long index(int b, int i, int j, int k) const {
    const DBox& x = h_boxes[b];
    return x.off + (i - x.lo[0] + 1) + (long long)(j - x.lo[1] + 1) * x.sy + (long long)(k - x.lo[2] + 1) * x.sz;
  }
};

enum Prec : int { PREC_F64 = 0, PREC_F32 = 1 };

// A LevelData<FArrayBox> resident in HBM (ghost width 1) with nc components: each
// component is one field of the layout's slab, the components cs elements apart (cs
// is the slab rounded up to 64 elements, so every component keeps the slab's
// alignment).  A component view (alias) is a one-component LevelData over one
// component's storage that does not own it; comp says which component it is.
template <class T>
struct LevelData {
  const DeviceLayout* lay = nullptr;
  T* d = nullptr;
  int nc = 1;
  int comp = 0;
  long long cs = 0;
  bool own = true;
  LevelData() = default;
  explicit LevelData(const DeviceLayout* l, int ncomp = 1) { define(l, ncomp); }
  ~LevelData() { release(); }
  LevelData(const LevelData&) = delete;
  LevelData& operator=(const LevelData&) = delete;
  void define(const DeviceLayout* l, int ncomp = 1);
  void alias(const LevelData& x, int c) {
    release();
    if (x.d && (c < 0 || c >= x.nc)) throw std::runtime_error("LevelData: component out of range");
    lay = x.lay;
    d = x.d ? x.d + (long long)c * x.cs : nullptr;
    nc = 1;
    comp = c;
    cs = x.cs;
    own = false;
  }
  void release();
  bool isDefined() const { return d != nullptr; }
  static long long compStride(const DeviceLayout* l) { return (l->slab + 63) / 64 * 64; }
};

void hip_check(hipError_t e, const char* what);

}  // namespace amrmg
