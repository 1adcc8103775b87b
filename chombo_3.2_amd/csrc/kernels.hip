// HIP kernels for gfx950 (MI355X) of the AMR Poisson/Helmholtz hot path, plus the
// device runtime (context, layouts, fields).
//
// Arithmetic in every FP64 kernel follows the expression order of the reference
// kernel it replaces (cited per kernel) and the file is compiled with
// -ffp-contract=off, so an FP64 result is bitwise the reference's result for the
// same inputs.  Kernels are HBM-bound 7-point stencils and BLAS-1 sweeps; none is
// a dense contraction, so none uses MFMA.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "fused.hpp"
#include "ops.hpp"
#include "zmarch.hpp"

namespace amrmg {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
  }
}

// ---------------------------------------------------------------------------
// runtime
// ---------------------------------------------------------------------------
void dev_alloc(Context* c, void** p, size_t bytes, const char* what) {
  *p = nullptr;
  if (c->dry || bytes == 0) return;
  hip_check(hipMalloc(p, bytes), what);
}
void dev_free(Context* c, void* p) {
  if (c->dry || !p) return;
  (void)hipFree(p);
}
void dev_h2d(Context* c, void* dst, const void* src, size_t bytes, const char* what) {
  if (c->dry || bytes == 0) return;
  hip_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), what);
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry points (amrmg_field_dot_noalloc): the
// per-box partial sums of the ordered dot product live in `scratch`, a
// file-scope device array of kNoAllocBoxes doubles (64 KiB; each GPU holds its
// own copy until the process ends).  Its device address is looked up once per
// device, when the first context on that device is created, and only that
// address is ever handed to kernels or copies.  Every call on one GPU shares
// the array: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[ops::kNoAllocBoxes];
static std::mutex g_scratch_mu;
static std::vector<double*> g_scratch_ptr;  // per device: the address of `scratch`

static void resolve_scratch(int dev) {  // the device must be current
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (g_scratch_ptr[dev]) return;
  double* p = nullptr;
  hip_check(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (!p) throw std::runtime_error("hipGetSymbolAddress(scratch) returned no address");
  g_scratch_ptr[dev] = p;
}
static double* device_scratch(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0 || dev >= (int)g_scratch_ptr.size() || !g_scratch_ptr[dev])
    throw std::runtime_error("no-allocation scratch not set up on this device");
  return g_scratch_ptr[dev];
}

Context::Context(int dev, bool dry_run) : device(dev), dry(dry_run) {
  if (dry) return;
  hip_check(hipSetDevice(dev), "hipSetDevice");
  resolve_scratch(dev);
  hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipMalloc(&d_scalars, 64 * sizeof(double)), "hipMalloc scalars");
  hip_check(hipHostMalloc(&h_scalars, 64 * sizeof(double)), "hipHostMalloc scalars");
  hip_check(hipMemset(d_scalars, 0, 64 * sizeof(double)), "hipMemset scalars");
}
// ---------------------------------------------------------------------------
// event profiler: start/stop events around each launch of a kernel class
// ---------------------------------------------------------------------------
struct Profiler {
  unsigned mask = ~0u;
  struct Rec { int kind; hipEvent_t a, b; double bytes, units; };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  double ms[KC_COUNT] = {0};
  double bytes[KC_COUNT] = {0};
  double units[KC_COUNT] = {0};   // work units (cell-relaxations for the GSRB class)
  long long count[KC_COUNT] = {0};
  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
  }
  void drain() {
    for (auto& r : recs) {
      hip_check(hipEventSynchronize(r.b), "hipEventSynchronize");
      float t = 0;
      hip_check(hipEventElapsedTime(&t, r.a, r.b), "hipEventElapsedTime");
      ms[r.kind] += t; bytes[r.kind] += r.bytes; units[r.kind] += r.units; count[r.kind] += 1;
      pool.push_back(r.a); pool.push_back(r.b);
    }
    recs.clear();
  }
  ~Profiler() {
    for (auto& r : recs) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};
static hipEvent_t prof_begin(Context* c, int kind) {
  if (!c->prof || !(c->prof->mask & (1u << kind))) return nullptr;
  hipEvent_t a = c->prof->get();
  hip_check(hipEventRecord(a, c->stream), "hipEventRecord");
  return a;
}
static void prof_end(Context* c, hipEvent_t a, int kind, double bytes, double units = 0.0) {
  if (!c->prof || !a) return;
  hipEvent_t b = c->prof->get();
  hip_check(hipEventRecord(b, c->stream), "hipEventRecord");
  c->prof->recs.push_back(Profiler::Rec{kind, a, b, bytes, units});
  if (c->prof->recs.size() > 4096) c->prof->drain();
}
hipEvent_t profile_begin(Context* c, int kind) { return prof_begin(c, kind); }
void profile_end(Context* c, hipEvent_t a, int kind, double bytes, double units) {
  prof_end(c, a, kind, bytes, units);
}
void profiler_enable(Context* c, unsigned mask) {
  if (mask && !c->prof) c->prof = new Profiler();
  if (mask && c->prof) c->prof->mask = mask;
  if (!mask && c->prof) { (void)hipStreamSynchronize(c->stream); delete c->prof; c->prof = nullptr; }
}
void profiler_reset(Context* c) {
  if (!c->prof) return;
  (void)hipStreamSynchronize(c->stream);
  c->prof->drain();
  for (int q = 0; q < KC_COUNT; q++) {
    c->prof->ms[q] = 0; c->prof->bytes[q] = 0; c->prof->count[q] = 0; c->prof->units[q] = 0;
  }
}
bool profiler_read(Context* c, int kind, double* ms, double* bytes, long long* count, double* units) {
  if (!c->prof || kind < 0 || kind >= KC_COUNT) return false;
  (void)hipStreamSynchronize(c->stream);
  c->prof->drain();
  *ms = c->prof->ms[kind]; *bytes = c->prof->bytes[kind]; *count = c->prof->count[kind];
  if (units) *units = c->prof->units[kind];
  return true;
}

Context::~Context() {
  if (dry) return;
  if (prof) profiler_enable(this, 0u);
  if (stream) (void)hipStreamSynchronize(stream);
  if (d_scalars) (void)hipFree(d_scalars);
  if (h_scalars) (void)hipHostFree(h_scalars);
  if (stream) (void)hipStreamDestroy(stream);
}
void Context::sync() { hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize"); }

static int pow2_at_least(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

DeviceLayout::DeviceLayout(Context* c, const DBL& d, const Box& dom, bool merge) : dbl(d), domain(dom), ctx(c) {
  nbox = d.size();
  long long cells = 0;
  for (int b = 0; b < nbox; b++) {
    cells += d[b].npts();
    if (b == 0) rect = d[b];
    for (int q = 0; q < 3; q++) {
      rect.lo[q] = std::min(rect.lo[q], d[b].lo[q]);
      rect.hi[q] = std::max(rect.hi[q], d[b].hi[q]);
    }
  }
  valid_cells = cells;
  // disjoint boxes whose volumes add up to the bounding box's tile it exactly
  merged = merge && nbox > 0 && cells == rect.npts();
  h_boxes.resize(nbox + (merged ? 1 : 0));
  int maxnx = 1, maxny = 1;
  if (merged) {
    const int sy = (rect.len(0) + 2 + 15) / 16 * 16;
    const long long sz = (long long)sy * (rect.len(1) + 2);
    if (sz > INT32_MAX) throw std::runtime_error("merged level plane too large");
    const long long base = 15;  // slab index of (rect.lo - 1); rect.lo[0] lands on a multiple of 16
    for (int b = 0; b <= nbox; b++) {
      const Box& bx = b < nbox ? d[b] : rect;
      DBox x;
      for (int q = 0; q < 3; q++) { x.lo[q] = bx.lo[q]; x.hi[q] = bx.hi[q]; }
      x.sy = sy;
      x.sz = (int)sz;
      x.off = base + (bx.lo[0] - rect.lo[0]) + (long long)(bx.lo[1] - rect.lo[1]) * sy +
              (long long)(bx.lo[2] - rect.lo[2]) * sz;
      h_boxes[b] = x;
    }
    slab = base + sz * (rect.len(2) + 2);
    maxnx = rect.len(0);
    maxny = rect.len(1);
  } else {
    long long off = 0;
    for (int b = 0; b < nbox; b++) {
      const Box& bx = d[b];
      DBox x;
      for (int q = 0; q < 3; q++) { x.lo[q] = bx.lo[q]; x.hi[q] = bx.hi[q]; }
      int nx = bx.len(0) + 2, ny = bx.len(1) + 2, nz = bx.len(2) + 2;
      x.off = off;
      x.sy = nx;
      x.sz = nx * ny;
      off += (long long)nx * ny * nz;
      h_boxes[b] = x;
      maxnx = std::max(maxnx, bx.len(0));
      maxny = std::max(maxny, bx.len(1));
    }
    slab = off;
  }
  TX = std::min(64, pow2_at_least(maxnx));
  TY = std::min(256 / TX, pow2_at_least(maxny));
  TZ = 256 / (TX * TY);
  std::vector<Tile> tiles;
  const int tb0 = merged ? nbox : 0, tb1 = merged ? nbox + 1 : nbox;
  for (int b = tb0; b < tb1; b++) {
    const DBox& bx = h_boxes[b];
    for (int k = bx.lo[2]; k <= bx.hi[2]; k += TZ)
      for (int j = bx.lo[1]; j <= bx.hi[1]; j += TY)
        for (int i = bx.lo[0]; i <= bx.hi[0]; i += TX) tiles.push_back(Tile{b, i, j, k});
  }
  if (tiles.size() > (size_t)INT32_MAX) throw std::runtime_error("too many tiles");
  ntiles = (int)tiles.size();
  // every tile cell index stays inside its box's grown FAB (threads past the
  // box's high end return), so tiles need only a valid box index
  for (const Tile& t : tiles) {
    if (t.box < 0 || t.box >= (int)h_boxes.size()) throw std::runtime_error("tile outside its box");
    const DBox& x = h_boxes[t.box];
    if (t.i0 < x.lo[0] || t.i0 > x.hi[0] || t.j0 < x.lo[1] || t.j0 > x.hi[1] || t.k0 < x.lo[2] || t.k0 > x.hi[2])
      throw std::runtime_error("tile outside its box");
  }
  // every box's grown region addresses stay inside the slab
  for (int b = 0; b < (int)h_boxes.size(); b++) {
    const DBox& x = h_boxes[b];
    const long long lo = index(b, x.lo[0] - 1, x.lo[1] - 1, x.lo[2] - 1);
    const long long hi = index(b, x.hi[0] + 1, x.hi[1] + 1, x.hi[2] + 1);
    if (lo < 0 || hi >= slab) throw std::runtime_error("box storage outside the level slab");
  }
  dev_alloc(c, (void**)&d_boxes, sizeof(DBox) * h_boxes.size(), "hipMalloc boxes");
  dev_h2d(c, d_boxes, h_boxes.data(), sizeof(DBox) * h_boxes.size(), "copy boxes");
  dev_alloc(c, (void**)&d_tiles, sizeof(Tile) * tiles.size(), "hipMalloc tiles");
  dev_h2d(c, d_tiles, tiles.data(), sizeof(Tile) * tiles.size(), "copy tiles");
  if (merged) {
    std::vector<Tile> ct;
    const DBox& r = h_boxes[nbox];
    for (int k = r.lo[2]; k <= r.hi[2]; k += kColumnPlanes)
      for (int j = r.lo[1]; j <= r.hi[1]; j += 4)
        for (int i = r.lo[0]; i <= r.hi[0]; i += 64) ct.push_back(Tile{nbox, i, j, k});
    if (ct.size() > (size_t)INT32_MAX - 8) throw std::runtime_error("too many column tiles");
    nctiles = (int)ct.size();
    dev_alloc(c, (void**)&d_ctiles, sizeof(Tile) * ct.size(), "hipMalloc column tiles");
    dev_h2d(c, d_ctiles, ct.data(), sizeof(Tile) * ct.size(), "copy column tiles");
  }
}
DeviceLayout::~DeviceLayout() {
  dev_free(ctx, d_boxes);
  dev_free(ctx, d_tiles);
  dev_free(ctx, d_ctiles);
}

template <class T>
void LevelData<T>::define(const DeviceLayout* l, int ncomp) {
  release();
  if (ncomp < 1) throw std::runtime_error("LevelData: ncomp must be at least 1");
  lay = l;
  nc = ncomp;
  comp = 0;
  own = true;
  cs = ncomp > 1 ? compStride(l) : l->slab;
  const long long n = cs * ncomp;
  if (n > 0 && !l->ctx->dry) {
    hip_check(hipMalloc(&d, sizeof(T) * n), "hipMalloc field");
    hip_check(hipMemsetAsync(d, 0, sizeof(T) * n, l->ctx->stream), "hipMemset field");
  }
}
template <class T>
void LevelData<T>::release() {
  if (d && own) {
    (void)hipStreamSynchronize(lay->ctx->stream);
    (void)hipFree(d);
  }
  d = nullptr;
  own = true;
}
template struct LevelData<double>;
template struct LevelData<float>;

namespace ops {

void GhostPlan::upload(Context* ctx) {
  release();
  npatch = (int)h_patches.size();
  std::vector<PatchWork> work;
  for (int p = 0; p < npatch; p++) {
    const Patch& x = h_patches[p];
    long long n = (long long)(x.hi[0] - x.lo[0] + 1) * (x.hi[1] - x.lo[1] + 1) * (x.hi[2] - x.lo[2] + 1);
    for (long long s = 0; s < n; s += 256) work.push_back(PatchWork{p, (int)s});
  }
  nwork = (int)work.size();
  owner = ctx;
  dev_alloc(ctx, (void**)&d_patches, sizeof(Patch) * npatch, "hipMalloc patches");
  dev_h2d(ctx, d_patches, h_patches.data(), sizeof(Patch) * npatch, "copy patches");
  dev_alloc(ctx, (void**)&d_work, sizeof(PatchWork) * nwork, "hipMalloc patch work");
  dev_h2d(ctx, d_work, work.data(), sizeof(PatchWork) * nwork, "copy patch work");
  if (!h_bcv.empty()) {
    dev_alloc(ctx, (void**)&d_bcv, sizeof(double) * h_bcv.size(), "hipMalloc BC values");
    dev_h2d(ctx, d_bcv, h_bcv.data(), sizeof(double) * h_bcv.size(), "copy BC values");
  }
}
void GhostPlan::release() {
  if (owner) {
    dev_free(owner, d_patches);
    dev_free(owner, d_work);
    dev_free(owner, d_bcv);
  }
  d_patches = nullptr;
  d_work = nullptr;
  d_bcv = nullptr;
}
void MotionPlan::upload(Context* ctx) {
  release();
  nitems = (int)h_items.size();
  std::vector<MotionWork> work;
  for (int p = 0; p < nitems; p++) {
    const Motion& x = h_items[p];
    long long n = (long long)(x.hi[0] - x.lo[0] + 1) * (x.hi[1] - x.lo[1] + 1) * (x.hi[2] - x.lo[2] + 1);
    for (long long s = 0; s < n; s += 256) work.push_back(MotionWork{p, (int)s});
  }
  nwork = (int)work.size();
  owner = ctx;
  dev_alloc(ctx, (void**)&d_items, sizeof(Motion) * nitems, "hipMalloc motion");
  dev_h2d(ctx, d_items, h_items.data(), sizeof(Motion) * nitems, "copy motion");
  dev_alloc(ctx, (void**)&d_work, sizeof(MotionWork) * nwork, "hipMalloc motion work");
  dev_h2d(ctx, d_work, work.data(), sizeof(MotionWork) * nwork, "copy motion work");
}
void MotionPlan::release() {
  if (owner) {
    dev_free(owner, d_items);
    dev_free(owner, d_work);
  }
  d_items = nullptr;
  d_work = nullptr;
}

// ---------------------------------------------------------------------------
// index helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ long long cidx(const DBox& b, int i, int j, int k) {
  return b.off + (i - b.lo[0] + 1) + (long long)(j - b.lo[1] + 1) * b.sy + (long long)(k - b.lo[2] + 1) * b.sz;
}

struct TileGeom {
  int lx, ly;  // log2 TX, log2 TY
};
static TileGeom geom(const DeviceLayout& L) {
  TileGeom g;
  g.lx = 0;
  while ((1 << g.lx) < L.TX) g.lx++;
  g.ly = 0;
  while ((1 << g.ly) < L.TY) g.ly++;
  return g;
}
// cell of this thread in its tile; false if outside the box
__device__ __forceinline__ bool tile_cell(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                          TileGeom g, DBox& b, int& i, int& j, int& k) {
  const Tile t = tiles[blockIdx.x];
  b = boxes[t.box];
  const int tid = threadIdx.x;
  i = t.i0 + (tid & ((1 << g.lx) - 1));
  j = t.j0 + ((tid >> g.lx) & ((1 << g.ly) - 1));
  k = t.k0 + (tid >> (g.lx + g.ly));
  return i <= b.hi[0] && j <= b.hi[1] && k <= b.hi[2];
}

#define LAUNCH_PU(kind, bytes, units, kernel, grid, ...)                                 \
  do {                                                                                   \
    if ((grid) > 0) {                                                                    \
      hipEvent_t ev_ = prof_begin(L.ctx, kind);                                          \
      hipLaunchKernelGGL(kernel, dim3(grid), dim3(256), 0, L.ctx->stream, __VA_ARGS__);   \
      hip_check(hipGetLastError(), #kernel);                                             \
      prof_end(L.ctx, ev_, kind, bytes, units);                                          \
    }                                                                                    \
  } while (0)
#define LAUNCH_P(kind, bytes, kernel, grid, ...) LAUNCH_PU(kind, bytes, 0.0, kernel, grid, __VA_ARGS__)
#define LAUNCH(kernel, grid, ...) LAUNCH_P(KC_BLAS, 0.0, kernel, grid, __VA_ARGS__)

// ---------------------------------------------------------------------------
// face-ghost fill: Copier exchange (Copier.cpp:1464-1541 with trimEdges),
// ConstBCFunction (BCFunc.cpp:58-158), homogeneous CF (AMRPoissonOp.cpp:1598-1794,
// InterpF.ChF:435-494).  All three write disjoint ghost cells and read valid cells
// only, so one launch performs them all.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) k_fill_ghosts(const DBox* __restrict__ boxes, const Patch* __restrict__ patches,
                                                     const PatchWork* __restrict__ work, const double* __restrict__ bcv,
                                                     T* __restrict__ phi, int doCF, int homog, double dx, double dxCrse) {
  const PatchWork w = work[blockIdx.x];
  const Patch p = patches[w.patch];
  // doCF: 0 exchange and BC, 1 exchange, BC and coarse-fine, 2 coarse-fine only,
  // 3 exchange only, 4 BC only
  const bool isCF = (p.kind == PK_CFHOMO || p.kind == PK_CFHOMOLIN);
  const bool isEx = p.kind == PK_EXCHANGE;
  if (isCF && doCF != 1 && doCF != 2) return;
  if (!isCF && doCF == 2) return;
  if (!isEx && doCF == 3) return;
  if ((isEx || isCF) && doCF == 4) return;
  const int nx = p.hi[0] - p.lo[0] + 1, ny = p.hi[1] - p.lo[1] + 1, nz = p.hi[2] - p.lo[2] + 1;
  const int q = w.start + threadIdx.x;
  if (q >= nx * ny * nz) return;
  const int i = p.lo[0] + q % nx, j = p.lo[1] + (q / nx) % ny, k = p.lo[2] + q / (nx * ny);
  const DBox bd = boxes[p.dst];
  const long long g = cidx(bd, i, j, k);
  if (p.kind == PK_EXCHANGE) {
    const DBox bs = boxes[p.src];
    phi[g] = phi[cidx(bs, i + p.sh[0], j + p.sh[1], k + p.sh[2])];
    return;
  }
  const int isign = p.side ? 1 : -1;
  int a[3] = {i, j, k};
  a[p.dir] -= isign;  // one cell inward
  const T nearv = phi[cidx(bd, a[0], a[1], a[2])];
  // the face value: the constant of the face, or this ghost cell's value from the
  // BC value function
  const double val = (homog || p.voff < 0) ? p.val : bcv[p.voff + q];
  if (p.kind == PK_DIRICHLET) {
    T inhomogVal = T(0.0);
    if (!homog) inhomogVal = T(val);
    phi[g] = T(2) * inhomogVal - nearv;
  } else if (p.kind == PK_DIRICHLET2) {  // DiriBC order 2 (BCFunc.cpp:34-40, 291-354)
    int b2[3] = {a[0], a[1], a[2]};
    b2[p.dir] -= isign;  // two cells inward
    T inhomogVal = T(0.0);
    if (!homog) inhomogVal = T(val);
    const T farv = phi[cidx(bd, b2[0], b2[1], b2[2])];
    phi[g] = T(8.0 / 3.0) * inhomogVal + T(1.0 / 3.0) * farv - T(2) * nearv;
  } else if (p.kind == PK_NEUMANN) {
    T v = nearv;
    if (!homog) v += T(double(isign) * dx) * T(val);
    phi[g] = v;
  } else if (p.kind == PK_CFHOMO) {
    int b2[3] = {a[0], a[1], a[2]};
    b2[p.dir] -= isign;  // two cells inward
    const T c1 = T(2.0 * (dxCrse - dx) / (dxCrse + dx));
    const T c2 = T(-((dxCrse - dx) / (dxCrse + 3.0 * dx)));
    const T pb = phi[cidx(bd, b2[0], b2[1], b2[2])];
    phi[g] = c1 * nearv + c2 * pb;
  } else {  // PK_CFHOMOLIN
    const T factor = T(1.0 - 2.0 * dx / (dx + dxCrse));
    phi[g] = factor * nearv;
  }
}
template <class T>
void fill_ghosts(const DeviceLayout& L, const GhostPlan& P, T* phi, int doCF, int homog, double dx, double dxCrse,
                 int comp) {
  const double* bcv = P.d_bcv ? P.d_bcv + (long long)comp * P.nbcv : nullptr;
  LAUNCH_P(KC_GHOST, 0.0, k_fill_ghosts<T>, P.nwork, L.d_boxes, P.d_patches, P.d_work, bcv, phi, doCF, homog, dx,
           dxCrse);
}

// ---------------------------------------------------------------------------
// GSRBLAPLACIAN (AMRPoissonOpF.ChF:122-198) and GSRBHELMHOLTZ (:53-120): one colour
// over all boxes; colour by global parity of i+j+k.
// ---------------------------------------------------------------------------
template <class T, bool LAP>
__global__ void __launch_bounds__(256) k_gsrb(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                              TileGeom g, T* __restrict__ phi, const T* __restrict__ rhs,
                                              int color, T dxinv, T lambda, T alpha, T beta) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  if ((i + j + k + color) & 1) return;
  const long long c = cidx(b, i, j, k);
  const long long sy = b.sy, sz = b.sz;
  const T s6 = phi[c + 1] + phi[c - 1] + phi[c + sy] + phi[c - sy] + phi[c + sz] + phi[c - sz];
  if (LAP) {
    const T lphi = s6 * dxinv;
    phi[c] = lambda * (rhs[c] - lphi);
  } else {
    const T p = phi[c];
    const T lphi = (s6 - T(6.0) * p) * dxinv;
    const T helmop = alpha * p + beta * lphi;
    phi[c] = p + lambda * (helmop - rhs[c]);
  }
}
template <class T>
void gsrb(const DeviceLayout& L, T* phi, const T* rhs, int color, double dx, double alpha, double beta) {
  double dxinv = 1.0 / (dx * dx);
  double sum_b = 0.0;
  for (int d = 0; d < 3; d++) sum_b = sum_b + 2.0 * dxinv;
  const bool lap = (alpha == 0.0 && beta == 1.0);
  double lambda = lap ? -1.0 / sum_b : -1.0 / (alpha - beta * sum_b);
  if (lap)
    LAUNCH_PU(KC_GSRB, 12.0 * sizeof(T) / 8.0 * L.valid_cells, 0.5 * L.valid_cells, (k_gsrb<T, true>), L.ntiles, L.d_boxes, L.d_tiles, geom(L), phi, rhs, color, T(dxinv), T(lambda), T(alpha), T(beta));
  else
    LAUNCH_PU(KC_GSRB, 12.0 * sizeof(T) / 8.0 * L.valid_cells, 0.5 * L.valid_cells, (k_gsrb<T, false>), L.ntiles, L.d_boxes, L.d_tiles, geom(L), phi, rhs, color, T(dxinv), T(lambda), T(alpha), T(beta));
}

// One GSRB colour over a region of every box, each box reading its neighbours'
// values from snap (per-box FABs hold copies made at the last ghost refresh, so
// on merged storage the cells outside the box come from a snapshot; on per-box
// storage snap is phi itself, whose ghost cells a pass never writes).  region 0:
// the box (looseGSRB, AMRPoissonOp.cpp:1333-1413); 1: the box grown by -1; 2: the
// one-cell layer of the box at side `side` of direction `dir`, shrunk by one at
// both ends in the directions below dir (overlapGSRB, :1416-1521).  Tiles cover
// each box.
template <class T, bool LAP>
__global__ void __launch_bounds__(256) k_gsrb_region(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                                     TileGeom g, T* phi, const T* snap, const T* __restrict__ rhs,
                                                     int color, int region, int dir, int side, T dxinv, T lambda,
                                                     T alpha, T beta) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  if ((i + j + k + color) & 1) return;
  const int x[3] = {i, j, k};
  if (region == 1) {
    for (int d = 0; d < 3; d++)
      if (x[d] <= b.lo[d] || x[d] >= b.hi[d]) return;
  } else if (region == 2) {
    if (x[dir] != (side ? b.hi[dir] : b.lo[dir])) return;
    for (int e = 0; e < dir; e++)
      if (x[e] <= b.lo[e] || x[e] >= b.hi[e]) return;
  }
  const long long c = cidx(b, i, j, k);
  const long long sy = b.sy, sz = b.sz;
  const T xp = i < b.hi[0] ? phi[c + 1] : snap[c + 1];
  const T xm = i > b.lo[0] ? phi[c - 1] : snap[c - 1];
  const T yp = j < b.hi[1] ? phi[c + sy] : snap[c + sy];
  const T ym = j > b.lo[1] ? phi[c - sy] : snap[c - sy];
  const T zp = k < b.hi[2] ? phi[c + sz] : snap[c + sz];
  const T zm = k > b.lo[2] ? phi[c - sz] : snap[c - sz];
  const T s6 = xp + xm + yp + ym + zp + zm;
  if (LAP) {
    const T lphi = s6 * dxinv;
    phi[c] = lambda * (rhs[c] - lphi);
  } else {
    const T p = phi[c];
    const T lphi = (s6 - T(6.0) * p) * dxinv;
    const T helmop = alpha * p + beta * lphi;
    phi[c] = p + lambda * (helmop - rhs[c]);
  }
}
template <class T>
void gsrb_region(const DeviceLayout& L, const Tile* tiles, int ntiles, T* phi, const T* snap, const T* rhs, int color,
                 int region, int dir, int side, bool lapOnly, double dx, double alpha, double beta) {
  double dxinv = 1.0 / (dx * dx);
  double sum_b = 0.0;
  for (int d = 0; d < 3; d++) sum_b = sum_b + 2.0 * dxinv;
  const bool lap = lapOnly || (alpha == 0.0 && beta == 1.0);
  double lambda = lap ? -1.0 / sum_b : -1.0 / (alpha - beta * sum_b);
  if (lap)
    LAUNCH_PU(KC_GSRB, 0.0, 0.0, (k_gsrb_region<T, true>), ntiles, L.d_boxes, tiles, geom(L), phi, snap, rhs, color,
              region, dir, side, T(dxinv), T(lambda), T(alpha), T(beta));
  else
    LAUNCH_PU(KC_GSRB, 0.0, 0.0, (k_gsrb_region<T, false>), ntiles, L.d_boxes, tiles, geom(L), phi, snap, rhs, color,
              region, dir, side, T(dxinv), T(lambda), T(alpha), T(beta));
}

// Eight-colour point updates: colour (cx, cy, cz) is the set of cells whose
// index parities are (i & 1, j & 1, k & 1) -- in two's complement i & 1 is the
// parity of |i|, the test the reference's coloured boxes use.
//
// GSMCAMRPOP (AMRPoissonOpF.ChF:10-51): phi += (rhs - lap phi) / (alpha - beta sum_b),
// the Laplacian summed direction by direction as (phi+ + phi-) - 2 phi.
template <class T>
__global__ void __launch_bounds__(256) k_gsmc(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                              TileGeom g, T* __restrict__ phi, const T* __restrict__ rhs,
                                              int cx, int cy, int cz, T dxinv, T lambda) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  if ((i & 1) != cx || (j & 1) != cy || (k & 1) != cz) return;
  const long long c = cidx(b, i, j, k);
  const long long sy = b.sy, sz = b.sz;
  const T p = phi[c];
  T lphi = ((phi[c + 1] + phi[c - 1] - T(2.0) * p) + (phi[c + sy] + phi[c - sy] - T(2.0) * p)) +
           (phi[c + sz] + phi[c - sz] - T(2.0) * p);
  lphi = lphi * dxinv;
  phi[c] = p + lambda * (rhs[c] - lphi);
}
template <class T>
void gsmc(const DeviceLayout& L, T* phi, const T* rhs, const int color[3], double dx, double alpha, double beta) {
  double dxinv = 1.0 / (dx * dx);
  double sum_b = 0.0;
  for (int d = 0; d < 3; d++) sum_b = sum_b + 2.0 * dxinv;
  double lambda = 1.0 / (alpha - beta * sum_b);
  LAUNCH_PU(KC_GSRB, 3.0 * sizeof(T) / 8.0 * L.valid_cells, 0.125 * L.valid_cells, (k_gsmc<T>), L.ntiles,
            L.d_boxes, L.d_tiles, geom(L), phi, rhs, color[0], color[1], color[2], T(dxinv), T(lambda));
}
// GSRBLAZY (AMRPoissonOpF.ChF:200-232): phi -= lambda (rhs - lphi), lambda = -1 / (alpha - beta sum_b),
// with lphi the operator applied to the whole level beforehand
template <class T>
__global__ void __launch_bounds__(256) k_gsrb_lazy(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                                   TileGeom g, T* __restrict__ phi, const T* __restrict__ lphi,
                                                   const T* __restrict__ rhs, int cx, int cy, int cz, T lambda) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  if ((i & 1) != cx || (j & 1) != cy || (k & 1) != cz) return;
  const long long c = cidx(b, i, j, k);
  phi[c] = phi[c] - lambda * (rhs[c] - lphi[c]);
}
template <class T>
void gsrb_lazy(const DeviceLayout& L, T* phi, const T* lphi, const T* rhs, const int color[3], double dx,
               double alpha, double beta) {
  double dxinv = 1.0 / (dx * dx);
  double sum_b = 0.0;
  for (int d = 0; d < 3; d++) sum_b = sum_b + 2.0 * dxinv;
  double lambda = -1.0 / (alpha - beta * sum_b);
  LAUNCH_PU(KC_GSRB, 4.0 * sizeof(T) / 8.0 * L.valid_cells, 0.125 * L.valid_cells, (k_gsrb_lazy<T>), L.ntiles,
            L.d_boxes, L.d_tiles, geom(L), phi, lphi, rhs, color[0], color[1], color[2], T(lambda));
}

// ---------------------------------------------------------------------------
// One whole GSRB iteration in one pass over HBM, on a merged level.
//
// The reference's levelGSRB (AMRPoissonOp.cpp:1216-1281) runs, for each colour,
// a ghost refresh (homogeneousCFInterp, exchange, homogeneous BC) and then
// FORT_GSRBLAPLACIAN / FORT_GSRBHELMHOLTZ on the cells of that colour.  On a
// merged level the exchange is the identity and each ghost cell outside the
// rectangle is the face ghost of exactly one valid cell c, with a value that
// depends only on c and on the cell one further inward, taken at the moment of
// the refresh.  Those ghosts are read only when c itself is updated, so this
// kernel forms them in registers at that moment, from the same values and with
// the same expressions as k_fill_ghosts; every update therefore sees exactly the
// operands the two-launch sequence gives it, and the result is bitwise the same.
//
// Blocking: a workgroup owns a 64 x 8 column of cells over a chunk of planes
// and marches up in z.  Step k loads plane k+2 of the old values into one of
// five LDS plane buffers (the column grown by 2 in x and y), updates the red
// cells of plane k+1 on the column grown by 1 (the ring is recomputed by the
// neighbouring workgroups too: every workgroup reads old values only), then the
// black cells of plane k, and stores plane k-1 (both colours final) to dst.
// Reads are old values from src, writes go to dst: workgroups never see each
// other's results, so no ordering between them is needed.  HBM traffic is one
// read of phi and rhs and one write of phi per cell (24 B in FP64).
// ---------------------------------------------------------------------------
template <class T, bool LAP, int FX, int FY, int NT>
__global__ void __launch_bounds__(NT) k_gsrb_fused(FusedArgs a, T* __restrict__ dst, const T* __restrict__ src,
                                                    const T* __restrict__ rhs, T dxinv, T lambda, T alpha, T beta,
                                                    T c1, T c2) {
  const int w = xcd_contiguous_block();
  if (w >= a.nwork) return;
  fused_column<T, LAP, FX, FY, NT>(a, w, dst, src, rhs, dxinv, lambda, alpha, beta, c1, c2);
}
// the same relaxation with the z neighbours in registers (zmarch.hpp)
template <class T, bool LAP, int FX, int FY, int NT>
__global__ void __launch_bounds__(NT) k_gsrb_zmarch(FusedArgs a, T* __restrict__ dst, const T* __restrict__ src,
                                                     const T* __restrict__ rhs, T dxinv, T lambda, T alpha, T beta,
                                                     T c1, T c2) {
  const int w = xcd_contiguous_block();
  if (w >= a.nwork) return;
  zmarch_column<T, LAP, FX, FY, NT>(a, w, dst, src, rhs, dxinv, lambda, alpha, beta, c1, c2);
}
// the register march with x quads (FP32)
template <class T, bool LAP, int FX, int FY, int NT>
__global__ void __launch_bounds__(NT) k_gsrb_zmarch4(FusedArgs a, T* __restrict__ dst, const T* __restrict__ src,
                                                      const T* __restrict__ rhs, T dxinv, T lambda, T alpha, T beta,
                                                      T c1, T c2) {
  static_assert(sizeof(T) == 4, "x-quad tiles are FP32 only");
  const int w = xcd_contiguous_block();
  if (w >= a.nwork) return;
  zmarch4_column<T, LAP, FX, FY, NT>(a, w, dst, src, rhs, dxinv, lambda, alpha, beta, c1, c2);
}

constexpr int kNumFusedShapes = kFusedShapeCount;
static_assert(sizeof(kFusedShapeTable) / sizeof(kFusedShapeTable[0]) == kNumFusedShapes, "shape table");

template <class T, bool LAP>
static void launch_fused(Context* ctx, int shape, int grid, const FusedArgs& a, T* dst, const T* src, const T* rhs,
                         T dxinv, T lambda, T alpha, T beta, T c1, T c2, double bytes, double units) {
  hipEvent_t ev = prof_begin(ctx, KC_GSRB);
#define FUSED_CASE(n, X, Y, NT)                                                                               \
  case n:                                                                                                     \
    hipLaunchKernelGGL((k_gsrb_fused<T, LAP, X, Y, NT>), dim3(grid), dim3(NT), 0, ctx->stream, a, dst, src, rhs, \
                       dxinv, lambda, alpha, beta, c1, c2);                                                   \
    break;
#define ZMARCH_CASE(n, X, Y, NT)                                                                                \
  case n:                                                                                                       \
    hipLaunchKernelGGL((k_gsrb_zmarch<T, LAP, X, Y, NT>), dim3(grid), dim3(NT), 0, ctx->stream, a, dst, src, rhs, \
                       dxinv, lambda, alpha, beta, c1, c2);                                                     \
    break;
  switch (shape) {
    FUSED_CASE(0, 64, 8, 256)
    FUSED_CASE(1, 32, 8, 256)
    FUSED_CASE(2, 64, 4, 256)
    FUSED_CASE(3, 32, 16, 256)
    FUSED_CASE(4, 128, 4, 256)
    FUSED_CASE(5, 64, 16, 256)
    FUSED_CASE(6, 32, 16, 512)
    FUSED_CASE(7, 64, 16, 512)
    FUSED_CASE(8, 64, 32, 1024)
    FUSED_CASE(9, 128, 8, 1024)
    FUSED_CASE(10, 64, 13, 512)
    FUSED_CASE(11, 32, 13, 256)
    FUSED_CASE(12, 64, 29, 1024)
    ZMARCH_CASE(13, 64, 26, 1024)
    ZMARCH_CASE(14, 32, 24, 512)
    ZMARCH_CASE(15, 64, 11, 512)
    ZMARCH_CASE(16, 32, 10, 256)
    ZMARCH_CASE(17, 128, 11, 1024)
    ZMARCH_CASE(18, 16, 21, 256)
    default:
      if constexpr (sizeof(T) == 4) {
#define ZMARCH4_CASE(n, X, Y, NT)                                                                                 \
  case n:                                                                                                         \
    hipLaunchKernelGGL((k_gsrb_zmarch4<T, LAP, X, Y, NT>), dim3(grid), dim3(NT), 0, ctx->stream, a, dst, src, rhs, \
                       dxinv, lambda, alpha, beta, c1, c2);                                                       \
    break;
        switch (shape) {
          ZMARCH4_CASE(19, 64, 11, 256)
          ZMARCH4_CASE(20, 128, 11, 512)
          ZMARCH4_CASE(21, 64, 26, 512)
          ZMARCH4_CASE(22, 32, 24, 256)
          ZMARCH4_CASE(23, 128, 26, 1024)
          ZMARCH4_CASE(24, 64, 56, 1024)
          ZMARCH4_CASE(25, 16, 47, 256)
          default: throw std::runtime_error("gsrb_fused: bad tile shape");
        }
#undef ZMARCH4_CASE
      } else {
        throw std::runtime_error("gsrb_fused: bad tile shape for FP64");
      }
  }
#undef FUSED_CASE
#undef ZMARCH_CASE
  hip_check(hipGetLastError(), "k_gsrb_fused");
  prof_end(ctx, ev, KC_GSRB, bytes, units);
}

static FusedArgs fused_args(const DeviceLayout& L, const FaceRules& fr, int shape, int kc, bool zero_src,
                            int elem_bytes) {
  if (!L.merged) throw std::runtime_error("gsrb_fused: level is not stored merged");
  if (shape < 0 || shape >= kNumFusedShapes) throw std::runtime_error("gsrb_fused: bad tile shape");
  if (!fused_shape_ok(shape, elem_bytes, L.rect.len(0)))
    throw std::runtime_error("gsrb_fused: x-quad tile shapes need FP32 and an x extent that is a multiple of 4");
  const int FX = kFusedShapeTable[shape][0], FY = kFusedShapeTable[shape][1];
  const DBox& R = L.h_boxes[L.nbox];
  FusedArgs a;
  for (int q = 0; q < 3; q++) { a.lo[q] = R.lo[q]; a.hi[q] = R.hi[q]; }
  for (int d = 0; d < 3; d++)
    for (int s = 0; s < 2; s++) {
      if (fr.kind[d][s] < 0 || fr.kind[d][s] > 2) throw std::runtime_error("gsrb_fused: bad face rule");
      a.kind[d][s] = fr.kind[d][s];
    }
  a.off = R.off;
  a.sy = R.sy;
  a.sz = R.sz;
  check_pair_rows(a);
  if (kFusedShapeTable[shape][3] == 2) check_quad_rows(a);
  a.zsrc = zero_src ? 1 : 0;
  a.ntx = (L.rect.len(0) + FX - 1) / FX;
  a.nty = (L.rect.len(1) + FY - 1) / FY;
  const int nz = L.rect.len(2);
  if (kc <= 0) {
    // z-chunks: long enough to amortise the 2-plane prologue, enough of them to fill the chip
    kc = 32;
    while (kc > 4 && (long long)a.ntx * a.nty * ((nz + kc - 1) / kc) < 2048) kc >>= 1;
  }
  a.kc = kc;
  const long long nwork = (long long)a.ntx * a.nty * ((nz + kc - 1) / kc);
  if (nwork > INT32_MAX - 8) throw std::runtime_error("gsrb_fused: level too large");
  a.nwork = (int)nwork;
  return a;
}

template <class T>
void gsrb_fused(const DeviceLayout& L, const FaceRules& fr, T* dst, const T* src, const T* rhs, double dx,
                double alpha, double beta, int shape, int kc, bool zero_src) {
  if (dst == src || dst == rhs) throw std::runtime_error("gsrb_fused: dst must be a separate field");
  check_pair_pointer(dst, "gsrb_fused");
  check_pair_pointer(src, "gsrb_fused");
  const FusedArgs a = fused_args(L, fr, shape, kc, zero_src, (int)sizeof(T));
  const int grid = (a.nwork + 7) / 8 * 8;
  double dxinv = 1.0 / (dx * dx);
  double sum_b = 0.0;
  for (int d = 0; d < 3; d++) sum_b = sum_b + 2.0 * dxinv;
  const bool lap = (alpha == 0.0 && beta == 1.0);
  double lambda = lap ? -1.0 / sum_b : -1.0 / (alpha - beta * sum_b);
  const T c1 = T(fr.c1), c2 = T(fr.c2);
  // algorithmic bytes: read phi (not when it is known zero), read rhs, write phi
  const double bytes = (zero_src ? 16.0 : 24.0) * sizeof(T) / 8.0 * L.valid_cells;
  const double units = (double)L.valid_cells;  // one cell-relaxation per valid cell
  if (lap)
    launch_fused<T, true>(L.ctx, shape, grid, a, dst, src, rhs, T(dxinv), T(lambda), T(alpha), T(beta), c1, c2, bytes,
                          units);
  else
    launch_fused<T, false>(L.ctx, shape, grid, a, dst, src, rhs, T(dxinv), T(lambda), T(alpha), T(beta), c1, c2,
                           bytes, units);
}

// Times every tile shape (and z-chunk lengths) of the one-launch relaxation on
// this level's own fields and returns the fastest as {shape, kc}; times[] (if
// given, kNumFusedShapes x 4 entries, ms per launch) receives all timings.
template <class T>
std::pair<int, int> tune_gsrb_fused(const DeviceLayout& L, const FaceRules& fr, T* dst, const T* src, const T* rhs,
                                    double dx, double alpha, double beta, double* times) {
  const int kcs[4] = {0, 8, 16, 32};
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  double best = 1e300;
  std::pair<int, int> pick{0, 0};
  for (int sh = 0; sh < kNumFusedShapes; sh++)
    for (int q = 0; q < 4; q++) {
      if (!fused_shape_ok(sh, (int)sizeof(T), L.rect.len(0))) {
        if (times) times[sh * 4 + q] = -1.0;  // does not apply to this level
        continue;
      }
      gsrb_fused<T>(L, fr, dst, src, rhs, dx, alpha, beta, sh, kcs[q]);  // warm
      hip_check(hipEventRecord(e0, L.ctx->stream), "record");
      const int reps = 3;
      for (int r = 0; r < reps; r++) gsrb_fused<T>(L, fr, dst, src, rhs, dx, alpha, beta, sh, kcs[q]);
      hip_check(hipEventRecord(e1, L.ctx->stream), "record");
      hip_check(hipEventSynchronize(e1), "sync");
      float ms = 0;
      hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
      const double t = ms / reps;
      if (times) times[sh * 4 + q] = t;
      if (t < best) { best = t; pick = {sh, kcs[q]}; }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return pick;
}

// OPERATORLAP (AMRPoissonOpF.ChF:287-324)
template <class T>
__global__ void __launch_bounds__(256) k_apply(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                               TileGeom g, T* __restrict__ lofphi, const T* __restrict__ phi,
                                               T dxinv, T alpha, T beta) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  const long long sy = b.sy, sz = b.sz;
  const T p = phi[c];
  const T lap = (phi[c + 1] + phi[c - 1] + phi[c + sy] + phi[c - sy] + phi[c + sz] + phi[c - sz] - T(6.0) * p) * dxinv;
  lofphi[c] = alpha * p + beta * lap;
}
// the same over a merged level's column tiles, marching in z: the z neighbours come
// from the planes already in registers, so each plane of phi is read once
template <class T>
__global__ void __launch_bounds__(256) k_apply_z(const DBox* __restrict__ boxes, const Tile* __restrict__ ctiles,
                                                 int nct, int kz, T* __restrict__ lofphi, const T* __restrict__ phi,
                                                 T dxinv, T alpha, T beta) {
  const int w = xcd_contiguous_block();
  if (w >= nct) return;
  const Tile t = ctiles[w];
  const DBox b = boxes[t.box];
  const int i = t.i0 + (threadIdx.x & 63), j = t.j0 + (threadIdx.x >> 6);
  if (i > b.hi[0] || j > b.hi[1]) return;
  const int kend = min(t.k0 + kz - 1, b.hi[2]);
  const long long sy = b.sy, sz = b.sz;
  long long c = cidx(b, i, j, t.k0);
  T pm = phi[c - sz], p = phi[c];
  for (int k = t.k0; k <= kend; k++, c += sz) {
    const T pz = phi[c + sz];
    const T lap = (phi[c + 1] + phi[c - 1] + phi[c + sy] + phi[c - sy] + pz + pm - T(6.0) * p) * dxinv;
    lofphi[c] = alpha * p + beta * lap;
    pm = p;
    p = pz;
  }
}
template <class T>
void apply(const DeviceLayout& L, T* lofphi, const T* phi, double dx, double alpha, double beta) {
  double dxinv = 1.0 / (dx * dx);
  if (L.merged && L.nctiles > 0) {
    LAUNCH_P(KC_APPLY, 2.0 * sizeof(T) * L.valid_cells, k_apply_z<T>, (L.nctiles + 7) / 8 * 8, L.d_boxes, L.d_ctiles,
             L.nctiles, DeviceLayout::kColumnPlanes, lofphi, phi, T(dxinv), T(alpha), T(beta));
    return;
  }
  LAUNCH_P(KC_APPLY, 2.0 * sizeof(T) * L.valid_cells, k_apply<T>, L.ntiles, L.d_boxes, L.d_tiles, geom(L), lofphi, phi, T(dxinv), T(alpha), T(beta));
}

// AMRResidual's operator and axby in one pass: res = (-1) * (alpha*phi + beta*lap) +
// rhs * 1, OPERATORLAP's expression followed by LevelDataOps::axby's, on valid cells
// (the cells that reflux corrects are rewritten by k_reflux_residual)
template <class T>
__device__ __forceinline__ T apply_axby(T lofphi, T r) {
  T v = lofphi;
  v = v * T(-1.0);
  return v + r * T(1.0);
}
template <class T>
__global__ void __launch_bounds__(256) k_apply_residual(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                                        TileGeom g, T* __restrict__ res, const T* __restrict__ phi,
                                                        const T* __restrict__ rhs, T dxinv, T alpha, T beta) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  const long long sy = b.sy, sz = b.sz;
  const T p = phi[c];
  const T lap = (phi[c + 1] + phi[c - 1] + phi[c + sy] + phi[c - sy] + phi[c + sz] + phi[c - sz] - T(6.0) * p) * dxinv;
  res[c] = apply_axby(alpha * p + beta * lap, rhs[c]);
}
template <class T>
__global__ void __launch_bounds__(256) k_apply_residual_z(const DBox* __restrict__ boxes,
                                                          const Tile* __restrict__ ctiles, int nct, int kz,
                                                          T* __restrict__ res, const T* __restrict__ phi,
                                                          const T* __restrict__ rhs, T dxinv, T alpha, T beta) {
  const int w = xcd_contiguous_block();
  if (w >= nct) return;
  const Tile t = ctiles[w];
  const DBox b = boxes[t.box];
  const int i = t.i0 + (threadIdx.x & 63), j = t.j0 + (threadIdx.x >> 6);
  if (i > b.hi[0] || j > b.hi[1]) return;
  const int kend = min(t.k0 + kz - 1, b.hi[2]);
  const long long sy = b.sy, sz = b.sz;
  long long c = cidx(b, i, j, t.k0);
  T pm = phi[c - sz], p = phi[c];
  for (int k = t.k0; k <= kend; k++, c += sz) {
    const T pz = phi[c + sz];
    const T lap = (phi[c + 1] + phi[c - 1] + phi[c + sy] + phi[c - sy] + pz + pm - T(6.0) * p) * dxinv;
    res[c] = apply_axby(alpha * p + beta * lap, rhs[c]);
    pm = p;
    p = pz;
  }
}
template <class T>
void apply_residual(const DeviceLayout& L, T* res, const T* phi, const T* rhs, double dx, double alpha, double beta) {
  double dxinv = 1.0 / (dx * dx);
  if (L.merged && L.nctiles > 0) {
    LAUNCH_P(KC_APPLY, 3.0 * sizeof(T) * L.valid_cells, k_apply_residual_z<T>, (L.nctiles + 7) / 8 * 8, L.d_boxes,
             L.d_ctiles, L.nctiles, DeviceLayout::kColumnPlanes, res, phi, rhs, T(dxinv), T(alpha), T(beta));
    return;
  }
  LAUNCH_P(KC_APPLY, 3.0 * sizeof(T) * L.valid_cells, k_apply_residual<T>, L.ntiles, L.d_boxes, L.d_tiles, geom(L), res,
           phi, rhs, T(dxinv), T(alpha), T(beta));
}

// OPERATORLAPRES (AMRPoissonOpF.ChF:341-377): r = -alpha*phi - beta*lap + rhs
template <class T>
__global__ void __launch_bounds__(256) k_residual(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                                  TileGeom g, T* res, const T* __restrict__ phi, const T* rhs,
                                                  T dxinv, T alpha, T beta) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  const long long sy = b.sy, sz = b.sz;
  const T p = phi[c];
  const T lap = (phi[c + 1] + phi[c - 1] + phi[c + sy] + phi[c - sy] + phi[c + sz] + phi[c - sz] - T(6.0) * p) * dxinv;
  res[c] = -(alpha * p) - beta * lap + rhs[c];
}
// the same over a merged level's column tiles, marching in z (as k_apply_z).
// nrm (optional): the max-norm of what is written is folded into *nrm as k_maxnorm
// folds it (per-block max, then an atomic max of the non-negative bit pattern, skipped
// when the block's max does not exceed the value already there)
template <class T>
__global__ void __launch_bounds__(256) k_residual_z(const DBox* __restrict__ boxes, const Tile* __restrict__ ctiles,
                                                    int nct, int kz, T* res, const T* __restrict__ phi, const T* rhs,
                                                    T dxinv, T alpha, T beta, double* nrm) {
  const int w = xcd_contiguous_block();
  if (w >= nct) return;  // uniform over the block
  const Tile t = ctiles[w];
  const DBox b = boxes[t.box];
  const int i = t.i0 + (threadIdx.x & 63), j = t.j0 + (threadIdx.x >> 6);
  double v = 0.0;
  if (i <= b.hi[0] && j <= b.hi[1]) {
    const int kend = min(t.k0 + kz - 1, b.hi[2]);
    const long long sy = b.sy, sz = b.sz;
    long long c = cidx(b, i, j, t.k0);
    T pm = phi[c - sz], p = phi[c];
    for (int k = t.k0; k <= kend; k++, c += sz) {
      const T pz = phi[c + sz];
      const T lap = (phi[c + 1] + phi[c - 1] + phi[c + sy] + phi[c - sy] + pz + pm - T(6.0) * p) * dxinv;
      const T r = -(alpha * p) - beta * lap + rhs[c];
      res[c] = r;
      v = fmax(v, fabs((double)r));
      pm = p;
      p = pz;
    }
  }
  if (nrm) {  // uniform over the launch
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __shared__ double sm[4];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      const double m = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
      if (m > 0.0 && m > *(volatile double*)nrm)
        atomicMax((unsigned long long*)nrm, (unsigned long long)__double_as_longlong(m));
    }
  }
}
template <class T>
void residual(const DeviceLayout& L, T* res, const T* phi, const T* rhs, double dx, double alpha, double beta) {
  double dxinv = 1.0 / (dx * dx);
  if (L.merged && L.nctiles > 0) {
    LAUNCH_P(KC_RESIDUAL, 3.0 * sizeof(T) * L.valid_cells, k_residual_z<T>, (L.nctiles + 7) / 8 * 8, L.d_boxes,
             L.d_ctiles, L.nctiles, DeviceLayout::kColumnPlanes, res, phi, rhs, T(dxinv), T(alpha), T(beta),
             (double*)nullptr);
    return;
  }
  LAUNCH_P(KC_RESIDUAL, 3.0 * sizeof(T) * L.valid_cells, k_residual<T>, L.ntiles, L.d_boxes, L.d_tiles, geom(L), res, phi, rhs, T(dxinv), T(alpha), T(beta));
}

// RESTRICTRES (AMRPoissonOpF.ChF:394-438) after res.setVal(0): one thread per coarse
// cell accumulates its eight children in the Fortran loop order.
template <class T>
__global__ void __launch_bounds__(256) k_restrict(const DBox* __restrict__ cboxes, const Tile* __restrict__ ctiles,
                                                  TileGeom g, const DBox* __restrict__ fboxes, T* __restrict__ resc,
                                                  const T* __restrict__ phi, const T* __restrict__ rhs, T dxinv,
                                                  T alpha, T beta) {
  DBox cb;
  int I, J, K;
  if (!tile_cell(cboxes, ctiles, g, cb, I, J, K)) return;
  const DBox fb = fboxes[blockIdx.x >= 0 ? ctiles[blockIdx.x].box : 0];
  const long long sy = fb.sy, sz = fb.sz;
  const T denom = T(8.0);
  T acc = T(0.0);
#pragma unroll
  for (int kk = 0; kk < 2; kk++)
#pragma unroll
    for (int jj = 0; jj < 2; jj++)
#pragma unroll
      for (int ii = 0; ii < 2; ii++) {
        const long long c = cidx(fb, 2 * I + ii, 2 * J + jj, 2 * K + kk);
        const T p = phi[c];
        const T lofphi = alpha * p + beta * (phi[c + 1] + phi[c - 1] + phi[c + sy] + phi[c - sy] + phi[c + sz] +
                                             phi[c - sz] - p * T(2.0) * T(3.0)) * dxinv;
        acc = acc + (rhs[c] - lofphi) / denom;
      }
  resc[cidx(cb, I, J, K)] = acc;
}
// Kernels over two layouts index the second layout's box table with the tile's box
// index: the layouts must pair box for box (one is the coarsening of the other).
static void check_paired(const DeviceLayout& Lf, const DeviceLayout& Lc, const char* what, int ratio = 2) {
  if (ratio != 2 && ratio != 4) throw std::runtime_error(std::string(what) + ": ratio must be 2 or 4");
  const int sh = ratio == 4 ? 2 : 1;
  if (Lf.nbox != Lc.nbox || Lf.merged != Lc.merged || Lf.h_boxes.size() != Lc.h_boxes.size())
    throw std::runtime_error(std::string(what) + ": fine and coarse layouts do not pair box for box");
  for (size_t b = 0; b < Lf.h_boxes.size(); b++)
    for (int q = 0; q < 3; q++)
      if ((Lf.h_boxes[b].lo[q] >> sh) != Lc.h_boxes[b].lo[q] || (Lf.h_boxes[b].hi[q] >> sh) != Lc.h_boxes[b].hi[q])
        throw std::runtime_error(std::string(what) + ": coarse box is not the coarsened fine box");
}
template <class T>
void restrict_res(const DeviceLayout& Lf, const DeviceLayout& Lc, T* resc, const T* phi, const T* rhs, double dx,
                  double alpha, double beta) {
  const DeviceLayout& L = Lc;
  check_paired(Lf, Lc, "restrict_res");
  double dxinv = 1.0 / (dx * dx);
  LAUNCH_P(KC_RESTRICT, (2.0 + 1.0 / 8.0) * sizeof(T) * Lf.valid_cells, k_restrict<T>, Lc.ntiles, Lc.d_boxes, Lc.d_tiles, geom(Lc), Lf.d_boxes, resc, phi, rhs, T(dxinv), T(alpha),
         T(beta));
}

// PROLONG (AMRPoissonOpF.ChF:453-478), refinement ratio 1 << SH (floor division of
// the fine index)
template <class T, int SH>
__global__ void __launch_bounds__(256) k_prolong(const DBox* __restrict__ fboxes, const Tile* __restrict__ ftiles,
                                                 TileGeom g, const DBox* __restrict__ cboxes, T* __restrict__ phi,
                                                 const T* __restrict__ coarse) {
  DBox fb;
  int i, j, k;
  if (!tile_cell(fboxes, ftiles, g, fb, i, j, k)) return;
  const DBox cb = cboxes[ftiles[blockIdx.x].box];
  const long long c = cidx(fb, i, j, k);
  phi[c] = phi[c] + coarse[cidx(cb, i >> SH, j >> SH, k >> SH)];
}
template <class T>
void prolong(const DeviceLayout& Lf, const DeviceLayout& Lc, T* phi, const T* coarse, int ratio) {
  const DeviceLayout& L = Lf;
  check_paired(Lf, Lc, "prolong", ratio);
  const double bytes = (2.0 + 1.0 / (ratio * ratio * ratio)) * sizeof(T) * Lf.valid_cells;
  if (ratio == 4)
    LAUNCH_P(KC_PROLONG, bytes, (k_prolong<T, 2>), Lf.ntiles, Lf.d_boxes, Lf.d_tiles, geom(Lf), Lc.d_boxes, phi, coarse);
  else
    LAUNCH_P(KC_PROLONG, bytes, (k_prolong<T, 1>), Lf.ntiles, Lf.d_boxes, Lf.d_tiles, geom(Lf), Lc.d_boxes, phi, coarse);
}

// average, refRatio 2 branch (lib/src/AMRTools/AverageF.ChF:40-57)
template <class T>
__global__ void __launch_bounds__(256) k_average(const DBox* __restrict__ cboxes, const Tile* __restrict__ ctiles,
                                                 TileGeom g, const DBox* __restrict__ fboxes, T* __restrict__ coarse,
                                                 const T* __restrict__ fine) {
  DBox cb;
  int I, J, K;
  if (!tile_cell(cboxes, ctiles, g, cb, I, J, K)) return;
  const DBox fb = fboxes[ctiles[blockIdx.x].box];
  const long long c = cidx(fb, 2 * I, 2 * J, 2 * K);
  const long long sy = fb.sy, sz = fb.sz;
  const T refScale = T(1.0 / 8.0);
  coarse[cidx(cb, I, J, K)] = refScale * (fine[c] + fine[c + 1] + fine[c + sy] + fine[c + 1 + sy] + fine[c + sz] +
                                          fine[c + 1 + sz] + fine[c + sy + sz] + fine[c + 1 + sy + sz]);
}
// average, refRatio 4 branch (AverageF.ChF:60-134): the 64 children summed left to
// right, x fastest, then times refScale
template <class T>
__global__ void __launch_bounds__(256) k_average4(const DBox* __restrict__ cboxes, const Tile* __restrict__ ctiles,
                                                  TileGeom g, const DBox* __restrict__ fboxes, T* __restrict__ coarse,
                                                  const T* __restrict__ fine) {
  DBox cb;
  int I, J, K;
  if (!tile_cell(cboxes, ctiles, g, cb, I, J, K)) return;
  const DBox fb = fboxes[ctiles[blockIdx.x].box];
  const long long c = cidx(fb, 4 * I, 4 * J, 4 * K);
  const long long sy = fb.sy, sz = fb.sz;
  T sum = fine[c];
#pragma unroll
  for (int t = 1; t < 64; t++) sum = sum + fine[c + (t & 3) + ((t >> 2) & 3) * sy + (t >> 4) * sz];
  const T refScale = T(1.0 / 64.0);
  coarse[cidx(cb, I, J, K)] = refScale * sum;
}
template <class T>
void average(const DeviceLayout& Lf, const DeviceLayout& Lc, T* coarse, const T* fine, int ratio) {
  const DeviceLayout& L = Lc;
  check_paired(Lf, Lc, "average", ratio);
  const double bytes = (1.0 + 1.0 / (ratio * ratio * ratio)) * sizeof(T) * Lf.valid_cells;
  if (ratio == 4)
    LAUNCH_P(KC_AVERAGE, bytes, k_average4<T>, Lc.ntiles, Lc.d_boxes, Lc.d_tiles, geom(Lc), Lf.d_boxes, coarse, fine);
  else
    LAUNCH_P(KC_AVERAGE, bytes, k_average<T>, Lc.ntiles, Lc.d_boxes, Lc.d_tiles, geom(Lc), Lf.d_boxes, coarse, fine);
}

// AMRRestrictS's residual and average in one pass: OPERATORLAPRES on the eight fine
// cells of each coarse cell of coarsen(fine, 2), then AVERAGE's refRatio-2 sum, with
// the residuals in registers.  The same expressions in the same order as k_residual
// followed by k_average, so the coarse values are bitwise theirs; the fine residual
// itself is not stored.
template <class T>
__global__ void __launch_bounds__(256) k_residual_average(const DBox* __restrict__ cboxes,
                                                          const Tile* __restrict__ ctiles, TileGeom g,
                                                          const DBox* __restrict__ fboxes, T* __restrict__ coarse,
                                                          const T* __restrict__ phi, const T* __restrict__ rhs,
                                                          T dxinv, T alpha, T beta) {
  DBox cb;
  int I, J, K;
  if (!tile_cell(cboxes, ctiles, g, cb, I, J, K)) return;
  const DBox fb = fboxes[ctiles[blockIdx.x].box];
  const long long c = cidx(fb, 2 * I, 2 * J, 2 * K);
  const long long sy = fb.sy, sz = fb.sz;
  auto res = [&](long long q) {
    const T p = phi[q];
    const T lap = (phi[q + 1] + phi[q - 1] + phi[q + sy] + phi[q - sy] + phi[q + sz] + phi[q - sz] - T(6.0) * p) * dxinv;
    return -(alpha * p) - beta * lap + rhs[q];
  };
  const T r0 = res(c), r1 = res(c + 1), r2 = res(c + sy), r3 = res(c + 1 + sy);
  const T r4 = res(c + sz), r5 = res(c + 1 + sz), r6 = res(c + sy + sz), r7 = res(c + 1 + sy + sz);
  const T refScale = T(1.0 / 8.0);
  coarse[cidx(cb, I, J, K)] = refScale * (r0 + r1 + r2 + r3 + r4 + r5 + r6 + r7);
}
// the same for ratio 4: the 64 fine residuals of each coarse cell summed left to
// right in AVERAGE's order (x fastest), then times refScale
template <class T>
__global__ void __launch_bounds__(256) k_residual_average4(const DBox* __restrict__ cboxes,
                                                           const Tile* __restrict__ ctiles, TileGeom g,
                                                           const DBox* __restrict__ fboxes, T* __restrict__ coarse,
                                                           const T* __restrict__ phi, const T* __restrict__ rhs,
                                                           T dxinv, T alpha, T beta) {
  DBox cb;
  int I, J, K;
  if (!tile_cell(cboxes, ctiles, g, cb, I, J, K)) return;
  const DBox fb = fboxes[ctiles[blockIdx.x].box];
  const long long c = cidx(fb, 4 * I, 4 * J, 4 * K);
  const long long sy = fb.sy, sz = fb.sz;
  auto res = [&](long long q) {
    const T p = phi[q];
    const T lap = (phi[q + 1] + phi[q - 1] + phi[q + sy] + phi[q - sy] + phi[q + sz] + phi[q - sz] - T(6.0) * p) * dxinv;
    return -(alpha * p) - beta * lap + rhs[q];
  };
  T sum = res(c);
  for (int t = 1; t < 64; t++) sum = sum + res(c + (t & 3) + ((t >> 2) & 3) * sy + (t >> 4) * sz);
  const T refScale = T(1.0 / 64.0);
  coarse[cidx(cb, I, J, K)] = refScale * sum;
}
template <class T>
void residual_average(const DeviceLayout& Lf, const DeviceLayout& Lc, T* coarse, const T* phi, const T* rhs,
                      double dx, double alpha, double beta, int ratio) {
  const DeviceLayout& L = Lc;
  check_paired(Lf, Lc, "residual_average", ratio);
  double dxinv = 1.0 / (dx * dx);
  const double bytes = (2.0 + 1.0 / (ratio * ratio * ratio)) * sizeof(T) * Lf.valid_cells;
  if (ratio == 4)
    LAUNCH_P(KC_RESTRICT, bytes, k_residual_average4<T>, Lc.ntiles, Lc.d_boxes, Lc.d_tiles, geom(Lc), Lf.d_boxes,
             coarse, phi, rhs, T(dxinv), T(alpha), T(beta));
  else
    LAUNCH_P(KC_RESTRICT, bytes, k_residual_average<T>, Lc.ntiles, Lc.d_boxes, Lc.d_tiles, geom(Lc), Lf.d_boxes,
             coarse, phi, rhs, T(dxinv), T(alpha), T(beta));
}

// ---------------------------------------------------------------------------
// LevelDataOps (LevelDataOps.H:93-294)
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) k_incr(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                              TileGeom g, T* lhs, const T* x, T scale) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  lhs[c] = lhs[c] + x[c] * scale;
}
template <class T>
void incr(const DeviceLayout& L, T* lhs, const T* x, double scale) {
  LAUNCH(k_incr<T>, L.ntiles, L.d_boxes, L.d_tiles, geom(L), lhs, x, T(scale));
}
template <class T>
__global__ void __launch_bounds__(256) k_incr_assign(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                                     TileGeom g, T* lhs, T* x) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  const T v = lhs[c] + x[c] * T(1.0);
  lhs[c] = v;
  x[c] = v;
}
template <class T>
void incr_assign(const DeviceLayout& L, T* lhs, T* x) {
  LAUNCH(k_incr_assign<T>, L.ntiles, L.d_boxes, L.d_tiles, geom(L), lhs, x);
}
// The V-cycle's correction updates, each the arithmetic of a sequence of the
// reference's LevelDataOps calls in one pass over the valid cells (ghost values are
// refilled before any later read):
// setToZero(lhs) then incr(lhs, x, 1):  lhs = 0 + x * 1 (lhs not read)
template <class T>
__global__ void __launch_bounds__(256) k_incr_from_zero(const DBox* __restrict__ boxes,
                                                        const Tile* __restrict__ tiles, TileGeom g, T* lhs,
                                                        const T* x) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  lhs[c] = T(0.0) + x[c] * T(1.0);
}
template <class T>
void incr_from_zero(const DeviceLayout& L, T* lhs, const T* x) {
  LAUNCH(k_incr_from_zero<T>, L.ntiles, L.d_boxes, L.d_tiles, geom(L), lhs, x);
}
// incr(m, d, 1), assignLocal(d, m), incr(phi, d, 1), setToZero(d)
template <class T>
__global__ void __launch_bounds__(256) k_incr_assign_update(const DBox* __restrict__ boxes,
                                                            const Tile* __restrict__ tiles, TileGeom g, T* m, T* d,
                                                            T* phi) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  const T v = m[c] + d[c] * T(1.0);
  m[c] = v;
  phi[c] = phi[c] + v * T(1.0);
  d[c] = T(0.0);
}
template <class T>
void incr_assign_update(const DeviceLayout& L, T* m, T* d, T* phi) {
  LAUNCH(k_incr_assign_update<T>, L.ntiles, L.d_boxes, L.d_tiles, geom(L), m, d, phi);
}
// setToZero(u), incr(u, x, 1), incr(phi, u, 1), setToZero(u): phi += (0 + x * 1) * 1
// (u is zero before and after and is not touched)
template <class T>
__global__ void __launch_bounds__(256) k_incr_phi_from_zero(const DBox* __restrict__ boxes,
                                                            const Tile* __restrict__ tiles, TileGeom g, T* phi,
                                                            const T* x) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  const T u = T(0.0) + x[c] * T(1.0);
  phi[c] = phi[c] + u * T(1.0);
}
template <class T>
void incr_phi_from_zero(const DeviceLayout& L, T* phi, const T* x) {
  LAUNCH(k_incr_phi_from_zero<T>, L.ntiles, L.d_boxes, L.d_tiles, geom(L), phi, x);
}
template <class T>
__global__ void __launch_bounds__(256) k_copy_valid(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                                    TileGeom g, T* dst, const T* src) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  dst[c] = src[c];
}
template <class T>
void copy_valid(const DeviceLayout& L, T* dst, const T* src) {
  LAUNCH(k_copy_valid<T>, L.ntiles, L.d_boxes, L.d_tiles, geom(L), dst, src);
}

static int flat_grid(long long n) { return (int)std::min<long long>((n + 255) / 256, 1 << 20); }

template <class T>
__global__ void __launch_bounds__(256) k_axby(long long n, T* lhs, const T* x, const T* y, T a, T b) {
  for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
    T v = x[q];
    v = v * a;
    lhs[q] = v + y[q] * b;
  }
}
template <class T>
void axby(const DeviceLayout& L, T* lhs, const T* x, const T* y, double a, double b) {
  LAUNCH(k_axby<T>, flat_grid(L.slab), L.slab, lhs, x, y, T(a), T(b));
}
template <class T>
__global__ void __launch_bounds__(256) k_scale(long long n, T* x, T s) {
  for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += (long long)gridDim.x * 256) x[q] = x[q] * s;
}
template <class T>
void scale(const DeviceLayout& L, T* x, double s) {
  LAUNCH(k_scale<T>, flat_grid(L.slab), L.slab, x, T(s));
}
template <class T>
__global__ void __launch_bounds__(256) k_setval(long long n, T* x, T v) {
  for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += (long long)gridDim.x * 256) x[q] = v;
}
template <class T>
void setval(const DeviceLayout& L, T* x, double v) {
  LAUNCH(k_setval<T>, flat_grid(L.slab), L.slab, x, T(v));
}
template <class T>
__global__ void __launch_bounds__(256) k_copy_all(long long n, T* dst, const T* src) {
  for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += (long long)gridDim.x * 256) dst[q] = src[q];
}
template <class T>
void copy_all(const DeviceLayout& L, T* dst, const T* src) {
  LAUNCH(k_copy_all<T>, flat_grid(L.slab), L.slab, dst, src);
}
// preCond (AMRPoissonOp.cpp:304-329): phi = rhs on the valid region, then phi *= mult
// on the whole FAB.  Ghost cells of phi are never read before the next refresh, so
// scaling valid cells gives the same valid values.
template <class T>
__global__ void __launch_bounds__(256) k_precond_init(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                                      TileGeom g, T* phi, const T* rhs, T mult) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  phi[c] = rhs[c] * mult;
}
template <class T>
void precond_init(const DeviceLayout& L, T* phi, const T* rhs, double mult) {
  LAUNCH(k_precond_init<T>, L.ntiles, L.d_boxes, L.d_tiles, geom(L), phi, rhs, T(mult));
}
__global__ void __launch_bounds__(256) k_d2f(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                             TileGeom g, float* dst, const double* src) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  dst[c] = (float)src[c];
}
void convert_valid_d2f(const DeviceLayout& L, float* dst, const double* src) {
  LAUNCH(k_d2f, L.ntiles, L.d_boxes, L.d_tiles, geom(L), dst, src);
}
__global__ void __launch_bounds__(256) k_incr_f2d(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                                  TileGeom g, double* lhs, const float* x) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  lhs[c] = lhs[c] + (double)x[c];
}
void incr_f2d(const DeviceLayout& L, double* lhs, const float* x) {
  LAUNCH(k_incr_f2d, L.ntiles, L.d_boxes, L.d_tiles, geom(L), lhs, x);
}
// phi + (0 + (double)x) * 1: k_incr_f2d on a zero field followed by k_incr(phi, ., 1)
__global__ void __launch_bounds__(256) k_update_f2d(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                                    TileGeom g, double* phi, const float* x) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  const double u = 0.0 + (double)x[c];
  phi[c] = phi[c] + u * 1.0;
}
void update_f2d(const DeviceLayout& L, double* phi, const float* x) {
  LAUNCH(k_update_f2d, L.ntiles, L.d_boxes, L.d_tiles, geom(L), phi, x);
}
__global__ void __launch_bounds__(256) k_f2d(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                             TileGeom g, double* dst, const float* src) {
  DBox b;
  int i, j, k;
  if (!tile_cell(boxes, tiles, g, b, i, j, k)) return;
  const long long c = cidx(b, i, j, k);
  dst[c] = (double)src[c];
}
void convert_valid_f2d(const DeviceLayout& L, double* dst, const float* src) {
  LAUNCH(k_f2d, L.ntiles, L.d_boxes, L.d_tiles, geom(L), dst, src);
}

// ---------------------------------------------------------------------------
// Copier motion (copyTo / copyToZero) between two layouts
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) k_motion(const Motion* __restrict__ items, const MotionWork* __restrict__ work,
                                                const DBox* __restrict__ sboxes, const DBox* __restrict__ dboxes,
                                                T* dst, const T* src, int zero) {
  const MotionWork w = work[blockIdx.x];
  const Motion m = items[w.item];
  const int nx = m.hi[0] - m.lo[0] + 1, ny = m.hi[1] - m.lo[1] + 1, nz = m.hi[2] - m.lo[2] + 1;
  const int q = w.start + threadIdx.x;
  if (q >= nx * ny * nz) return;
  const int i = m.lo[0] + q % nx, j = m.lo[1] + (q / nx) % ny, k = m.lo[2] + q / (nx * ny);
  const long long dc = cidx(dboxes[m.to], i, j, k);
  if (zero) dst[dc] = T(0.0);
  else dst[dc] = src[cidx(sboxes[m.from], i, j, k)];
}
template <class T>
void motion(const MotionPlan& P, const DeviceLayout& Lsrc, const DeviceLayout& Ldst, T* dst, const T* src, int zero) {
  const DeviceLayout& L = Ldst;
  LAUNCH_P(KC_MOTION, 0.0, k_motion<T>, P.nwork, P.d_items, P.d_work, Lsrc.d_boxes, Ldst.d_boxes, dst, src, zero);
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
__global__ void k_zero_scalar(double* s, int slot) {
  if (threadIdx.x == 0 && blockIdx.x == 0) s[slot] = 0.0;
}
// FArrayBox::norm(p = 0) over valid boxes; max is exact, so any order gives the
// reference's value.  Non-negative doubles order like their bit patterns.
template <class T>
__global__ void __launch_bounds__(256) k_maxnorm(const DBox* __restrict__ boxes, const Tile* __restrict__ tiles,
                                                 TileGeom g, int ntiles, const T* __restrict__ x, double* out) {
  // grid-stride over tiles: one atomic per block, a few thousand blocks per launch
  double v = 0.0;
  const int tid = threadIdx.x;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const DBox b = boxes[tl.box];
    const int i = tl.i0 + (tid & ((1 << g.lx) - 1));
    const int j = tl.j0 + ((tid >> g.lx) & ((1 << g.ly) - 1));
    const int k = tl.k0 + (tid >> (g.lx + g.ly));
    if (i <= b.hi[0] && j <= b.hi[1] && k <= b.hi[2]) v = fmax(v, fabs((double)x[cidx(b, i, j, k)]));
  }
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  __shared__ double sm[4];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
    if (m > 0.0) atomicMax((unsigned long long*)out, (unsigned long long)__double_as_longlong(m));
  }
}
template <class T>
void maxnorm_async(const DeviceLayout& L, const T* x, int slot, int accumulate) {
  if (!accumulate) LAUNCH(k_zero_scalar, 1, L.ctx->d_scalars, slot);
  const int grid = std::min(L.ntiles, 2048);
  LAUNCH_P(KC_REDUCE, 1.0 * sizeof(T) * L.valid_cells, k_maxnorm<T>, grid, L.d_boxes, L.d_tiles, geom(L), L.ntiles, x,
           L.ctx->d_scalars + slot);
}
template <class T>
void residual_norm(const DeviceLayout& L, T* res, const T* phi, const T* rhs, double dx, double alpha, double beta,
                   int slot, int accumulate) {
  if (!(L.merged && L.nctiles > 0)) {
    residual<T>(L, res, phi, rhs, dx, alpha, beta);
    maxnorm_async<T>(L, res, slot, accumulate);
    return;
  }
  if (!accumulate) LAUNCH(k_zero_scalar, 1, L.ctx->d_scalars, slot);
  const double dxinv = 1.0 / (dx * dx);
  LAUNCH_P(KC_RESIDUAL, 3.0 * sizeof(T) * L.valid_cells, k_residual_z<T>, (L.nctiles + 7) / 8 * 8, L.d_boxes,
           L.d_ctiles, L.nctiles, DeviceLayout::kColumnPlanes, res, phi, rhs, T(dxinv), T(alpha), T(beta),
           L.ctx->d_scalars + slot);
}

template <class T>
double maxnorm(const DeviceLayout& L, const T* x, int slot, int nc, long long cs) {
  for (int c = 0; c < nc; c++) maxnorm_async(L, x + c * cs, slot, c > 0 ? 1 : 0);
  hip_check(hipMemcpyAsync(L.ctx->h_scalars + slot, L.ctx->d_scalars + slot, sizeof(double), hipMemcpyDeviceToHost,
                           L.ctx->stream),
            "copy scalar");
  L.ctx->sync();
  return L.ctx->h_scalars[slot];
}
// LevelDataOps::dotProduct: per box a serial sum in Fortran order, components
// outermost (FArrayBox::dotProduct, ForAllThisCBNNXC), boxes summed in layout order.
template <class T>
__global__ void __launch_bounds__(256) k_dot_ordered(const DBox* __restrict__ boxes, int nbox, const T* a, const T* b,
                                                     double* partial, double* out, int nc, long long cs) {
  for (int q = threadIdx.x; q < nbox; q += blockDim.x) {
    const DBox bx = boxes[q];
    double dot = 0.0;
    for (int n = 0; n < nc; n++)
      for (int k = bx.lo[2]; k <= bx.hi[2]; k++)
        for (int j = bx.lo[1]; j <= bx.hi[1]; j++)
          for (int i = bx.lo[0]; i <= bx.hi[0]; i++) {
            const long long c = cidx(bx, i, j, k) + n * cs;
            dot += (double)a[c] * (double)b[c];
          }
    partial[q] = dot;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double val = 0.0;
    for (int q = 0; q < nbox; q++) val += partial[q];
    *out = val;
  }
}
// norm(.., 2): per box FArrayBox::sumPow(2) (column sums over rows, then over
// columns, FArrayBox.cpp:165-230), boxes summed in layout order, then sqrt.
// ABS: norm(.., 1), FArrayBox::norm(1) (FArrayBox.cpp:110-141): the same column
// sums of |x|, no root.
template <bool ABS, class T>
__device__ __forceinline__ double sumpow_term(T v) {
  const double r = (double)v;
  return ABS ? fabs(r) : r * r;
}
template <class T, bool ABS = false>
__global__ void __launch_bounds__(256) k_sumpow_ordered(const DBox* __restrict__ boxes, int nbox, const T* x,
                                                        double* out, int nc, long long cs) {
  __shared__ double tmp[1024];
  double norm = 0.0;
  for (int q = 0; q < nbox; q++) {
    const DBox bx = boxes[q];
    const int nx = bx.hi[0] - bx.lo[0] + 1;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) {
      double t = 0.0;
      bool first = true;
      for (int n = 0; n < nc; n++)  // pencils: component, then z, then y (ForAllThisCPencil)
        for (int k = bx.lo[2]; k <= bx.hi[2]; k++)
          for (int j = bx.lo[1]; j <= bx.hi[1]; j++) {
            const double r = sumpow_term<ABS>(x[cidx(bx, bx.lo[0] + i, j, k) + n * cs]);
            if (first) { t = r; first = false; }
            else t += r;
          }
      tmp[i] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double sum = tmp[0];
      for (int i = 1; i < nx; i++) sum += tmp[i];
      norm += sum;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = norm;
}
// The same sums with every (box, column) chain in parallel, for levels whose
// nbox x (widest box) chains fit in LDS: each chain runs in its thread in the
// order above, then one thread adds the column sums box by box in the same order.
constexpr int kSumpowChains = 4096;
template <class T, bool ABS = false>
__global__ void __launch_bounds__(256) k_sumpow_chains(const DBox* __restrict__ boxes, int nbox, int W, const T* x,
                                                       double* out, int nc, long long cs) {
  __shared__ double tmp[kSumpowChains];
  for (int p = threadIdx.x; p < nbox * W; p += blockDim.x) {
    const int q = p / W, i = p % W;
    const DBox bx = boxes[q];
    if (i >= bx.hi[0] - bx.lo[0] + 1) continue;
    double t = 0.0;
    bool first = true;
    for (int n = 0; n < nc; n++)
      for (int k = bx.lo[2]; k <= bx.hi[2]; k++)
        for (int j = bx.lo[1]; j <= bx.hi[1]; j++) {
          const double r = sumpow_term<ABS>(x[cidx(bx, bx.lo[0] + i, j, k) + n * cs]);
          if (first) { t = r; first = false; }
          else t += r;
        }
    tmp[p] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double norm = 0.0;
    for (int q = 0; q < nbox; q++) {
      const int nx = boxes[q].hi[0] - boxes[q].lo[0] + 1;
      double sum = tmp[q * W];
      for (int i = 1; i < nx; i++) sum += tmp[q * W + i];
      norm += sum;
    }
    *out = norm;
  }
}
static double* scratch_partials(Context* ctx, int n) {
  static thread_local double* buf = nullptr;
  static thread_local int cap = 0;
  static thread_local Context* owner = nullptr;
  if (n > cap || owner != ctx) {
    if (buf) { (void)hipStreamSynchronize(ctx->stream); (void)hipFree(buf); }
    cap = std::max(n, 1024);
    hip_check(hipMalloc(&buf, sizeof(double) * cap), "hipMalloc partials");
    owner = ctx;
  }
  return buf;
}
template <class T>
double dot_ordered(const DeviceLayout& L, const T* a, const T* b, int slot, int nc, long long cs) {
  double* part = scratch_partials(L.ctx, L.nbox);
  hipLaunchKernelGGL(k_dot_ordered<T>, dim3(1), dim3(256), 0, L.ctx->stream, L.d_boxes, L.nbox, a, b,
                     part, L.ctx->d_scalars + slot, nc, cs);
  hip_check(hipGetLastError(), "k_dot_ordered");
  hip_check(hipMemcpyAsync(L.ctx->h_scalars + slot, L.ctx->d_scalars + slot, sizeof(double), hipMemcpyDeviceToHost,
                           L.ctx->stream),
            "copy scalar");
  L.ctx->sync();
  return L.ctx->h_scalars[slot];
}
// dot_ordered with its partial sums in the device's `scratch` array: no device
// memory is allocated or freed (the result goes through the context's scalar
// slot and its pinned host mirror, both made when the context was created)
template <class T>
double dot_ordered_noalloc(const DeviceLayout& L, const T* a, const T* b, int slot, int nc, long long cs) {
  if (L.nbox > kNoAllocBoxes)
    throw std::runtime_error("dot without allocation: the level has " + std::to_string(L.nbox) +
                             " boxes, at most " + std::to_string(kNoAllocBoxes) + " are supported");
  double* part = device_scratch(L.ctx->device);
  hipLaunchKernelGGL(k_dot_ordered<T>, dim3(1), dim3(256), 0, L.ctx->stream, L.d_boxes, L.nbox, a, b, part,
                     L.ctx->d_scalars + slot, nc, cs);
  hip_check(hipGetLastError(), "k_dot_ordered");
  hip_check(hipMemcpyAsync(L.ctx->h_scalars + slot, L.ctx->d_scalars + slot, sizeof(double), hipMemcpyDeviceToHost,
                           L.ctx->stream),
            "copy scalar");
  L.ctx->sync();
  return L.ctx->h_scalars[slot];
}
template double dot_ordered_noalloc<double>(const DeviceLayout&, const double*, const double*, int, int, long long);
template <class T, bool ABS>
static double sum_ordered(const DeviceLayout& L, const T* x, int slot, int nc, long long cs) {
  int W = 0;
  for (int q = 0; q < L.nbox; q++) {
    W = std::max(W, L.h_boxes[q].hi[0] - L.h_boxes[q].lo[0] + 1);
    if (L.h_boxes[q].hi[0] - L.h_boxes[q].lo[0] + 1 > 1024) throw std::runtime_error("norm: box wider than 1024");
  }
  if ((long long)L.nbox * W <= kSumpowChains) {
    hipLaunchKernelGGL((k_sumpow_chains<T, ABS>), dim3(1), dim3(256), 0, L.ctx->stream, L.d_boxes, L.nbox, W, x,
                       L.ctx->d_scalars + slot, nc, cs);
    hip_check(hipGetLastError(), "k_sumpow_chains");
  } else {
    hipLaunchKernelGGL((k_sumpow_ordered<T, ABS>), dim3(1), dim3(256), 0, L.ctx->stream, L.d_boxes, L.nbox, x,
                       L.ctx->d_scalars + slot, nc, cs);
    hip_check(hipGetLastError(), "k_sumpow_ordered");
  }
  hip_check(hipMemcpyAsync(L.ctx->h_scalars + slot, L.ctx->d_scalars + slot, sizeof(double), hipMemcpyDeviceToHost,
                           L.ctx->stream),
            "copy scalar");
  L.ctx->sync();
  return L.ctx->h_scalars[slot];
}
template <class T>
double norm2_ordered(const DeviceLayout& L, const T* x, int slot, int nc, long long cs) {
  return std::sqrt(sum_ordered<T, false>(L, x, slot, nc, cs));
}
template <class T>
double norm1_ordered(const DeviceLayout& L, const T* x, int slot, int nc, long long cs) {
  return sum_ordered<T, true>(L, x, slot, nc, cs);
}

// ---------------------------------------------------------------------------
// QuadCFInterp::coarseFineInterp: derivatives per coarse record
// (QuadCFStencil.cpp:36-268), then phistar (QuadCFInterp.cpp:570-764,
// QuadCFInterpF.ChF:50-126) and the quadratic (QuadCFInterp.cpp:775-839,
// QuadCFInterpF.ChF:9-47) per fine ghost cell.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T gen_sum(const CFGen& g, int q, const T* phic) {
  T s = T(0.0);
  for (int t = 0; t < g.n[q]; t++) s += T(g.w[q][t]) * phic[g.off[q][t]];
  return s;
}
template <class T>
__global__ void __launch_bounds__(256) k_cf_coarse(const CFCoarse* __restrict__ recs, int n, const CFGen* __restrict__ gens,
                                                   const T* __restrict__ phic, T* __restrict__ out, T dx) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const CFCoarse c = recs[r];
  T* o = out + 6LL * r;
  o[0] = phic[c.pc];
  if (c.std) {
    const T pcc = phic[c.pc];
    T lo = phic[c.nbr[0]], hi = phic[c.nbr[1]];
    o[1] = (hi - lo) / (T(2.0) * dx);
    o[2] = (hi + lo - T(2.0) * pcc) / (dx * dx);
    lo = phic[c.nbr[2]];
    hi = phic[c.nbr[3]];
    o[3] = (hi - lo) / (T(2.0) * dx);
    o[4] = (hi + lo - T(2.0) * pcc) / (dx * dx);
    const T ur = phic[c.nbr[4]], ul = phic[c.nbr[5]], lr = phic[c.nbr[6]], ll = phic[c.nbr[7]];
    o[5] = (ur + ll - lr - ul) / (T(4.) * dx * dx);
    return;
  }
  const CFGen g = gens[c.gen];
  // first derivatives divide by dx, second and mixed by dx*dx
  for (int q = 0; q < 5; q++) {
    T v;
    if (g.kind[q] == DK_ZERO) v = T(0.0);
    else {
      v = gen_sum(g, q, phic);
      if (q == 0 || q == 2) v /= dx;
      else v /= (dx * dx);
    }
    o[1 + q] = v;
  }
}
// np1 = nRef + 1, np3 = nRef + 3, den = nRef^2 + 4 nRef + 3 of FORT_QUADINTERP
// (QuadCFInterpF.ChF:9-47) and interpOnIVS (QuadCFInterp.cpp:775-839)
template <class T>
__global__ void __launch_bounds__(256) k_cf_fine(const CFFine* __restrict__ recs, int n, const T* __restrict__ deriv,
                                                 T* __restrict__ phif, T h, T np1, T np3, T den) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const CFFine f = recs[r];
  const T* d = deriv + 6LL * f.coarse;
  const T x1 = T(f.x1), x2 = T(f.x2);
  const T update1 = x1 * d[1] + T(0.5) * x1 * x1 * d[2];
  const T update2 = x2 * d[3] + T(0.5) * x2 * x2 * d[4];
  const T update3 = x1 * x2 * d[5];
  const T ps = d[0] + update1 + update2 + update3;
  const T pa = phif[f.a2];
  const T pb = phif[f.a1];
  T g;
  if (f.packed) {
    const T frac = T(2.0) / (h * h);
    const T denom = den;
    const T mult = frac / denom;
    const T invh = T(1.0) / h;
    const T x = T(2.0) * h;
    const T xsquared = T(4.0) * h * h;
    const T a = mult * (T(2.0) * ps + np1 * pa - np3 * pb);
    const T b = (pb - pa) * invh - a * h;
    g = xsquared * a + b * x + pa;
  } else {
    const T a = (T(2.) / h / h) * (T(2.) * ps + pa * np1 - pb * np3) / den;
    const T b = (pb - pa) / h - a * h;
    const T x = T(2.) * h;
    g = a * x * x + b * x + pa;
  }
  phif[f.g] = g;
}
template <class T>
void cf_inhomog(Context* ctx, const CFCoarse* dc, int ncoarse, const CFGen* dg, const CFFine* df, int nfine,
                T* deriv, T* phif, const T* phic, double dxf, double dxc, int nref) {
  hipEvent_t ev = prof_begin(ctx, KC_CF);
  if (ncoarse > 0) {
    hipLaunchKernelGGL(k_cf_coarse<T>, dim3((ncoarse + 255) / 256), dim3(256), 0, ctx->stream, dc, ncoarse, dg, phic,
                       deriv, T(dxc));
    hip_check(hipGetLastError(), "k_cf_coarse");
  }
  if (nfine > 0) {
    hipLaunchKernelGGL(k_cf_fine<T>, dim3((nfine + 255) / 256), dim3(256), 0, ctx->stream, df, nfine, deriv, phif,
                       T(dxf), T(nref + 1.0), T(nref + 3.0), T(nref * nref + 4 * nref + 3.0));
    hip_check(hipGetLastError(), "k_cf_fine");
  }
  prof_end(ctx, ev, KC_CF, 0.0);
}

// ---------------------------------------------------------------------------
// reflux: LevelFluxRegister::incrementCoarse/incrementFine/reflux
// (LevelFluxRegister.cpp:373-670) with NEWGETFLUX (AMRPoissonOpF.ChF:557-586)
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256) k_reflux(const RefluxCell* __restrict__ cells, int n,
                                                const RefluxCoarse* __restrict__ rc, const RefluxFine* __restrict__ rf,
                                                T* u, const T* __restrict__ phic, const T* __restrict__ phif,
                                                T beta_dx_c, T beta_ref_dx_c, T minus_s) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const RefluxCell c = cells[r];
  T val = u[c.u];
  if (c.nc > 0) {
    T reg = T(0.0);
    for (int q = c.c0; q < c.c0 + c.nc; q++) {
      const RefluxCoarse x = rc[q];
      const T F = (phic[x.a] - phic[x.b]) * beta_dx_c;
      reg += F * T(x.sign);
    }
    val = val + reg * minus_s;
  }
  // a fine register is one record (ratio 2) or a chain of first..last records of
  // four fluxes each (ratio 4), accumulated in INCREMENTFINE's order
  T reg = T(0.0);
  for (int q = c.f0; q < c.f0 + c.nf; q++) {
    const RefluxFine x = rf[q];
    if (x.first) reg = T(0.0);
    for (int t = 0; t < 4; t++) {
      const T F = (phif[x.a[t]] - phif[x.b[t]]) * beta_ref_dx_c;
      reg = reg + T(x.scale) * F;
    }
    if (x.last) val = val + reg * minus_s;
  }
  u[c.u] = val;
}
// the refluxed cells of AMRResidual: the operator value re-formed from phi (bitwise
// k_apply's), the reflux corrections of k_reflux in its order, then axby's expression
template <class T>
__global__ void __launch_bounds__(256) k_reflux_residual(const RefluxCell* __restrict__ cells, int n,
                                                         const RefluxCoarse* __restrict__ rc,
                                                         const RefluxFine* __restrict__ rf, T* res,
                                                         const T* __restrict__ phic, const T* __restrict__ phif,
                                                         const T* __restrict__ rhs, T beta_dx_c, T beta_ref_dx_c,
                                                         T minus_s, T dxinv, T alpha, T beta) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const RefluxCell c = cells[r];
  const long long q = c.u, sy = c.sy, sz = c.sz;
  const T p = phic[q];
  const T lap = (phic[q + 1] + phic[q - 1] + phic[q + sy] + phic[q - sy] + phic[q + sz] + phic[q - sz] - T(6.0) * p) *
                dxinv;
  T val = alpha * p + beta * lap;
  if (c.nc > 0) {
    T reg = T(0.0);
    for (int t = c.c0; t < c.c0 + c.nc; t++) {
      const RefluxCoarse x = rc[t];
      const T F = (phic[x.a] - phic[x.b]) * beta_dx_c;
      reg += F * T(x.sign);
    }
    val = val + reg * minus_s;
  }
  T freg = T(0.0);
  for (int t = c.f0; t < c.f0 + c.nf; t++) {
    const RefluxFine x = rf[t];
    if (x.first) freg = T(0.0);
    for (int m = 0; m < 4; m++) {
      const T F = (phif[x.a[m]] - phif[x.b[m]]) * beta_ref_dx_c;
      freg = freg + T(x.scale) * F;
    }
    if (x.last) val = val + freg * minus_s;
  }
  res[q] = apply_axby(val, rhs[q]);
}
template <class T>
void reflux_residual(Context* ctx, const RefluxCell* cells, int ncells, const RefluxCoarse* rc, const RefluxFine* rf,
                     T* res, const T* phic, const T* phif, const T* rhs, double beta_dx_c, double beta_ref_dx_c,
                     double minus_inv_dx, double dx, double alpha, double beta) {
  if (ncells <= 0) return;
  const double dxinv = 1.0 / (dx * dx);
  hipEvent_t ev = prof_begin(ctx, KC_REFLUX);
  hipLaunchKernelGGL(k_reflux_residual<T>, dim3((ncells + 255) / 256), dim3(256), 0, ctx->stream, cells, ncells, rc,
                     rf, res, phic, phif, rhs, T(beta_dx_c), T(beta_ref_dx_c), T(minus_inv_dx), T(dxinv), T(alpha),
                     T(beta));
  hip_check(hipGetLastError(), "k_reflux_residual");
  prof_end(ctx, ev, KC_REFLUX, 0.0);
}
template <class T>
void reflux(Context* ctx, const RefluxCell* cells, int ncells, const RefluxCoarse* rc, const RefluxFine* rf, T* u,
            const T* phic, const T* phif, double beta_dx_c, double beta_ref_dx_c, double minus_inv_dx) {
  if (ncells <= 0) return;
  hipEvent_t ev = prof_begin(ctx, KC_REFLUX);
  hipLaunchKernelGGL(k_reflux<T>, dim3((ncells + 255) / 256), dim3(256), 0, ctx->stream, cells, ncells, rc, rf, u, phic,
                     phif, T(beta_dx_c), T(beta_ref_dx_c), T(minus_inv_dx));
  hip_check(hipGetLastError(), "k_reflux");
  prof_end(ctx, ev, KC_REFLUX, 0.0);
}

// ---------------------------------------------------------------------------
// explicit instantiations
// ---------------------------------------------------------------------------
#define INST(T)                                                                                                   \
  template void fill_ghosts<T>(const DeviceLayout&, const GhostPlan&, T*, int, int, double, double, int);      \
  template void gsrb<T>(const DeviceLayout&, T*, const T*, int, double, double, double);                       \
  template void gsmc<T>(const DeviceLayout&, T*, const T*, const int*, double, double, double);                \
  template void gsrb_region<T>(const DeviceLayout&, const Tile*, int, T*, const T*, const T*, int, int, int, int, bool, \
                               double, double, double);                                                      \
  template void gsrb_lazy<T>(const DeviceLayout&, T*, const T*, const T*, const int*, double, double, double); \
  template void gsrb_fused<T>(const DeviceLayout&, const FaceRules&, T*, const T*, const T*, double, double,      \
                              double, int, int, bool);                                                        \
  template std::pair<int, int> tune_gsrb_fused<T>(const DeviceLayout&, const FaceRules&, T*, const T*, const T*,  \
                                                  double, double, double, double*);                          \
  template void apply<T>(const DeviceLayout&, T*, const T*, double, double, double);                           \
  template void residual<T>(const DeviceLayout&, T*, const T*, const T*, double, double, double);              \
  template void residual_norm<T>(const DeviceLayout&, T*, const T*, const T*, double, double, double, int, int); \
  template void restrict_res<T>(const DeviceLayout&, const DeviceLayout&, T*, const T*, const T*, double, double, \
                                double);                                                                      \
  template void prolong<T>(const DeviceLayout&, const DeviceLayout&, T*, const T*, int);                       \
  template void average<T>(const DeviceLayout&, const DeviceLayout&, T*, const T*, int);                       \
  template void residual_average<T>(const DeviceLayout&, const DeviceLayout&, T*, const T*, const T*, double,     \
                                    double, double, int);                                                     \
  template void apply_residual<T>(const DeviceLayout&, T*, const T*, const T*, double, double, double);        \
  template void reflux_residual<T>(Context*, const RefluxCell*, int, const RefluxCoarse*, const RefluxFine*, T*, \
                                   const T*, const T*, const T*, double, double, double, double, double, double); \
  template void incr<T>(const DeviceLayout&, T*, const T*, double);                                            \
  template void incr_assign<T>(const DeviceLayout&, T*, T*);                                                   \
  template void incr_from_zero<T>(const DeviceLayout&, T*, const T*);                                          \
  template void incr_assign_update<T>(const DeviceLayout&, T*, T*, T*);                                        \
  template void incr_phi_from_zero<T>(const DeviceLayout&, T*, const T*);                                                   \
  template void axby<T>(const DeviceLayout&, T*, const T*, const T*, double, double);                          \
  template void scale<T>(const DeviceLayout&, T*, double);                                                     \
  template void setval<T>(const DeviceLayout&, T*, double);                                                    \
  template void copy_valid<T>(const DeviceLayout&, T*, const T*);                                              \
  template void copy_all<T>(const DeviceLayout&, T*, const T*);                                                \
  template void precond_init<T>(const DeviceLayout&, T*, const T*, double);                                    \
  template void motion<T>(const MotionPlan&, const DeviceLayout&, const DeviceLayout&, T*, const T*, int);     \
  template double maxnorm<T>(const DeviceLayout&, const T*, int, int, long long);                              \
  template void maxnorm_async<T>(const DeviceLayout&, const T*, int, int);                                      \
  template double dot_ordered<T>(const DeviceLayout&, const T*, const T*, int, int, long long);                \
  template double norm2_ordered<T>(const DeviceLayout&, const T*, int, int, long long);                        \
  template double norm1_ordered<T>(const DeviceLayout&, const T*, int, int, long long);                        \
  template void cf_inhomog<T>(Context*, const CFCoarse*, int, const CFGen*, const CFFine*, int, T*, T*, const T*, \
                              double, double, int);                                                           \
  template void reflux<T>(Context*, const RefluxCell*, int, const RefluxCoarse*, const RefluxFine*, T*, const T*, \
                          const T*, double, double, double);
INST(double)
INST(float)

}  // namespace ops
}  // namespace amrmg

// ---------------------------------------------------------------------------
// plain streaming copy of known size: calibrates HBM byte counters (FETCH_SIZE,
// WRITE_SIZE) and measures the achievable copy bandwidth on the device
// ---------------------------------------------------------------------------
namespace amrmg {
__global__ void __launch_bounds__(256) k_copy_probe(double* __restrict__ dst, const double* __restrict__ src,
                                                    long long n) {
  for (long long q = blockIdx.x * 256LL + threadIdx.x; q < n; q += (long long)gridDim.x * 256) dst[q] = src[q];
}
double copy_probe(Context* c, long long nbytes, int iters) {
  long long n = nbytes / 8;
  double *a = nullptr, *b = nullptr;
  hip_check(hipMalloc(&a, n * 8), "hipMalloc probe");
  hip_check(hipMalloc(&b, n * 8), "hipMalloc probe");
  hip_check(hipMemsetAsync(a, 0, n * 8, c->stream), "memset probe");
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  int grid = (int)std::min<long long>((n + 255) / 256, 256 * 1024);
  hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, c->stream, b, a, n);  // warm
  hip_check(hipEventRecord(e0, c->stream), "record");
  for (int i = 0; i < iters; i++) hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, c->stream, b, a, n);
  hip_check(hipEventRecord(e1, c->stream), "record");
  hip_check(hipEventSynchronize(e1), "sync");
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return ms / iters;
}
}  // namespace amrmg
