// Level-wide launchers of the HIP kernels (implemented in kernels.hip).
//
// Each launcher replaces a per-box loop of FORT_* calls in the reference
// (lib/src/AMRElliptic/AMRPoissonOp.cpp) by one launch over the tile table of a
// level.  All launches go to the context's stream; none synchronises except the
// reductions that return a host scalar.
#pragma once
#include <utility>

#include "device.hpp"

namespace amrmg {
void profiler_enable(Context* c, unsigned class_mask);  // 0 = off
void profiler_reset(Context* c);
// event pair around one launch of a kernel class (no-op when that class is not timed)
hipEvent_t profile_begin(Context* c, int kernel_class);
void profile_end(Context* c, hipEvent_t begin, int kernel_class, double bytes, double units = 0.0);
bool profiler_read(Context* c, int kind, double* ms, double* bytes, long long* count, double* units = nullptr);
// average ms of one copy of nbytes (read nbytes + write nbytes) over iters launches
double copy_probe(Context* c, long long nbytes, int iters);

namespace ops {

struct GhostPlan {  // face-ghost fill rules of one level (exchange + BC + homogeneous CF)
  Patch* d_patches = nullptr;
  PatchWork* d_work = nullptr;
  double* d_bcv = nullptr;  // per-face-cell BC values (null: every rule uses its face constant)
  long long nbcv = 0;       // values per component (component c's block starts at c * nbcv)
  int npatch = 0, nwork = 0;
  Context* owner = nullptr;
  std::vector<Patch> h_patches;
  std::vector<double> h_bcv;
  void upload(Context* ctx);
  void release();
};

struct MotionPlan {  // Copier between two layouts
  Motion* d_items = nullptr;
  MotionWork* d_work = nullptr;
  int nitems = 0, nwork = 0;
  Context* owner = nullptr;
  std::vector<Motion> h_items;
  void upload(Context* ctx);
  void release();
};

// doCF: 0 exchange and BC rules, 1 exchange, BC and homogeneous coarse-fine, 2
// coarse-fine only, 3 exchange only, 4 BC only
// comp: the component phi is (selects its block of per-face-cell BC values)
template <class T>
void fill_ghosts(const DeviceLayout& L, const GhostPlan& P, T* phi, int doCF, int homog, double dx,
                 double dxCrse, int comp = 0);
template <class T>
void gsrb(const DeviceLayout& L, T* phi, const T* rhs, int color, double dx, double alpha, double beta);
// One GSRB colour over a region of every box (tiles over each box), neighbours
// outside the box read from snap: region 0 the box, 1 the box grown by -1, 2 the
// face layer (dir, side) shrunk in the directions below dir.  lapOnly forces
// GSRBLAPLACIAN; otherwise the levelGSRB selection by alpha and beta.
template <class T>
void gsrb_region(const DeviceLayout& L, const Tile* tiles, int ntiles, T* phi, const T* snap, const T* rhs, int color,
                 int region, int dir, int side, bool lapOnly, double dx, double alpha, double beta);
// Eight-colour point updates on the cells with index parities color[0..2]:
// GSMCAMRPOP (levelMultiColor) and GSRBLAZY (levelGSRBLazy, lphi = L phi)
template <class T>
void gsmc(const DeviceLayout& L, T* phi, const T* rhs, const int color[3], double dx, double alpha, double beta);
template <class T>
void gsrb_lazy(const DeviceLayout& L, T* phi, const T* lphi, const T* rhs, const int color[3], double dx,
               double alpha, double beta);

// Ghost rule of each face of a merged level's rectangle during relaxation
// (homogeneous): 0 Neumann, 1 Dirichlet, 2 quadratic coarse-fine with c1, c2.
struct FaceRules {
  int kind[3][2];
  double c1 = 0, c2 = 0;
};
// One GSRB iteration (red, refresh, black) of a merged level in one launch:
// dst <- relaxed src.  Bitwise equal to fill_ghosts + gsrb(0) + fill_ghosts + gsrb(1).
// shape: tile shape index (kFusedShapes in kernels.hip), kc: z-chunk length (0 = automatic);
// zero_src: src is known to be zero everywhere and is not read (same result as reading it)
template <class T>
void gsrb_fused(const DeviceLayout& L, const FaceRules& fr, T* dst, const T* src, const T* rhs, double dx,
                double alpha, double beta, int shape = 0, int kc = 0, bool zero_src = false);
// Autotuner of the one-launch relaxation: times every tile shape with z-chunk
// lengths {auto, 8, 16, 32} on the given fields, returns the fastest {shape, kc};
// times (kFusedShapeCount x 4 entries, ms per launch) receives all timings when not null.
constexpr int kFusedShapeCount = 26;
template <class T>
std::pair<int, int> tune_gsrb_fused(const DeviceLayout& L, const FaceRules& fr, T* dst, const T* src, const T* rhs,
                                    double dx, double alpha, double beta, double* times);
template <class T>
void apply(const DeviceLayout& L, T* lofphi, const T* phi, double dx, double alpha, double beta);
template <class T>
void residual(const DeviceLayout& L, T* res, const T* phi, const T* rhs, double dx, double alpha,
              double beta);
// residual, with its max-norm folded into ctx->d_scalars[slot] (accumulate 0: the slot
// is cleared first), as maxnorm_async on the result would
template <class T>
void residual_norm(const DeviceLayout& L, T* res, const T* phi, const T* rhs, double dx, double alpha, double beta,
                   int slot, int accumulate);
template <class T>
void restrict_res(const DeviceLayout& Lf, const DeviceLayout& Lc, T* resc, const T* phi, const T* rhs,
                  double dx, double alpha, double beta);
// PROLONG and AVERAGE between a layout and its coarsening by ratio (2 or 4)
template <class T>
void prolong(const DeviceLayout& Lf, const DeviceLayout& Lc, T* phi, const T* coarse, int ratio = 2);
template <class T>
void average(const DeviceLayout& Lf, const DeviceLayout& Lc, T* coarse, const T* fine, int ratio = 2);
// coarse = AVERAGE(rhs - L(phi)) over coarsen(fine, ratio) without storing the fine
// residual (bitwise residual followed by average); phi's ghosts must be filled
template <class T>
void residual_average(const DeviceLayout& Lf, const DeviceLayout& Lc, T* coarse, const T* phi, const T* rhs,
                      double dx, double alpha, double beta, int ratio = 2);

template <class T>
void incr(const DeviceLayout& L, T* lhs, const T* x, double scale);
// lhs = lhs + x * 1.0; x = lhs (valid cells)
template <class T>
void incr_assign(const DeviceLayout& L, T* lhs, T* x);
// lhs = 0 + x * 1 (setToZero then incr, lhs not read)
template <class T>
void incr_from_zero(const DeviceLayout& L, T* lhs, const T* x);
// m += d; d = m; phi += d; d = 0 (incr, assignLocal, incr, setToZero)
template <class T>
void incr_assign_update(const DeviceLayout& L, T* m, T* d, T* phi);
// phi += (0 + x * 1) * 1 (an incr into a zero field, then into phi)
template <class T>
void incr_phi_from_zero(const DeviceLayout& L, T* phi, const T* x);
template <class T>
void axby(const DeviceLayout& L, T* lhs, const T* x, const T* y, double a, double b);
template <class T>
void scale(const DeviceLayout& L, T* x, double s);
template <class T>
void setval(const DeviceLayout& L, T* x, double v);
template <class T>
void copy_valid(const DeviceLayout& L, T* dst, const T* src);
template <class T>
void copy_all(const DeviceLayout& L, T* dst, const T* src);
template <class T>
void precond_init(const DeviceLayout& L, T* phi, const T* rhs, double mult);
// mixed precision: dst (float) <- src (double) on valid cells, and back with incr
void convert_valid_d2f(const DeviceLayout& L, float* dst, const double* src);
void incr_f2d(const DeviceLayout& L, double* lhs, const float* x);
void convert_valid_f2d(const DeviceLayout& L, double* dst, const float* src);
// phi += widened x, as incr_f2d into a zero field followed by incr(phi, field, 1)
void update_f2d(const DeviceLayout& L, double* phi, const float* x);

template <class T>
void motion(const MotionPlan& P, const DeviceLayout& Lsrc, const DeviceLayout& Ldst, T* dst, const T* src,
            int zero);

// reductions; result left in ctx->d_scalars[slot], returned after a stream sync.
// nc components cs elements apart, taken in FArrayBox order (component outermost)
template <class T>
double maxnorm(const DeviceLayout& L, const T* x, int slot, int nc = 1, long long cs = 0);
template <class T>
double dot_ordered(const DeviceLayout& L, const T* a, const T* b, int slot, int nc = 1, long long cs = 0);
// the same with the per-box partial sums in a file-scope device array (no device
// allocation); levels of at most kNoAllocBoxes boxes
constexpr int kNoAllocBoxes = 8192;
template <class T>
double dot_ordered_noalloc(const DeviceLayout& L, const T* a, const T* b, int slot, int nc = 1, long long cs = 0);
template <class T>
double norm2_ordered(const DeviceLayout& L, const T* x, int slot, int nc = 1, long long cs = 0);
// norm(.., 1): the same per-box column sums of |x|, no root
template <class T>
double norm1_ordered(const DeviceLayout& L, const T* x, int slot, int nc = 1, long long cs = 0);
// asynchronous variant: max-norm into d_scalars[slot] without a sync (value read later)
template <class T>
void maxnorm_async(const DeviceLayout& L, const T* x, int slot, int accumulate);

// inhomogeneous coarse-fine interpolation (QuadCFInterp::coarseFineInterp)
template <class T>
void cf_inhomog(Context* ctx, const CFCoarse* dc, int ncoarse, const CFGen* dg, const CFFine* df, int nfine,
                T* deriv_scratch, T* phif, const T* phic, double dxf, double dxc, int nref = 2);

// AMRResidual's AMROperator + axby(res, res, rhs, -1, 1) in two launches: every valid
// cell from its stencil, then the refluxed cells again with their corrections;
// bitwise the operator, reflux and axby launches
template <class T>
void apply_residual(const DeviceLayout& L, T* res, const T* phi, const T* rhs, double dx, double alpha, double beta);
template <class T>
void reflux_residual(Context* ctx, const RefluxCell* cells, int ncells, const RefluxCoarse* rc, const RefluxFine* rf,
                     T* res, const T* phic, const T* phif, const T* rhs, double beta_dx_c, double beta_ref_dx_c,
                     double minus_inv_dx, double dx, double alpha, double beta);
// reflux (AMRPoissonOp::reflux after the fine level's CF ghosts are filled)
template <class T>
void reflux(Context* ctx, const RefluxCell* cells, int ncells, const RefluxCoarse* rc, const RefluxFine* rf,
            T* u, const T* phic, const T* phif, double beta_dx_c, double beta_ref_dx_c, double minus_inv_dx);

}  // namespace ops
}  // namespace amrmg
