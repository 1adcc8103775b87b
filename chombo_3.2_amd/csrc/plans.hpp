// Host construction of the device tables that replace the reference's per-box
// metadata: Copier motion plans, BC/exchange/CF ghost rules, QuadCFStencil
// derivative stencils and LevelFluxRegister reflux lists.  Built once per
// hierarchy and reused by every solve.
#pragma once
#include <vector>

#include "device.hpp"
#include "ops.hpp"

namespace amrmg {

// BCValueFunc (BCFunc.H:205-208) with a caller pointer: the value at face position
// pos of the face (dir, side), side 0 low and 1 high
typedef void (*BCValueFn)(double* pos, int* dir, int* side, double* value, void* user);

struct BCSpec {
  int type[2][3];     // [side][dir]: 0 Neumann, 1 Dirichlet (linear), 2 Dirichlet (quadratic, DiriBC order 2),
                      // 3 periodic (both sides of the direction)
  double val[2][3];
  // When set, the inhomogeneous rules take their value per face cell from fn (at the
  // face centre, as DiriBC / NeumBC call their BCValueHolder) instead of val.
  BCValueFn fn = nullptr;
  void* user = nullptr;
};

// point -> box lookup over a layout (bucketed)
class BoxFinder {
 public:
  explicit BoxFinder(const DBL& d);
  int find(const IV& p) const;
  // boxes intersecting region r (each once, ascending index)
  std::vector<int> overlapping(const Box& r) const;

 private:
  const DBL* dbl_;
  Box grid_;
  int B_ = 1;
  IV n_{1, 1, 1};
  std::vector<int> start_, list_;
  int bucket(int i, int j, int k) const;
};

// Face-ghost rules of one level: exchange from neighbours, BCs on the domain
// boundary, homogeneous CF interpolation on coarse-fine faces (CFRegion/CFIVS).
ops::GhostPlan build_ghost_plan(const DBL& dbl, const Box& domain, const BCSpec& bc, bool withCF, bool merged);

// The face values of the plan's BC rules from bc.fn at level spacing dx (voff and
// h_bcv; bc.fn null restores the face constants).  bc.fn fills one value per
// component; component c's values are the block at c * nbcv.  The plan must be
// uploaded again.
void attach_bc_values(ops::GhostPlan& P, double dx, const BCSpec& bc, int ncomp = 1);

// Copier from src layout to dst layout (Copier::define with zero ghost).
ops::MotionPlan build_copier(const DBL& src, const DBL& dst);

// QuadCFInterp tables of a fine level over its coarser level.
struct CFPlanHost {
  std::vector<CFCoarse> coarse;
  std::vector<CFGen> gens;
  std::vector<CFFine> fine;
};
CFPlanHost build_cf_plan(const DBL& fine, const DeviceLayout& fineLay, const DBL& coarse,
                         const DeviceLayout& coarseLay, const Box& fineDomain, double dxFine, int ref);

// The face of a reflux contribution (variable coefficients weight each flux by
// bCoef there): box, direction and the face index (the low face of cell iv) of
// each of its fluxes (one for a coarse record, four for a fine one).
struct RefluxFace {
  int box, dir;
  int iv[4][3];
};
// Reflux tables of a coarse level with its finer level.
struct RefluxPlanHost {
  std::vector<RefluxCell> cells;
  std::vector<RefluxCoarse> coarse;
  std::vector<RefluxFine> fine;
  std::vector<RefluxFace> coarseFace, fineFace;  // beside coarse / fine, record for record
};
RefluxPlanHost build_reflux_plan(const DBL& coarse, const DeviceLayout& coarseLay, const DBL& fine,
                                 const DeviceLayout& fineLay, const Box& coarseDomain, int ref);

}  // namespace amrmg
