// Variable-coefficient operator of VCAMRPoissonOp2 (lib/src/AMRElliptic/VCAMRPoissonOp2.{H,cpp},
// VCAMRPoissonOpF.ChF): L phi = alpha a phi - beta div(b grad phi) on one level.
//
// Storage: aCoef and the relaxation coefficient lambda (resetLambda) are level
// fields with the level's box table; bCoef (a FluxBox in the reference) is one
// array per direction d holding, per box, the faces normal to d over the box grown
// by one face in d, Fortran order.  On a merged level one face array covers the
// rectangle, so a face shared by two boxes is stored once (the host checks that
// the two boxes' values for it agree).  A face table per direction, entry for
// entry beside the level's box table (boxes, then the rectangle when merged),
// gives the index of face (lo0, lo1, lo2) and the face array's strides.
#pragma once
#include <memory>
#include <vector>

#include "device.hpp"
#include "ops.hpp"

namespace amrmg {

// faces of one reflux contribution in a level's face arrays (x[0] for a coarse
// record, x[0..3] for a fine one)
struct VCFaceIdx {
  long long x[4];
  int dir, pad;
};

template <class Real>
struct VCCoefs {
  Context* ctx = nullptr;
  const DeviceLayout* lay = nullptr;
  LevelData<Real> a, lambda;
  Real* b[3] = {nullptr, nullptr, nullptr};
  long long fslab[3] = {0, 0, 0};
  std::vector<DBox> h_fbox[3];
  DBox* d_fbox[3] = {nullptr, nullptr, nullptr};

  explicit VCCoefs(const DeviceLayout* L);
  ~VCCoefs();
  VCCoefs(const VCCoefs&) = delete;
  VCCoefs& operator=(const VCCoefs&) = delete;
  // entries of the packed face layout of direction d (per box, its face box)
  long long packedFaces(int d) const;
  // host data in the packed layouts: a valid-packed, faces per box in Fortran order
  void setCells(const double* valid_packed);
  void setFaces(int d, const double* packed);
  long long bytes() const;
};

namespace vc {
// GSRBHELMHOLTZVC (VCAMRPoissonOpF.ChF:57-155), one colour
template <class T>
void gsrb(const DeviceLayout& L, const VCCoefs<T>& C, T* phi, const T* rhs, int color, double dx, double alpha,
          double beta);
// One GSRB iteration (red, refresh, black) of a merged level in one launch:
// dst <- relaxed src, bitwise equal to the ghost refresh + gsrb(0) + ghost refresh
// + gsrb(1) sequence.  shape: tile shape (-1 automatic, 0..3).
template <class T>
void gsrb_fused(const DeviceLayout& L, const ops::FaceRules& fr, const VCCoefs<T>& C, T* dst, const T* src,
                const T* rhs, double dx, double alpha, double beta, int shape = -1);
// VCCOMPUTEOP (VCAMRPoissonOpF.ChF:197-265)
template <class T>
void apply(const DeviceLayout& L, const VCCoefs<T>& C, T* lofphi, const T* phi, double dx, double alpha,
           double beta);
// VCCOMPUTERES (VCAMRPoissonOpF.ChF:311-382)
template <class T>
void residual(const DeviceLayout& L, const VCCoefs<T>& C, T* res, const T* phi, const T* rhs, double dx,
              double alpha, double beta);
// RESTRICTRESVC (VCAMRPoissonOpF.ChF:422-496)
template <class T>
void restrict_res(const DeviceLayout& Lf, const DeviceLayout& Lc, const VCCoefs<T>& C, T* resc, const T* phi,
                  const T* rhs, double dx, double alpha, double beta);
// resetLambda (VCAMRPoissonOp2.cpp:283-318) with SUMFACES (VCAMRPoissonOpF.ChF:498-525)
template <class T>
void lambda(VCCoefs<T>& C, double dx, double alpha, double beta);
// preCond's first half (VCAMRPoissonOp2.cpp:93-127): phi = rhs * lambda on valid cells
template <class T>
void precond_init(const DeviceLayout& L, const VCCoefs<T>& C, T* phi, const T* rhs);
// MGnewOp's coefficient coarsening (VCAMRPoissonOp2.cpp:1024-1060): C = the level-0
// coefficients F averaged with ratio r (CoarseAverage / CoarseAverageFace, arithmetic)
template <class T>
void average_coefs(const VCCoefs<T>& F, VCCoefs<T>& C, int r, int harmonic = 0);
// VCAMRPoissonOp2::reflux (VCAMRPoissonOp2.cpp:339-430): the refluxed residual with
// fluxes -b * gradient (getFlux, :787-837), from the tables of ops::reflux and the
// faces of their records
template <class T>
void reflux(Context* ctx, const RefluxCell* cells, int ncells, const RefluxCoarse* rc, const RefluxFine* rf,
            const VCFaceIdx* fc, const VCFaceIdx* ff, T* u, const T* phic, const T* phif, const VCCoefs<T>& C,
            const VCCoefs<T>& F, double beta_dx_c, double beta_ref_dx_c, double minus_inv_dx);
// diagonalScale / divideByIdentityCoef (VCAMRPoissonOp2.H:120-140): x *= aCoef,
// resp. x /= aCoef on valid cells
template <class T>
void scale_by_a(const DeviceLayout& L, const VCCoefs<T>& C, T* x, bool divide);
// x *= lambda on valid cells (VCAMRPoissonOp2::levelJacobi)
template <class T>
void scale_by_lambda(const DeviceLayout& L, const VCCoefs<T>& C, T* x);
// FP32 copy of FP64 coefficients on the same level (a and bCoef; lambda is recomputed)
void narrow(const VCCoefs<double>& src, VCCoefs<float>& dst);
}  // namespace vc

}  // namespace amrmg
