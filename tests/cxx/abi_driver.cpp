// A C++ caller of libamrmg.so through include/amrmg.h only (no Python): the host side
// of the boundary as a C++ code would bind it.  It solves one single-level problem
// with the whole-solve entry points, then evaluates the operator on the result through
// the operator-level handles, wrapped in a class shaped like an AMRLevelOp subclass
// whose virtuals each make one call (lib/src/AMRElliptic/AMRMultiGrid.H:38-226,
// LinearSolver.H:27-149).
//
// usage: abi_driver <input> <output>
//   input:  int32 n, int32 ispec[n], double rspec[9], int32 sparams[6],
//           double rparams[3], int64 ncells, double rhs[ncells]
//   output: int32 {iterations, exit status, bottom iterations}, double hist[iterations+1],
//           double phi[ncells], double residual[ncells], double {dot(res, res),
//           norm(res, 0), norm(res, 2)}
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "amrmg.h"

static void chk(int rc) {
  if (rc != 0) throw std::runtime_error(amrmg_last_error());
}

struct DeviceLD {  // the operator's data type T: one device-resident field
  amrmg_field* f = nullptr;
  ~DeviceLD() {
    if (f) amrmg_field_destroy(f);
  }
};

// One AMR level's operator; each method is the virtual of the same name.
class DevicePoissonOp {
 public:
  amrmg_op* m_op = nullptr;
  amrmg_level* m_lev = nullptr;
  void create(DeviceLD& lhs, const DeviceLD&) { chk(amrmg_field_create(m_lev, &lhs.f)); }
  void residual(DeviceLD& lhs, DeviceLD& phi, const DeviceLD& rhs, bool homogeneous) {
    chk(amrmg_op_residual(m_op, lhs.f, phi.f, rhs.f, homogeneous ? 1 : 0));
  }
  void applyOp(DeviceLD& lhs, DeviceLD& phi, bool homogeneous) {
    chk(amrmg_op_apply(m_op, lhs.f, phi.f, homogeneous ? 1 : 0));
  }
  void relax(DeviceLD& e, const DeviceLD& r, int n) { chk(amrmg_op_relax(m_op, e.f, r.f, n)); }
  double dotProduct(const DeviceLD& a, const DeviceLD& b) {
    double d = 0;
    chk(amrmg_field_dot(a.f, b.f, &d));
    return d;
  }
  double norm(const DeviceLD& x, int p) {
    double n = 0;
    chk(amrmg_field_norm(x.f, p, &n));
    return n;
  }
  ~DevicePoissonOp() {
    if (m_op) amrmg_op_destroy(m_op);
  }
};

template <class T>
static void get(FILE* f, T* p, size_t n) {
  if (std::fread(p, sizeof(T), n, f) != n) throw std::runtime_error("short input");
}
template <class T>
static void put(FILE* f, const T* p, size_t n) {
  if (std::fwrite(p, sizeof(T), n, f) != n) throw std::runtime_error("short write");
}

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <input> <output>\n", argv[0]);
    return 2;
  }
  try {
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) throw std::runtime_error("cannot open input");
    int32_t n = 0;
    get(in, &n, 1);
    std::vector<int> ispec((size_t)n);
    get(in, ispec.data(), ispec.size());
    double rspec[9];
    get(in, rspec, 9);
    int sp[6];
    double rp[3];
    get(in, sp, 6);
    get(in, rp, 3);
    int64_t ncells = 0;
    get(in, &ncells, 1);
    std::vector<double> rhs((size_t)ncells);
    get(in, rhs.data(), rhs.size());
    std::fclose(in);
    if (ispec[0] != 1) throw std::runtime_error("the driver takes one level");

    amrmg_ctx* ctx = nullptr;
    chk(amrmg_ctx_create(0, &ctx));
    // whole solve: AMRMultiGrid::solve
    amrmg_solver* s = nullptr;
    chk(amrmg_solver_create(ctx, ispec.data(), rspec, 0, &s));
    chk(amrmg_solver_set_params(s, sp, rp));
    chk(amrmg_solver_set_field(s, 1, rhs.data()));
    std::vector<double> hist(256);
    int info[3] = {0, 0, 0};
    chk(amrmg_solver_solve(s, 1, hist.data(), (int)hist.size(), info));
    std::vector<double> phi((size_t)ncells);
    chk(amrmg_solver_get_field(s, 0, phi.data()));
    chk(amrmg_solver_destroy(s));

    // the operator on the solution through the handles: residual rhs - L(phi)
    const int nbox = ispec[13];
    std::vector<int> boxes(ispec.begin() + 14, ispec.begin() + 14 + 6 * nbox);
    const int domain[6] = {ispec[1], ispec[2], ispec[3], ispec[4], ispec[5], ispec[6]};
    const int bc_type[6] = {ispec[7], ispec[8], ispec[9], ispec[10], ispec[11], ispec[12]};
    const double bc_val[6] = {rspec[3], rspec[4], rspec[5], rspec[6], rspec[7], rspec[8]};
    amrmg_level* lev = nullptr;
    chk(amrmg_level_register(ctx, boxes.data(), nbox, domain, 1, &lev));
    std::vector<double> res((size_t)ncells);
    double dot = 0, n0 = 0, n2 = 0;
    {
      DevicePoissonOp op;
      op.m_lev = lev;
      chk(amrmg_op_create(lev, nullptr, rspec[0], -1.0, rspec[1], rspec[2], bc_type, bc_val, &op.m_op));
      DeviceLD p, r, l;
      op.create(p, p);
      op.create(r, p);
      op.create(l, p);
      chk(amrmg_field_upload(p.f, phi.data(), 0));
      chk(amrmg_field_upload(r.f, rhs.data(), 0));
      op.residual(l, p, r, false);
      chk(amrmg_field_download(l.f, res.data(), 0));
      dot = op.dotProduct(l, l);
      n0 = op.norm(l, 0);
      n2 = op.norm(l, 2);
    }
    chk(amrmg_level_destroy(lev));
    chk(amrmg_ctx_destroy(ctx));

    FILE* out = std::fopen(argv[2], "wb");
    if (!out) throw std::runtime_error("cannot open output");
    put(out, info, 3);
    put(out, hist.data(), (size_t)info[0] + 1);
    put(out, phi.data(), phi.size());
    put(out, res.data(), res.size());
    const double sc[3] = {dot, n0, n2};
    put(out, sc, 3);
    std::fclose(out);
    std::printf("abi_driver: %d iterations, exit status %d\n", info[0], info[1]);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "abi_driver: %s\n", e.what());
    return 1;
  }
  return 0;
}
