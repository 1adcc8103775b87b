// Handle types behind the opaque pointers of include/amrmg.h that more than one
// translation unit of the C ABI uses.
#pragma once
#include <memory>
#include <string>

#include "amrpoissonop.hpp"

struct amrmg_ctx {
  std::unique_ptr<amrmg::Context> c;
};

// A DisjointBoxLayout with its ProblemDomain resident in HBM.  refs counts the fields
// and operators defined on it; a level that an operator created (its multigrid child
// or its coarsened layout) belongs to that operator.
struct amrmg_level {
  amrmg_ctx* ctx = nullptr;
  std::unique_ptr<amrmg::DeviceLayout> lay;
  int refs = 0;
  bool ownedByOp = false;
};

// One component, one ghost cell.
struct amrmg_field {
  amrmg_level* lev = nullptr;
  amrmg::LevelData<double> ld;
};

struct amrmg_op {
  amrmg_ctx* ctx = nullptr;
  amrmg_level* lev = nullptr;
  std::unique_ptr<amrmg::AMRPoissonOp<double>> op;
  std::unique_ptr<amrmg_level> ownLevel;   // the level of an operator made by amrmg_op_create_coarser
  std::unique_ptr<amrmg_level> cfLevel;    // coarsen(level, 2) of an operator with a coarser AMR level
  amrmg_op* coarser = nullptr;             // AMR links (AMRnewOp)
  amrmg_op* finer = nullptr;
  amrmg_op* mgParent = nullptr;            // multigrid links (MGnewOp)
  amrmg_op* mgChild = nullptr;
};

struct amrmg_plan {
  amrmg_level* src = nullptr;
  amrmg_level* dst = nullptr;
  amrmg::ops::MotionPlan P;
};

// Error convention of the C ABI: 0 on success, -1 with the message kept per thread
// for amrmg_last_error; no exception crosses the boundary.
extern thread_local std::string amrmg_g_err;
#define API_BEGIN try {
#define API_END                       \
  }                                   \
  catch (const std::exception& e) {   \
    amrmg_g_err = e.what();           \
    return -1;                        \
  }                                   \
  catch (...) {                       \
    amrmg_g_err = "unknown error";    \
    return -1;                        \
  }                                   \
  return 0;
