// Definitions of the opaque C-ABI handle types shared by the host and device
// halves of the library.
#pragma once

#include <memory>
#include <string>

#include "../../../include/sc23ad.h"
#include "hierarchy.hpp"

struct sc23ad_csrmat_s {
  sc23ad::HostCsr a;
};

// The precision-independent part of a setup, kept so that hierarchies with
// other smoother parameters can be derived without coarsening again.
struct Sc23adSetupState {
  sc23ad::HostCsr fine;
  sc23ad::AmgChain chain;
  sc23ad::AmgParams prm;
};

struct sc23ad_hier_s {
  int precision = 2;
  int warnings = 0;
  int fine_nnz = 0;
  std::shared_ptr<const Sc23adSetupState> setup;
  sc23ad::IntHierarchy ih;
  sc23ad::FpHierarchy<double> dh;
  sc23ad::FpHierarchy<float> fh;
};

namespace sc23ad {
extern thread_local std::string g_last_error;
int fail(int code, const std::string& msg);
// sc23ad_hier_build_multi[_dev]: one coarsening chain, then a hierarchy per
// requested precision; device_setup runs the interpolation weights and the
// Galerkin products on the GPU.
int build_multi_impl(const sc23ad_csrmat* a, const sc23ad_amg_param* p, const int* precisions,
                     int count, sc23ad_hier** outs, bool device_setup);
}  // namespace sc23ad
