// Internal layout of the host operator handle shared by the host and device
// halves of the C ABI.
#pragma once

#include <memory>
#include <string>

#include "../host/gemslr_host.hpp"

struct pgs_gemslr_s {
  int is_complex = 0;
  std::unique_ptr<pgs::HostOp<double>> op_d;
  std::unique_ptr<pgs::HostOp<pgs::zdouble>> op_z;
  std::unique_ptr<pgs::HostApplyWork<double>> work_d;
  std::unique_ptr<pgs::HostApplyWork<pgs::zdouble>> work_z;
  std::string msg;
};
