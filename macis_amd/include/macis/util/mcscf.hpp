// MCSCFSettings (restating include/macis/util/mcscf.hpp:15-30); only the CI
// fields are used by the accelerated path.
#pragma once
#include <cstddef>
#include <limits>

namespace macis {
struct MCSCFSettings {
  size_t max_macro_iter = 100;
  double max_orbital_step = 0.5;
  double orb_grad_tol_mcscf = 5e-6;
  bool enable_diis = true;
  size_t diis_start_iter = 3;
  size_t diis_nkeep = 10;
  double ci_res_tol = 1e-8;
  size_t ci_max_subspace = 20;
  double ci_matel_tol = std::numeric_limits<double>::epsilon();
};
}  // namespace macis
