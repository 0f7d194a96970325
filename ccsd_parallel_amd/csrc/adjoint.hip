// Adjoint of the CCSD energy (placeholder until the Lambda solver lands).
#include "adjoint.hpp"

namespace ccsd {

int adjoint_run(Ctx* c, const GradOut& g, ccsd_grad_info* info) {
  (void)c; (void)g; (void)info;
  return CCSD_ERR_STATE;
}

void adjoint_free(Ctx* c) { (void)c; }

}  // namespace ccsd
