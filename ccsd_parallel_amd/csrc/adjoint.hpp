// Adjoint (backward) of the converged CCSD energy: Lambda equations and the
// contraction of the unrelaxed densities into dE/df and dE/d<pq||rs>.
#pragma once
#include "engine.hpp"

namespace ccsd {

struct GradOut {
  double *fock, *oooo, *ooov, *oovv, *ovov, *ovvv, *vvvv;
};

int adjoint_run(Ctx* c, const GradOut& g, ccsd_grad_info* info);
void adjoint_free(Ctx* c);

}  // namespace ccsd
