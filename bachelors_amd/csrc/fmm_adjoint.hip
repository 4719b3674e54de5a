// Transposed FMM (adjoint of the weight -> B map); see run_adjoint.
#include "fmm_adjoint.hpp"

namespace fmm {

void run_adjoint(AdjointState&, const AdjointArgs&, const double*, double*, hipStream_t) {
    throw Error(5, "adjoint not available yet");
}

}  // namespace fmm
