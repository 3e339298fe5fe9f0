// Hellmann-Feynman gradients of exciton energies.
#include "context.hpp"

namespace xatu_dev {

void hf_gradients(Context&, int, const double2*, int, const int*, double*) {
    throw Error(XATU_ERR_INVALID_ARGUMENT, "gradients not available yet");
}

}  // namespace xatu_dev
