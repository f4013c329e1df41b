// FP64 batched (ensemble) instantiation of the compiled-plan nonlinear-term
// kernels (nonlinear_ct.h): one member per blockIdx.y.
#include "nonlinear_ct.h"

namespace nsp {
NSP_CT_INSTANTIATE(double, 1)
}  // namespace nsp
