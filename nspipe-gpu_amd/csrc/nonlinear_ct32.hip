// FP32 instantiation of the compiled-plan nonlinear-term kernels
// (nonlinear_ct.h); the transform precision of the mixed form.
#include "nonlinear_ct.h"

namespace nsp {
template bool ct_theta_inv<float>(const NlVariant&, const Geom&, const cpx<float>*, const double*, const double2*,
                                  const double2*, cpx<float>*, size_t, hipStream_t);
template bool ct_theta_fwd<float>(const NlVariant&, const Geom&, const cpx<float>*, const cpx<float>*, size_t,
                                  const double*, const double2*, double2*, hipStream_t);
template bool ct_zpass<float>(const NlVariant&, const Geom&, const cpx<float>*, const cpx<float>*, const double*,
                              const cpx<float>*, size_t, cpx<float>*, hipStream_t);
}  // namespace nsp
