// FP32 instantiation of the compiled-plan nonlinear-term kernels
// (nonlinear_ct.h); the transform precision of the mixed form.
#include "nonlinear_ct.h"

namespace nsp {
NSP_CT_EXTERN(float, 1)
NSP_CT_INSTANTIATE(float, 0)
template bool ct_theta_inv<float>(const NlVariant&, const Geom&, const cpx<float>*, const double*, const double2*,
                                  const double2*, cpx<float>*, size_t, hipStream_t, const Batch*);
template bool ct_theta_fwd<float>(const NlVariant&, const Geom&, const cpx<float>*, const cpx<float>*, size_t,
                                  const double*, const double2*, double2*, hipStream_t, const Batch*);
template bool ct_zpass<float>(const NlVariant&, const Geom&, const cpx<float>*, const cpx<float>*, const double*,
                              const cpx<float>*, size_t, cpx<float>*, hipStream_t, const Batch*);
}  // namespace nsp
