// FP64 instantiation of the compiled-plan nonlinear-term kernels
// (nonlinear_ct.h); the transform precision of the FP64 form.
#include "nonlinear_ct.h"

namespace nsp {
NSP_CT_EXTERN(double, 1)
NSP_CT_INSTANTIATE(double, 0)
template bool ct_theta_inv<double>(const NlVariant&, const Geom&, const cpx<double>*, const double*, const double2*,
                                  const double2*, cpx<double>*, size_t, hipStream_t, const Batch*);
template bool ct_theta_fwd<double>(const NlVariant&, const Geom&, const cpx<double>*, const cpx<double>*, size_t,
                                  const double*, const double2*, double2*, hipStream_t, const Batch*);
template bool ct_zpass<double>(const NlVariant&, const Geom&, const cpx<double>*, const cpx<double>*, const double*,
                              const cpx<double>*, size_t, cpx<double>*, hipStream_t, const Batch*);

bool ct_has_theta(int L) {
#define X(LL, ...) if (L == LL) return true;
  NSP_THETA_PLANS(X)
#undef X
  return false;
}
// the register form of the z pass (one first-pass butterfly per thread of 256)
bool ct_has_zreg(int L) {
#define X(LL, ...) if (L == LL) return LL / __VA_ARGS__::rad[0] <= 256;
  NSP_ZHALF_PLANS(X)
#undef X
  return false;
}
bool ct_has_zhalf(int L) {
#define X(LL, ...) if (L == LL) return true;
  NSP_ZHALF_PLANS(X)
#undef X
  return false;
}
// radix list of a compiled plan (host side, for twiddle tables)
std::vector<int> ct_radices(int L, bool theta) {
#define X(LL, ...)                                                                 \
  if (L == LL) {                                                                   \
    using PP = __VA_ARGS__;                                                        \
    return std::vector<int>(PP::rad, PP::rad + PP::P);                             \
  }
  if (theta) {
    NSP_THETA_PLANS(X)
  } else {
    NSP_ZHALF_PLANS(X)
  }
#undef X
  return {};
}
// radix list of the pair form's full-length plan for half length NH (empty if none)
std::vector<int> ct_pair_radices(int nh) {
#define X(LL, ...)                                                                 \
  if (nh == LL) {                                                                  \
    using PP = __VA_ARGS__;                                                        \
    return std::vector<int>(PP::rad, PP::rad + PP::P);                             \
  }
  NSP_ZPAIR_PLANS(X)
#undef X
  return {};
}
}  // namespace nsp
