// State conversion and output kernels outside the time step: remesh onto
// another grid (tools/Re_Mesh), physical-space fields and the cross-section
// integrals of the reference's diagnostics files (mod_io_nonlinear.cu).
#include "nsp_internal.h"

namespace nsp {

// Remesh gather (rmsh_k, tools/Re_Mesh/src/mod_remesh.cu:8-30, then
// modeZZ_k :35-41).  One thread per (new mode, component): a mode present
// on the old grid gets, at every new radial row, the remesh row's weights
// times the old column, summed over the row's nonzero band in increasing
// column order; other modes are zero.  ur, ut take the odd-in-r half for even
// m, uz the even half (parity index |(|m| mod 2) - flag|, flag 1 for ur, ut).
__global__ __launch_bounds__(256) void remesh_kernel(Geom gn, Geom go, const double* __restrict__ RM,
                                                     const int2* __restrict__ band,
                                                     const double2* __restrict__ uo, double2* __restrict__ un) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (k >= gn.modes) return;
  const int it = k / gn.nz, l = k - it * gn.nz;
  const int m = it < gn.nt / 2 ? it : it - gn.nt;
  const bool src = m >= -go.nt / 2 && m < go.nt / 2 && l < go.nz;
  const int ko = src ? (m >= 0 ? m : m + go.nt) * go.nz + l : 0;
  const int par = abs((abs(m) % 2) - (c < 2 ? 1 : 0));
  const double* R = RM + (size_t)par * gn.nr * go.nr;
  const double2* col = uo + (size_t)c * go.n + ko;
  double2* dst = un + (size_t)c * gn.n + k;
  for (int i = 0; i < gn.nr; i++) {
    double2 v = make_double2(0.0, 0.0);
    if (src) {
      const int2 bd = band[i];
      for (int j = bd.x; j <= bd.y; j++) {
        const double w = R[(size_t)i * go.nr + j];
        const double2 a = col[(size_t)j * go.modes];
        v.x += w * a.x;
        v.y += w * a.y;
      }
    }
    if (k == 0) {  // mode (0,0): no radial velocity, real ut and uz
      if (c == 0) v = make_double2(0.0, 0.0);
      else v.y = 0.0;
    }
    dst[(size_t)i * gn.modes] = v;
  }
}

void launch_remesh(const Geom& gn, const Geom& go, const double* RM, const int2* band, const double2* uo,
                   double2* un, hipStream_t s) {
  const dim3 grid((gn.modes + 255) / 256, 3);
  remesh_kernel<<<grid, 256, 0, s>>>(gn, go, RM, band, uo, un);
}

}  // namespace nsp
