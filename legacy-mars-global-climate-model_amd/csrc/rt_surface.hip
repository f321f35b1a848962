// rt_surface.hip -- DSOLFLUX (code/dsolflux.f) and COMP3's surface VIS flux for
// TEMPGR (code/comp3.f:336-348) over a grid, from the DETAU/DIFFVT outputs of the
// radiation: one thread per column, the glibc-exact exp, the reference's
// summation order, so the result equals the reference's bit for bit.
#include "rt_ctx.h"
#include "rt_libm.h"

namespace mgcm {

struct SurfArgs {
  int ncol;
  double sol[NV];
  const double *acosz, *detau, *diffvt;
  double *directsol, *dnvflux;
};

__global__ __launch_bounds__(64) void k_dsolflux(const RtConst *__restrict__ rc, SurfArgs a) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= a.ncol) return;
  const double ubar0 = a.acosz[col];
  double directsol = 0.0;
  if (ubar0 >= F32LIT(1.0e-5)) {  // code/comp3.f:336-345: sun down gives 0
    const double *detau = a.detau + (size_t)col * NG * NV;  // DETAU(nw, ng)
    for (int nw = 0; nw < NV; nw++) {
      const double factor = ubar0 * a.sol[nw];
      for (int ng = 0; ng < NG - 1; ng++) {
        const double d = detau[ng * NV + nw];
        if (d <= 5.0)
          directsol = directsol + factor * mgcm_libm::mgcm_exp(-d / ubar0) * rc->gw[nw][ng] * (1.0 - rc->fzero[nw]);
      }
      const double d = detau[(NG - 1) * NV + nw];
      if (d <= 5.0) directsol = directsol + factor * mgcm_libm::mgcm_exp(-d / ubar0) * rc->fzero[nw];
    }
  }
  a.directsol[col] = directsol;
  if (a.dnvflux) a.dnvflux[col] = a.diffvt[col] + directsol;  // DNVFLUX = DNDIFFV + DIRECTSOL
}

hipError_t launch_dsolflux(const RtConst *rc, int ncol, const double *sol, const double *acosz,
                           const double *detau, const double *diffvt, double *directsol, double *dnvflux,
                           hipStream_t st) {
  SurfArgs a;
  a.ncol = ncol;
  for (int b = 0; b < NV; b++) a.sol[b] = sol[b];
  a.acosz = acosz;
  a.detau = detau;
  a.diffvt = diffvt;
  a.directsol = directsol;
  a.dnvflux = dnvflux;
  hipLaunchKernelGGL(k_dsolflux, dim3((ncol + 63) / 64), dim3(64), 0, st, rc, a);
  return hipGetLastError();
}

}  // namespace mgcm

extern "C" int mgcm_rt_dsolflux(mgcm_rt_ctx *c, int ncol, const double *sol, const double *acosz,
                                const double *detau, const double *diffvt, double *directsol, double *dnvflux,
                                void *stream) {
  if (!c || !sol || ncol < 0) return MGCM_RT_ERR_ARG;
  if (ncol == 0) return MGCM_RT_OK;
  if (!acosz || !detau || !directsol || (dnvflux && !diffvt)) return MGCM_RT_ERR_ARG;
  return mgcm::launch_dsolflux(c->d_const, ncol, sol, acosz, detau, diffvt, directsol, dnvflux,
                               (hipStream_t)stream) == hipSuccess
             ? MGCM_RT_OK
             : MGCM_RT_ERR_DEVICE;
}
