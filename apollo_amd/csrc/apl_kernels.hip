// apl_kernels.hip -- instantiates the RHS kernels of one polynomial order.
// Compiled once per order with -DAPL_ORDER=<N> (see Makefile), so the eight
// orders build in parallel.
#include "apl_adjoint.h"
#include "apl_device.h"

#ifndef APL_ORDER
#error "compile with -DAPL_ORDER=<N>"
#endif

namespace apl {

template <int N, int EQN, int FLUX, typename T>
static int launch_typed(int linear, hipStream_t s, const RhsArgs &a, int nblocks) {
  if constexpr (EQN == APL_EQN_PHMAXWELL8 || EQN == APL_EQN_MAXWELL_TM3) {
    if (linear) {
      const size_t lds = sizeof(double) * LinLds<N, EQN, T>::doubles;
      auto kern = rhs_linear_kernel<N, EQN, FLUX, T>;
      if (lds > 65536 && hipFuncSetAttribute((const void *)kern,
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds) != hipSuccess)
        return APL_ERR_UNSUPPORTED;
      hipLaunchKernelGGL(kern, dim3(nblocks), dim3(64), lds, s, a);
      return hipGetLastError() == hipSuccess ? APL_OK : APL_ERR_HIP;
    }
  }
  if (linear) return APL_ERR_ARG;
  const size_t lds = sizeof(double) * Lds<N, EQN, T>::doubles;
  auto kern = rhs_kernel<N, EQN, FLUX, T>;
  if (lds > 65536) {
    if (hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess)
      return APL_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL(kern, dim3(nblocks), dim3(64), lds, s, a);
  return hipGetLastError() == hipSuccess ? APL_OK : APL_ERR_HIP;
}

template <int N, int EQN, int FLUX>
static int launch_prec(int prec, int linear, hipStream_t s, const RhsArgs &a, int nblocks) {
  switch (prec) {
    case APL_PREC_FP64: return launch_typed<N, EQN, FLUX, double>(linear, s, a, nblocks);
    case APL_PREC_FP32_MFMA: return launch_typed<N, EQN, FLUX, float>(linear, s, a, nblocks);
    case APL_PREC_BF16X3: return launch_typed<N, EQN, FLUX, Bf16x3>(linear, s, a, nblocks);
    default: return APL_ERR_UNSUPPORTED;
  }
}

template <int N>
int launch_rhs(int eqn, int flux, int prec, int linear, hipStream_t s, const RhsArgs &a,
               int nblocks) {
  switch (eqn) {
    case APL_EQN_EULER5:
      return launch_prec<N, APL_EQN_EULER5, APL_FLUX_LF>(prec, linear, s, a, nblocks);
    case APL_EQN_EULER4:
      return launch_prec<N, APL_EQN_EULER4, APL_FLUX_LF>(prec, linear, s, a, nblocks);
    case APL_EQN_PHMAXWELL8:
      return launch_prec<N, APL_EQN_PHMAXWELL8, APL_FLUX_LF>(prec, linear, s, a, nblocks);
    case APL_EQN_MAXWELL_TM3:
      if (flux == APL_FLUX_UPWIND)
        return launch_prec<N, APL_EQN_MAXWELL_TM3, APL_FLUX_UPWIND>(prec, linear, s, a, nblocks);
      return launch_prec<N, APL_EQN_MAXWELL_TM3, APL_FLUX_LF>(prec, linear, s, a, nblocks);
    default: return APL_ERR_ARG;
  }
}

template int launch_rhs<APL_ORDER>(int, int, int, int, hipStream_t, const RhsArgs &, int);

template <typename K>
static bool allow_lds(K kern, size_t lds) {
  return lds <= 65536 || hipFuncSetAttribute((const void *)kern,
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds) == hipSuccess;
}

template <int N, int EQN, int FLUX>
static int launch_adjoint_typed(hipStream_t s, const AdjArgs &a, int nblocks) {
  auto ka = adj_face_kernel<N, EQN>;
  auto kb = adj_rhs_kernel<N, EQN, FLUX>;
  const size_t la = sizeof(double) * AdjFaceLds<N, EQN>::doubles;
  const size_t lb = sizeof(double) * AdjLds<N, EQN>::doubles;
  if (!allow_lds(ka, la) || !allow_lds(kb, lb)) return APL_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(ka, dim3(nblocks), dim3(64), la, s, a);
  if (hipGetLastError() != hipSuccess) return APL_ERR_HIP;
  hipLaunchKernelGGL(kb, dim3(nblocks), dim3(64), lb, s, a);
  return hipGetLastError() == hipSuccess ? APL_OK : APL_ERR_HIP;
}

template <int N>
int launch_adjoint(int eqn, int flux, hipStream_t s, const AdjArgs &a, int nblocks) {
  switch (eqn) {
    case APL_EQN_EULER5: return launch_adjoint_typed<N, APL_EQN_EULER5, APL_FLUX_LF>(s, a, nblocks);
    case APL_EQN_EULER4: return launch_adjoint_typed<N, APL_EQN_EULER4, APL_FLUX_LF>(s, a, nblocks);
    case APL_EQN_PHMAXWELL8:
      return launch_adjoint_typed<N, APL_EQN_PHMAXWELL8, APL_FLUX_LF>(s, a, nblocks);
    case APL_EQN_MAXWELL_TM3:
      if (flux == APL_FLUX_UPWIND)
        return launch_adjoint_typed<N, APL_EQN_MAXWELL_TM3, APL_FLUX_UPWIND>(s, a, nblocks);
      return launch_adjoint_typed<N, APL_EQN_MAXWELL_TM3, APL_FLUX_LF>(s, a, nblocks);
    default: return APL_ERR_ARG;
  }
}

template <int N>
int launch_objective(int m, hipStream_t s, const ObjArgs &a, int nblocks) {
  switch (m) {
    case 3: hipLaunchKernelGGL((objective_kernel<N, 3>), dim3(nblocks), dim3(64), 0, s, a); break;
    case 4: hipLaunchKernelGGL((objective_kernel<N, 4>), dim3(nblocks), dim3(64), 0, s, a); break;
    case 5: hipLaunchKernelGGL((objective_kernel<N, 5>), dim3(nblocks), dim3(64), 0, s, a); break;
    case 8: hipLaunchKernelGGL((objective_kernel<N, 8>), dim3(nblocks), dim3(64), 0, s, a); break;
    default: return APL_ERR_ARG;
  }
  return hipGetLastError() == hipSuccess ? APL_OK : APL_ERR_HIP;
}

template int launch_adjoint<APL_ORDER>(int, int, hipStream_t, const AdjArgs &, int);
template int launch_objective<APL_ORDER>(int, hipStream_t, const ObjArgs &, int);

#if APL_ORDER == 1
__global__ void advance_time_kernel(double *t, double dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *t = *t + dt;
}

int launch_advance_time(hipStream_t s, double *t, double dt) {
  hipLaunchKernelGGL(advance_time_kernel, dim3(1), dim3(64), 0, s, t, dt);
  return hipGetLastError() == hipSuccess ? APL_OK : APL_ERR_HIP;
}
#endif

}  // namespace apl
