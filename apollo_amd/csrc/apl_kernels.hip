// apl_kernels.hip -- instantiates the RHS kernels of one polynomial order.
// Compiled once per order with -DAPL_ORDER=<N> (see Makefile), so the eight
// orders build in parallel.
#include "apl_adjoint.h"
#include "apl_device.h"
#include "apl_limiter.h"
#include "apl_valu.h"

#ifndef APL_ORDER
#error "compile with -DAPL_ORDER=<N>"
#endif

namespace apl {

template <typename K>
static bool allow_lds(K kern, size_t lds) {
  return lds <= 65536 || hipFuncSetAttribute((const void *)kern,
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds) == hipSuccess;
}

// Launch `kern` over `nbatches` 16-element batches in workgroups of W waves
// (one batch per wave, LDS_DOUBLES doubles of LDS per wave).
template <int LDS_DOUBLES, int W, typename K, typename A>
static int launch_batches(K kern, hipStream_t s, const A &a, int nbatches) {
  const size_t lds = sizeof(double) * (size_t)LDS_DOUBLES * W;
  if (!allow_lds(kern, lds)) return APL_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(kern, dim3((nbatches + W - 1) / W), dim3(64 * W), lds, s, a);
  return hipGetLastError() == hipSuccess ? APL_OK : APL_ERR_HIP;
}

// kind: 0 the cubature form on MFMA, 1 the nodal form (linear fluxes) on
// MFMA, 2 the cubature form on the VALU kernel (N <= 3, FP64 and FP32),
// 3 the nodal form on the VALU kernel (TM, FP64 and FP32)
template <int N, int EQN, int FLUX, typename T>
static int launch_typed(int kind, hipStream_t s, const RhsArgs &a, int nblocks) {
  if constexpr (EQN == APL_EQN_PHMAXWELL8 || EQN == APL_EQN_MAXWELL_TM3) {
    if (kind == 1)
      return launch_batches<LinLds<N, EQN, T>::doubles, LinGroup<N, EQN, T>::W>(
          rhs_linear_kernel<N, EQN, FLUX, T>, s, a, nblocks);
  }
  if constexpr (EQN == APL_EQN_MAXWELL_TM3 && sizeof(T) != 1) {
    if (kind == 3) {
      // 64 elements per workgroup
      auto kern = rhs_valu_linear_kernel<N, EQN, FLUX, T>;
      const size_t lds = sizeof(double) * (size_t)ValuLinLds<N, EQN, T>::doubles;
      if (!allow_lds(kern, lds)) return APL_ERR_UNSUPPORTED;
      hipLaunchKernelGGL(kern, dim3((a.K + 63) / 64), dim3(64 * ValuLinCfg<N, EQN>::W), lds, s, a);
      return hipGetLastError() == hipSuccess ? APL_OK : APL_ERR_HIP;
    }
  }
  if (kind == 1 || kind == 3) return APL_ERR_ARG;
  if (kind == 2) {
    if constexpr (N <= 3 && sizeof(T) != 1) {
      // 64 elements per workgroup of four waves, or of eight on small meshes
      const int groups = (a.K + 63) / 64;
      if (groups < VALU_SMALL_GROUPS && ValuLds<N, EQN, T, 8>::doubles * 8 <= 163840) {
        auto kern = rhs_valu_kernel<N, EQN, FLUX, T, 8>;
        const size_t lds = sizeof(double) * (size_t)ValuLds<N, EQN, T, 8>::doubles;
        if (!allow_lds(kern, lds)) return APL_ERR_UNSUPPORTED;
        hipLaunchKernelGGL(kern, dim3(groups), dim3(512), lds, s, a);
      } else {
        auto kern = rhs_valu_kernel<N, EQN, FLUX, T, 4>;
        const size_t lds = sizeof(double) * (size_t)ValuLds<N, EQN, T, 4>::doubles;
        if (!allow_lds(kern, lds)) return APL_ERR_UNSUPPORTED;
        hipLaunchKernelGGL(kern, dim3(groups), dim3(256), lds, s, a);
      }
      return hipGetLastError() == hipSuccess ? APL_OK : APL_ERR_HIP;
    } else {
      return APL_ERR_UNSUPPORTED;
    }
  }
  return launch_batches<Lds<N, EQN, T>::doubles, RhsGroup<N, EQN, T>::W>(rhs_kernel<N, EQN, FLUX, T>, s,
                                                                          a, nblocks);
}

template <int N, int EQN, int FLUX>
static int launch_prec(int prec, int kind, hipStream_t s, const RhsArgs &a, int nblocks) {
  switch (prec) {
    case APL_PREC_FP64: return launch_typed<N, EQN, FLUX, double>(kind, s, a, nblocks);
    case APL_PREC_FP32_MFMA: return launch_typed<N, EQN, FLUX, float>(kind, s, a, nblocks);
    case APL_PREC_BF16X3: return launch_typed<N, EQN, FLUX, Bf16x3>(kind, s, a, nblocks);
    default: return APL_ERR_UNSUPPORTED;
  }
}

template <int N>
int launch_rhs(int eqn, int flux, int prec, int linear, hipStream_t s, const RhsArgs &a,
               int nblocks) {
  switch (eqn) {
    case APL_EQN_EULER5:
      if (flux == APL_FLUX_HLL) return launch_prec<N, APL_EQN_EULER5, APL_FLUX_HLL>(prec, linear, s, a, nblocks);
      if (flux == APL_FLUX_ROE) return launch_prec<N, APL_EQN_EULER5, APL_FLUX_ROE>(prec, linear, s, a, nblocks);
      if (flux == APL_FLUX_WAVE) return launch_prec<N, APL_EQN_EULER5, APL_FLUX_WAVE>(prec, linear, s, a, nblocks);
      return launch_prec<N, APL_EQN_EULER5, APL_FLUX_LF>(prec, linear, s, a, nblocks);
    case APL_EQN_EULER4:
      if (flux == APL_FLUX_HLL) return launch_prec<N, APL_EQN_EULER4, APL_FLUX_HLL>(prec, linear, s, a, nblocks);
      if (flux == APL_FLUX_ROE) return launch_prec<N, APL_EQN_EULER4, APL_FLUX_ROE>(prec, linear, s, a, nblocks);
      if (flux == APL_FLUX_WAVE) return launch_prec<N, APL_EQN_EULER4, APL_FLUX_WAVE>(prec, linear, s, a, nblocks);
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

template <int N, int EQN, int FLUX>
static int launch_adjoint_typed(hipStream_t s, const AdjArgs &a, int nblocks) {
  if (a.face_pass) {
    const int rc = launch_batches<AdjFaceLds<N, EQN>::doubles, AdjFaceGroup<N, EQN>::W>(
        adj_face_kernel<N, EQN>, s, a, nblocks);
    if (rc) return rc;
  }
  if (a.mode == 1)
    return launch_batches<AdjLds<N, EQN>::doubles, AdjGroup<N, EQN>::W>(
        adj_rhs_kernel<N, EQN, FLUX, true>, s, a, nblocks);
  return launch_batches<AdjLds<N, EQN>::doubles, AdjGroup<N, EQN>::W>(
      adj_rhs_kernel<N, EQN, FLUX, false>, s, a, nblocks);
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

template <int N, int EQN>
static int launch_limiter_typed(hipStream_t s, const LimArgs &a) {
  const dim3 grid((a.K + 255) / 256), block(256), wgrid((a.K + LE - 1) / LE),
      rgrid((a.K + LER - 1) / LER), wave(64);
  hipLaunchKernelGGL((lim_average_kernel<N, EQN>), wgrid, wave, 0, s, a);
  hipLaunchKernelGGL((lim_gradient_kernel<N, EQN>), grid, block, 0, s, a);
  hipLaunchKernelGGL((lim_reconstruct_kernel<N, EQN>), rgrid, wave, 0, s, a);
  return hipGetLastError() == hipSuccess ? APL_OK : APL_ERR_HIP;
}

template <int N>
int launch_limiter(int eqn, hipStream_t s, const LimArgs &a) {
  if constexpr (N > 4) {
    return APL_ERR_UNSUPPORTED;
  } else {
    switch (eqn) {
      case APL_EQN_EULER5: return launch_limiter_typed<N, APL_EQN_EULER5>(s, a);
      case APL_EQN_EULER4: return launch_limiter_typed<N, APL_EQN_EULER4>(s, a);
      default: return APL_ERR_UNSUPPORTED;
    }
  }
}

template int launch_limiter<APL_ORDER>(int, hipStream_t, const LimArgs &);

#if APL_ORDER == 1
__global__ void __launch_bounds__(256) lim_patch_kernel(LimArgs a) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= a.K) return;
  const double *ave = a.tab, *vk = a.vert + e * 6;
  // J from the vertices as the geometry computes it (wxnodaldggeometry2d.cc:412-426)
  const double xr = (vk[2] - vk[0]) / 2, xs = (vk[4] - vk[0]) / 2;
  const double yr = (vk[3] - vk[1]) / 2, ys = (vk[5] - vk[1]) / 2;
  const double J = -xs * yr + xr * ys;
  double xc = 0.0, yc = 0.0, A0 = 0.0;
#pragma unroll
  for (int kk = 0; kk < 3; ++kk) {
    xc += ave[kk] * vk[2 * kk];
    yc += ave[kk] * vk[2 * kk + 1];
    A0 += 2. / 3. * ave[kk] * J;
  }
  double *o = a.avg + e * 8;
  o[5] = xc;
  o[6] = yc;
  o[7] = A0;
}

int launch_limiter_patch(hipStream_t s, const LimArgs &a) {
  hipLaunchKernelGGL(lim_patch_kernel, dim3((a.K + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? APL_OK : APL_ERR_HIP;
}

__global__ void advance_time_kernel(double *t, double dt) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *t = *t + dt;
}

int launch_advance_time(hipStream_t s, double *t, double dt) {
  hipLaunchKernelGGL(advance_time_kernel, dim3(1), dim3(64), 0, s, t, dt);
  return hipGetLastError() == hipSuccess ? APL_OK : APL_ERR_HIP;
}

// One 256-lane workgroup: strided maxima, then a tree over LDS.
__global__ void __launch_bounds__(256) max_reduce_kernel(const double *in, int n, double *out) {
  __shared__ double part[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) m = in[i] > m ? in[i] : m;
  part[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double o = part[threadIdx.x + w];
      part[threadIdx.x] = o > part[threadIdx.x] ? o : part[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

int launch_max_reduce(hipStream_t s, const double *in, int n, double *out) {
  hipLaunchKernelGGL(max_reduce_kernel, dim3(1), dim3(256), 0, s, in, n, out);
  return hipGetLastError() == hipSuccess ? APL_OK : APL_ERR_HIP;
}
#endif

}  // namespace apl
