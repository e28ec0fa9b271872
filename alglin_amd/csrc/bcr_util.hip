// Small elementwise helpers (type conversion / zero fill / axpy).  Fills are
// kernels rather than hipMemsetAsync so that a launch sequence stays valid
// under HIP graph capture.
#include "bcr_launch.hh"

namespace bcr {

template <typename TS, typename TD>
__global__ void __launch_bounds__(256) convert_kernel(const TS* __restrict__ src, TD* __restrict__ dst, int64_t count) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x)
    dst[i] = src ? TD(src[i]) : TD(0);
}

__global__ void __launch_bounds__(256) axpy_kernel(double alpha, const double* __restrict__ src,
                                                   double* __restrict__ dst, int64_t count) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x)
    dst[i] += alpha * src[i];
}

static int grid_for(int64_t count) {
  const int64_t b = (count + 255) / 256;
  return int(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

template <typename TS, typename TD>
hipError_t launch_convert(const TS* src, TD* dst, int64_t count, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL((convert_kernel<TS, TD>), dim3(grid_for(count)), dim3(256), 0, st, src, dst, count);
  return hipGetLastError();
}

hipError_t launch_axpy(double alpha, const double* src, double* dst, int64_t count, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(count)), dim3(256), 0, st, alpha, src, dst, count);
  return hipGetLastError();
}

template hipError_t launch_convert<double, double>(const double*, double*, int64_t, hipStream_t);
template hipError_t launch_convert<float, float>(const float*, float*, int64_t, hipStream_t);
template hipError_t launch_convert<double, float>(const double*, float*, int64_t, hipStream_t);
template hipError_t launch_convert<float, double>(const float*, double*, int64_t, hipStream_t);
template hipError_t launch_convert<int32_t, int32_t>(const int32_t*, int32_t*, int64_t, hipStream_t);

}  // namespace bcr
