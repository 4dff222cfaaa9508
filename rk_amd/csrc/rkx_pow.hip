// The step control's pow (rkx_pow.hpp) evaluated elementwise, on the device
// and on the host CPU: the entry points that pin libm_pow against the C
// library's pow over whole argument sets.
#include "rkx_device.hpp"

namespace rkx {
namespace {

constexpr int kPowBlock = 256;

__global__ __launch_bounds__(kPowBlock) void k_pow(const double *__restrict__ x,
                                                    const double *__restrict__ y,
                                                    double *__restrict__ p, size_t n)
{
  const size_t stride = static_cast<size_t>(gridDim.x) * kPowBlock;
  for (size_t i = static_cast<size_t>(blockIdx.x) * kPowBlock + threadIdx.x; i < n; i += stride)
    p[i] = libm_pow(x[i], y[i], dev_pow_tables);
}

const pow_tables host_pow_tables =
#include "rkx_pow_tables.inc"
    ;

}  // namespace
}  // namespace rkx

using namespace rkx;

extern "C" {

int rkx_pow_device(const double *x_dev, const double *y_dev, double *p_dev, size_t n,
                   void *stream)
{
  if (n == 0)
    return RKX_OK;
  if (!x_dev || !y_dev || !p_dev)
    return fail(RKX_EINVAL, "rkx_pow_device: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!s)
    if (int rc = default_stream(&s))
      return rc;
  const size_t blocks = (n + kPowBlock - 1) / kPowBlock;
  const unsigned int grid = static_cast<unsigned int>(blocks < 65536 ? blocks : 65536);
  hipLaunchKernelGGL(k_pow, dim3(grid), dim3(kPowBlock), 0, s, x_dev, y_dev, p_dev, n);
  RKX_HIP(hipGetLastError());
  RKX_HIP(hipStreamSynchronize(s));
  return RKX_OK;
}

int rkx_pow_host(const double *x, const double *y, double *p, size_t n)
{
  if (n > 0 && (!x || !y || !p))
    return fail(RKX_EINVAL, "rkx_pow_host: null pointer");
  for (size_t i = 0; i < n; ++i)
    p[i] = libm_pow(x[i], y[i], host_pow_tables);
  return RKX_OK;
}

}  // extern "C"
