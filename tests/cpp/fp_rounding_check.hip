// fp_rounding_check.hip -- checks on the GPU that the float division / square root forms used
// by the FP32 ray set-up round like the host's IEEE arithmetic (test infrastructure).
// Prints one JSON line with mismatch counts per form over N random operand pairs.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

__global__ void ops(const float* a, const float* b, int n, float* out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = a[i], y = b[i];
  out[6 * i + 0] = __fsqrt_rn(x);
  out[6 * i + 1] = sqrtf(x);
  out[6 * i + 2] = (float)sqrt((double)x);
  out[6 * i + 3] = __fdiv_rn(x, y);
  out[6 * i + 4] = x / y;
  out[6 * i + 5] = (float)((double)x / (double)y);
}

int main()
{
  const int n = 1 << 22;
  std::mt19937 rng(12345);
  std::uniform_real_distribution<float> ua(1.0e-3f, 3.0e6f), ub(1.0e-3f, 2.0e3f);
  std::vector<float> a(n), b(n), out(6 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    a[i] = ua(rng);
    b[i] = ub(rng);
  }
  float *da, *db, *dout;
  if (hipMalloc(&da, n * 4) || hipMalloc(&db, n * 4) || hipMalloc(&dout, 6 * (size_t)n * 4)) return 3;
  hipMemcpy(da, a.data(), n * 4, hipMemcpyHostToDevice);
  hipMemcpy(db, b.data(), n * 4, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(ops, dim3((n + 255) / 256), dim3(256), 0, 0, da, db, n, dout);
  if (hipDeviceSynchronize() != hipSuccess) return 4;
  hipMemcpy(out.data(), dout, 6 * (size_t)n * 4, hipMemcpyDeviceToHost);
  long mism[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < n; ++i) {
    const float s = std::sqrt(a[i]);
    const float q = a[i] / b[i];
    for (int k = 0; k < 3; ++k) mism[k] += (out[6 * (size_t)i + k] != s);
    for (int k = 3; k < 6; ++k) mism[k] += (out[6 * (size_t)i + k] != q);
  }
  std::printf("{\"n\": %d, \"fsqrt_rn\": %ld, \"sqrtf\": %ld, \"sqrt_via_f64\": %ld, \"fdiv_rn\": %ld, "
              "\"div\": %ld, \"div_via_f64\": %ld}\n",
              n, mism[0], mism[1], mism[2], mism[3], mism[4], mism[5]);
  hipFree(da);
  hipFree(db);
  hipFree(dout);
  return 0;
}
