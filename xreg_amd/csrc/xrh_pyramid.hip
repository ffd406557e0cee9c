// xrh_pyramid.hip -- fixed-image downsampling for multi-resolution registration.
//
// DownsampleImage (lib/itk/xregITKResampleUtils.h:51-120, the default B-spline overload
// :181-186), as DownsampleProjData applies it to each fixed image of a level
// (lib/image/xregProjData.cpp:40-50) with the detector downsampled alongside
// (DownsampleCameraModel). The steps are ITK 5.1's (the reference's pinned ITK):
//   1. itk::DiscreteGaussianImageFilter, variance sigma^2 in pixels (UseImageSpacing off),
//      sigma = 0.5 / factor by default, MaximumError 0.01, MaximumKernelWidth 32: the 1-D
//      kernel of itk::GaussianOperator (Lindeberg's discrete Gaussian e^-t I_n(t) built by the
//      Bessel recurrence and truncated once it holds 1 - MaximumError of the mass, then
//      normalised), applied along x then y with edge replication (ZeroFluxNeumann), double
//      intermediate, float result;
//   2. itk::BSplineDecompositionImageFilter (order 3, pole sqrt(3) - 2, mirror boundary,
//      tolerance 1e-10), along x then y, in double;
//   3. itk::ResampleImageFilter with an identity transform: output size
//      floor(n * factor + 0.5), spacing / factor, same origin; each output pixel evaluates the
//      cubic B-spline at its continuous input index (itk::BSplineInterpolateImageFunction,
//      mirror boundary) and is rounded to float.
// Smoothing is skipped when factor >= 1 or |sigma| <= 1e-6 (as DownsampleImage does).
#include "xrh_internal.h"

#include <cmath>
#include <vector>

namespace xrh {
namespace {

// itk::GaussianOperator::ModifiedBesselI0 / I1: the polynomial approximations of Abramowitz and
// Stegun 9.8.1-9.8.4 in Horner form.
double bessel_i0(double y)
{
  const double d = std::fabs(y);
  if (d < 3.75) {
    double m = y / 3.75;
    m *= m;
    return 1.0 + m * (3.5156229 + m * (3.0899424 + m * (1.2067492 + m * (0.2659732 + m * (0.360768e-1 +
                                                                                              m * 0.45813e-2)))));
  }
  const double m = 3.75 / d;
  return (std::exp(d) / std::sqrt(d)) *
         (0.39894228 +
          m * (0.1328592e-1 +
               m * (0.225319e-2 +
                    m * (-0.157565e-2 +
                         m * (0.916281e-2 +
                              m * (-0.2057706e-1 + m * (0.2635537e-1 + m * (-0.1647633e-1 + m * 0.392377e-2))))))));
}

double bessel_i1(double y)
{
  const double d = std::fabs(y);
  double acc;
  if (d < 3.75) {
    double m = y / 3.75;
    m *= m;
    acc = d * (0.5 + m * (0.87890594 + m * (0.51498869 + m * (0.15084934 + m * (0.2658733e-1 +
                                                                                m * (0.301532e-2 + m * 0.32411e-3))))));
  } else {
    const double m = 3.75 / d;
    acc = 0.2282967e-1 + m * (-0.2895312e-1 + m * (0.1787654e-1 - m * 0.420059e-2));
    acc = 0.39894228 + m * (-0.3988024e-1 + m * (-0.362018e-2 + m * (0.163801e-2 + m * (-0.1031555e-1 + m * acc))));
    acc *= (std::exp(d) / std::sqrt(d));
  }
  return (y < 0.0) ? -acc : acc;
}

// itk::GaussianOperator::GenerateCoefficients: symmetric kernel of 2r+1 taps (index r is the
// centre).
std::vector<double> gaussian_kernel(double variance, double max_error, unsigned max_width)
{
  std::vector<double> half;
  const double et = std::exp(-variance);
  const double cap = 1.0 - max_error;
  double sum = 0.0;
  half.push_back(et * bessel_i0(variance));
  sum += half[0];
  half.push_back(et * bessel_i1(variance));
  sum += half[1] * 2.0;
  for (int i = 2; sum < cap; ++i) {
    half.push_back(half[i - 2] - 2 * (i - 1) / variance * half[i - 1]);
    sum += half[i] * 2;
    if (half[i] <= 0.0) break;
    if (half.size() > max_width) break;
  }
  for (double& c : half) c /= sum;
  const int r = (int)half.size() - 1;
  std::vector<double> k(2 * r + 1);
  for (int i = 0; i <= r; ++i) k[r + i] = k[r - i] = half[i];
  return k;
}

__global__ void gauss_x_kernel(const float* __restrict__ img, double* __restrict__ out, int rows, int cols,
                               const double* __restrict__ w, int R)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (c >= cols || r >= rows) return;
  double s = 0.0;
  for (int k = -R; k <= R; ++k) {
    const int cc = min(max(c + k, 0), cols - 1);
    s += w[k + R] * (double)img[(int64_t)r * cols + cc];
  }
  out[(int64_t)r * cols + c] = s;
}

__global__ void gauss_y_kernel(const double* __restrict__ in, float* __restrict__ out, int rows, int cols,
                               const double* __restrict__ w, int R)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
  if (c >= cols || r >= rows) return;
  double s = 0.0;
  for (int k = -R; k <= R; ++k) {
    const int rr = min(max(r + k, 0), rows - 1);
    s += w[k + R] * in[(int64_t)rr * cols + c];
  }
  out[(int64_t)r * cols + c] = (float)s;
}

__global__ void to_double_kernel(const float* __restrict__ in, double* __restrict__ out, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (double)in[i];
}

// BSplineDecompositionImageFilter::DataToCoefficients1D for the cubic pole, on the line
// c[0], c[stride], ..., c[(n-1) * stride].
__device__ void bspline_line(double* c, int n, int64_t stride)
{
  if (n == 1) return;
  const double z = 1.7320508075688772 - 2.0;  // sqrt(3) - 2
#define C_(i) c[(int64_t)(i) * stride]
  const double gain = (1.0 - z) * (1.0 - 1.0 / z);
  for (int i = 0; i < n; ++i) C_(i) *= gain;
  // initial causal coefficient (SetInitialCausalCoefficient), tolerance 1e-10
  const int horizon = (int)ceil(log(1.0e-10) / log(fabs(z)));
  double zn = z;
  if (horizon < n) {
    double sum = C_(0);
    for (int i = 1; i < horizon; ++i) {
      sum += zn * C_(i);
      zn *= z;
    }
    C_(0) = sum;
  } else {
    const double iz = 1.0 / z;
    double z2n = pow(z, (double)(n - 1));
    double sum = C_(0) + z2n * C_(n - 1);
    z2n *= z2n * iz;
    for (int i = 1; i <= n - 2; ++i) {
      sum += (zn + z2n) * C_(i);
      zn *= z;
      z2n *= iz;
    }
    C_(0) = sum / (1.0 - zn * zn);
  }
  for (int i = 1; i < n; ++i) C_(i) += z * C_(i - 1);
  // initial anti-causal coefficient, then the anti-causal recursion
  C_(n - 1) = (z / (z * z - 1.0)) * (z * C_(n - 2) + C_(n - 1));
  for (int i = n - 2; i >= 0; --i) C_(i) = z * (C_(i + 1) - C_(i));
#undef C_
}

__global__ void bspline_rows_kernel(double* __restrict__ coef, int rows, int cols)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rows) bspline_line(coef + (int64_t)r * cols, cols, 1);
}

__global__ void bspline_cols_kernel(double* __restrict__ coef, int rows, int cols)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < cols) bspline_line(coef + c, rows, cols);
}

__device__ inline int mirror(int i, int n)
{
  if (n == 1) return 0;
  const int n2 = 2 * n - 2;
  if (i < 0) i = (-i) % n2;
  if (i >= n) {
    i = i % n2;
    if (i >= n) i = n2 - i;
  }
  return i;
}

__device__ inline void cubic_weights(double x, int& base, double w[4])
{
  base = (int)floor(x) - 1;
  const double t = x - (double)(base + 1);
  w[3] = (1.0 / 6.0) * t * t * t;
  w[0] = (1.0 / 6.0) + 0.5 * t * (t - 1.0) - w[3];
  w[2] = t + w[0] - 2.0 * w[3];
  w[1] = 1.0 - w[0] - w[2] - w[3];
}

// One thread per output pixel: continuous input index (phys / spacing with the same origin),
// the region test of ImageFunction::IsInsideBuffer, then the 16-term B-spline sum with the
// x index varying fastest (BSplineInterpolateImageFunction's point order).
__global__ void resample_kernel(const double* __restrict__ coef, int rows, int cols, float* __restrict__ out,
                                int orows, int ocols, double sp_r, double sp_c, double osp_r, double osp_c)
{
#pragma clang fp contract(off)
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (j >= ocols || i >= orows) return;
  const double x = (osp_c * (double)j) * (1.0 / sp_c);
  const double y = (osp_r * (double)i) * (1.0 / sp_r);
  double v = 0.0;
  if (x >= -0.5 && x < cols - 0.5 && y >= -0.5 && y < rows - 0.5) {
    int bx, by;
    double wx[4], wy[4];
    cubic_weights(x, bx, wx);
    cubic_weights(y, by, wy);
    int ix[4], iy[4];
    for (int k = 0; k < 4; ++k) {
      ix[k] = mirror(bx + k, cols);
      iy[k] = mirror(by + k, rows);
    }
    for (int p = 0; p < 16; ++p) {
      const int a = p & 3, b = p >> 2;
      double w = 1.0;
      w *= wx[a];
      w *= wy[b];
      v += w * coef[(int64_t)iy[b] * cols + ix[a]];
    }
  }
  out[(int64_t)i * ocols + j] = (float)v;
}

}  // namespace
}  // namespace xrh

using namespace xrh;

extern "C" {

xrh_status xrh_downsample_size(int32_t rows, int32_t cols, double factor, int32_t* out_rows, int32_t* out_cols)
{
  return guarded([&] {
    XRH_REQUIRE(out_rows && out_cols, "null argument");
    XRH_REQUIRE(rows > 0 && cols > 0 && factor > 0.0, "bad size or factor");
    *out_rows = (int32_t)(unsigned long)((rows * factor) + 0.5);
    *out_cols = (int32_t)(unsigned long)((cols * factor) + 0.5);
    XRH_REQUIRE(*out_rows > 0 && *out_cols > 0, "downsampled image is empty");
  });
}

xrh_status xrh_downsample_image(xrh_ctx* ctx, const float* img, int32_t rows, int32_t cols, double row_spacing,
                                double col_spacing, double factor, double sigma, float* out, int32_t out_rows,
                                int32_t out_cols)
{
  return guarded([&] {
    XRH_REQUIRE(ctx && img && out, "null argument");
    XRH_REQUIRE(rows > 0 && cols > 0 && rows <= 32768 && cols <= 32768, "bad image size");
    XRH_REQUIRE(factor > 0.0 && row_spacing > 0.0 && col_spacing > 0.0, "bad factor or spacing");
    int32_t er = 0, ec = 0;
    if (xrh_downsample_size(rows, cols, factor, &er, &ec) != XRH_OK) fail(XRH_ERR_INVALID, "bad size");
    XRH_REQUIRE(er == out_rows && ec == out_cols, "output size must be xrh_downsample_size's");
    const int64_t n = (int64_t)rows * cols, on = (int64_t)out_rows * out_cols;
    std::vector<double> w;
    const bool smooth = (factor < 1.0) && (std::fabs(sigma) > 1.0e-6);
    if (smooth) {
      const double s = (sigma < 0.0) ? 0.5 / factor : sigma;
      w = gaussian_kernel(s * s, 0.01, 32);
    }
    float* d_img = nullptr;
    float* d_sm = nullptr;
    double* d_tmp = nullptr;
    double* d_w = nullptr;
    float* d_out = nullptr;
    auto cleanup = [&] {
      (void)hipStreamSynchronize(ctx->stream);
      for (void* p : {(void*)d_img, (void*)d_sm, (void*)d_tmp, (void*)d_w, (void*)d_out})
        if (p) ctx_free(ctx, p);
    };
    try {
      d_img = (float*)ctx_alloc(ctx, (uint64_t)n * 4);
      d_sm = (float*)ctx_alloc(ctx, (uint64_t)n * 4);
      d_tmp = (double*)ctx_alloc(ctx, (uint64_t)n * 8);
      d_out = (float*)ctx_alloc(ctx, (uint64_t)on * 4);
      XRH_HIP_CHECK(hipMemcpyAsync(d_img, img, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
      const float* src = d_img;
      if (smooth) {
        d_w = (double*)ctx_alloc(ctx, w.size() * 8);
        XRH_HIP_CHECK(hipMemcpyAsync(d_w, w.data(), w.size() * 8, hipMemcpyHostToDevice, ctx->stream));
        const int R = (int)(w.size() / 2);
        const dim3 g((unsigned)((cols + 255) / 256), (unsigned)rows);
        hipLaunchKernelGGL(gauss_x_kernel, g, dim3(256), 0, ctx->stream, d_img, d_tmp, rows, cols, d_w, R);
        hipLaunchKernelGGL(gauss_y_kernel, g, dim3(256), 0, ctx->stream, d_tmp, d_sm, rows, cols, d_w, R);
        XRH_HIP_CHECK(hipGetLastError());
        src = d_sm;
      }
      // B-spline coefficients of the (smoothed) float image, in double
      hipLaunchKernelGGL(to_double_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, src, d_tmp,
                         n);
      hipLaunchKernelGGL(bspline_rows_kernel, dim3((unsigned)((rows + 63) / 64)), dim3(64), 0, ctx->stream, d_tmp,
                         rows, cols);
      hipLaunchKernelGGL(bspline_cols_kernel, dim3((unsigned)((cols + 63) / 64)), dim3(64), 0, ctx->stream, d_tmp,
                         rows, cols);
      hipLaunchKernelGGL(resample_kernel, dim3((unsigned)((out_cols + 255) / 256), (unsigned)out_rows), dim3(256), 0,
                         ctx->stream, d_tmp, rows, cols, d_out, out_rows, out_cols, row_spacing, col_spacing,
                         row_spacing / factor, col_spacing / factor);
      XRH_HIP_CHECK(hipGetLastError());
      XRH_HIP_CHECK(hipMemcpyAsync(out, d_out, (size_t)on * 4, hipMemcpyDeviceToHost, ctx->stream));
      XRH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  });
}

}  // extern "C"
