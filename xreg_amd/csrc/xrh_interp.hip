// xrh_interp.hip -- line integrals with the CPU caster's other sampling options.
//
// RayCasterLineIntCPU builds one ITK interpolator per range and evaluates it at every sample
// (lib/ray_cast/xregRayCastLineIntCPU.cpp:115-147, 270-277): nearest neighbour, windowed sinc
// (radius 4, Lanczos window, constant-zero boundary) or an order-3 B-spline besides the linear
// one; with an anti-aliasing factor k >= 1 it casts k rays per pixel through jittered detector
// indices and averages them (:150-183). The rays here are the reference's own float set-up
// (setup_ray_f32) and samples advance by repeated float addition, so every sample is taken at
// the reference's continuous index.
//
// Mapping onto CDNA4: each interpolator reads a grid built once per volume from the stored
// voxels, with an apron wide enough that no tap needs a bounds test:
//   - nearest neighbour and sinc: the float voxels with a zero apron of 4 (the sinc taps reach
//     floor(x) - 3 .. floor(x) + 4; zeros are the constant boundary);
//   - B-spline: the cubic B-spline coefficients in double (mirror-boundary recursive filter run
//     on the device, one thread per grid line per axis) with an apron of 2 holding the mirrored
//     coefficients, so the 4x4x4 taps are plain loads.
// Taps are x-contiguous, so each grid row of a sample is one run of consecutive buffer loads.
// Sinc and B-spline weights and sums are FP64, as ITK evaluates them, and each sample is
// rounded to float once: the sums cancel (negative sinc lobes, B-spline coefficients that
// overshoot at edges), and a faint sample beside bright tissue would otherwise lose digits
// that a maximum-intensity projection keeps. Weights are separable and the sums run row ->
// plane -> volume. Grids go through buffer resources (a stray offset reads 0, never faults).
// One lane casts one ray over the same band-major, XCD-aware tile order as the trilinear
// kernels.
#include "xrh_internal.h"
#include "xrh_ray.h"

#include <climits>
#include <cmath>
#include <mutex>

namespace xrh {
namespace {

// The value the layout stores for voxel (x, y, z) (0 <= x < X, ...), as a float: every layout
// keeps the voxel itself (FP32, or its 16-bit rounding) as the first component of the entry
// at padded index (x + 1, y + 1, z + 1).
__device__ inline float layout_voxel(const VolDesc& d, int x, int y, int z)
{
  const int64_t a = x + 1, b = y + 1, c = z + 1;
  if (d.layout == kLayoutBrick) {
    const int64_t brick = (c >> 2) * (int64_t)d.pitch_z + (b >> 2) * (int64_t)d.pitch_y + (a >> 2);
    const int64_t local = (c & 3) * (kBrickEdge + 1) * kBrickEdge + (b & 3) * (kBrickEdge + 1) + (a & 3);
    return ((const float4*)d.data)[brick * kBrickFaces + local].x;
  }
  const int64_t e = a + (int64_t)d.pitch_y * b + (int64_t)d.pitch_z * c;
  if (d.layout == kLayoutLinear) {
    if (d.storage == XRH_F32) return ((const float*)d.data)[e];
    const uint32_t bits = ((const uint16_t*)d.data)[e];
    return (d.storage == XRH_F16) ? h2f(bits) : bf2f(bits);
  }
  if (d.layout == kLayoutQuad || d.layout == kLayoutFace) return ((const float4*)d.data)[e].x;
  if (d.layout == kLayoutCoef) return ((const float4*)d.data)[2 * e].x;
  const uint32_t bits = ((const uint4*)d.data)[e].x & 0xFFFFu;  // oct: corner (0, 0, 0) in the low half
  return (d.storage == XRH_F16) ? h2f(bits) : bf2f(bits);
}

// Voxels with a zero apron of A: one thread per grid element.
__global__ void voxel_grid_kernel(VolDesc d, int A, float* __restrict__ out, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t GX = d.X + 2 * A, GY = d.Y + 2 * A;
  const int x = (int)(i % GX) - A, y = (int)((i / GX) % GY) - A, z = (int)(i / (GX * GY)) - A;
  const bool in = x >= 0 && x < d.X && y >= 0 && y < d.Y && z >= 0 && z < d.Z;
  out[i] = in ? layout_voxel(d, x, y, z) : 0.0f;
}

// The voxels in double, x fastest: the B-spline filter's input.
__global__ void voxels_to_double_kernel(VolDesc d, double* __restrict__ out, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = (int)(i % d.X), y = (int)((i / d.X) % d.Y), z = (int)(i / ((int64_t)d.X * d.Y));
  out[i] = (double)layout_voxel(d, x, y, z);
}

// Cubic B-spline coefficients of one grid line in place (itk::BSplineDecompositionImageFilter,
// spline order 3: one pole z = sqrt(3) - 2, mirror boundaries -- Unser, "Splines: a perfect
// fit for signal and image processing", IEEE SPM 1999, Box 2). Line t of `lines` starts at
// element (t % na) * sa + (t / na) * sb and has n elements `stride` apart. The causal start
// sums the mirrored series over `horizon` terms when that is shorter than the line (terms
// below the 1e-10 tolerance), else in closed form over the whole line (z_n1 = z^(n-1)).
__global__ void bspline_lines_kernel(double* __restrict__ c, int64_t n, int64_t stride, int64_t lines,
                                     int64_t na, int64_t sa, int64_t sb, double z, int64_t horizon, double z_n1)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= lines || n == 1) return;
  double* p = c + (t % na) * sa + (t / na) * sb;
  const double gain = (1.0 - z) * (1.0 - 1.0 / z);
  for (int64_t i = 0; i < n; ++i) p[i * stride] *= gain;
  if (horizon < n) {
    double sum = p[0], zk = z;
    for (int64_t i = 1; i < horizon; ++i) {
      sum += zk * p[i * stride];
      zk *= z;
    }
    p[0] = sum;
  } else {
    const double iz = 1.0 / z;
    double zk = z, z2k = z_n1;
    double sum = p[0] + z2k * p[(n - 1) * stride];
    z2k *= z2k * iz;
    for (int64_t i = 1; i <= n - 2; ++i) {
      sum += (zk + z2k) * p[i * stride];
      zk *= z;
      z2k *= iz;
    }
    p[0] = sum / (1.0 - zk * zk);
  }
  for (int64_t i = 1; i < n; ++i) p[i * stride] += z * p[(i - 1) * stride];
  p[(n - 1) * stride] = (z / (z * z - 1.0)) * (z * p[(n - 2) * stride] + p[(n - 1) * stride]);
  for (int64_t i = n - 2; i >= 0; --i) p[i * stride] = z * (p[(i + 1) * stride] - p[i * stride]);
}

// ITK's mirror boundary for the B-spline support (BSplineInterpolateImageFunction::
// ApplyMirrorBoundaryConditions): one reflection about index 0 and one about n - 1. ITK leaves
// the range for lines shorter than 4; the clamp keeps those in it.
__device__ inline int64_t mirror_index(int64_t i, int64_t n)
{
  if (n == 1) return 0;
  if (i < 0) i = -i;
  if (i > n - 1) i = 2 * (n - 1) - i;
  return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
}

// B-spline coefficients with a mirrored apron of A.
__global__ void mirror_grid_kernel(const double* __restrict__ c, int X, int Y, int Z, int A,
                                   double* __restrict__ out, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t GX = X + 2 * A, GY = Y + 2 * A;
  const int64_t x = mirror_index(i % GX - A, X), y = mirror_index((i / GX) % GY - A, Y),
                z = mirror_index(i / (GX * GY) - A, Z);
  out[i] = c[x + (int64_t)X * (y + (int64_t)Y * z)];
}

// ---------------------------------------------------------------------------------------
// samplers

struct GridView {
  __amdgpu_buffer_rsrc_t rsrc;
  uint32_t py, pz;
  int X, Y, Z;
};

__device__ inline float grid_load(const GridView& g, uint32_t e)
{
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(g.rsrc, (int)(e * 4u), 0, 0));
}

__device__ inline double grid_load_d(const GridView& g, uint32_t e)
{
  typedef unsigned int u2 __attribute__((ext_vector_type(2)));
  const u2 r = __builtin_amdgcn_raw_buffer_load_b64(g.rsrc, (int)(e * 8u), 0, 0);
  return __hiloint2double((int)r.y, (int)r.x);
}

// itk::Math::RoundHalfIntegerUp of a float index as ITK 5.1 computes it on x86-64 with SSE2:
// round-to-nearest-even of 2x + 0.5 (two float additions), then an arithmetic shift right.
__device__ inline int round_half_up(float x) { return ((int)__builtin_rintf(x + x + 0.5f)) >> 1; }

// itk::Math::Floor of a float index, the same way: round-to-nearest-even of 2x - 0.5, shifted.
__device__ inline int itk_floor(float x) { return ((int)__builtin_rintf(x + x - 0.5f)) >> 1; }

// LinearInterpolateImageFunction::EvaluateAtContinuousIndex (zero apron A = 4): per axis the
// base floor(x) clamped to the volume, the float fraction x - base, and an upper neighbour that
// falls back to the base past the last voxel (a fraction <= 0 weighs it 0); corners and lerps
// in double, x then y then z, rounded to float once. Only the anti-aliased casts use it: plain
// linear casts run the trilinear kernels.
__device__ inline float sample_linear(const GridView& g, float x, float y, float z)
{
  const float p[3] = {x, y, z};
  const int n[3] = {g.X, g.Y, g.Z};
  int b[3], u[3];
  double w[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int bk = min(max((int)floorf(p[k]), 0), n[k] - 1);
    const float dk = p[k] - (float)bk;
    b[k] = bk + 4;
    u[k] = ((bk + 1 > n[k] - 1) ? bk : bk + 1) + 4;
    w[k] = (dk > 0.0f) ? (double)dk : 0.0;
  }
  auto V = [&](int i, int j, int k) -> double {
    return (double)grid_load(g, (uint32_t)i + g.py * (uint32_t)j + g.pz * (uint32_t)k);
  };
  const double v000 = V(b[0], b[1], b[2]), v100 = V(u[0], b[1], b[2]);
  const double v010 = V(b[0], u[1], b[2]), v110 = V(u[0], u[1], b[2]);
  const double v001 = V(b[0], b[1], u[2]), v101 = V(u[0], b[1], u[2]);
  const double v011 = V(b[0], u[1], u[2]), v111 = V(u[0], u[1], u[2]);
  const double x00 = v000 + (v100 - v000) * w[0];
  const double x10 = v010 + (v110 - v010) * w[0];
  const double x01 = v001 + (v101 - v001) * w[0];
  const double x11 = v011 + (v111 - v011) * w[0];
  const double xx0 = x00 + (x10 - x00) * w[1];
  const double xx1 = x01 + (x11 - x01) * w[1];
  return (float)(xx0 + (xx1 - xx0) * w[2]);
}

// NearestNeighborInterpolateImageFunction::EvaluateAtContinuousIndex: the voxel at the rounded
// index (zero apron A = 4).
__device__ inline float sample_nn(const GridView& g, float x, float y, float z)
{
  const uint32_t ix = (uint32_t)(round_half_up(x) + 4), iy = (uint32_t)(round_half_up(y) + 4),
                 iz = (uint32_t)(round_half_up(z) + 4);
  return grid_load(g, ix + g.py * iy + g.pz * iz);
}

// Lanczos-4 windowed-sinc weights of the 8 taps floor(x) - 3 .. floor(x) + 4 at fractional
// offset d = x - floor(x): w[i] = sinc(t) * sinc(t / 4) at t = d + 3 - i, with sinc(u) =
// sin(pi u) / (pi u), which is WindowedSincInterpolateImageFunction's Sinc(t) times
// LanczosWindowFunction<4>(t). sin(pi t) = (-1)^k sin(pi d) and sin(pi t / 4) follows from
// sin(pi d / 4), cos(pi d / 4) and the angle k pi / 4 (k = 3 - i), so three sinpi/cospi calls
// give all eight. At d = 0 the weights are the unit impulse on tap 3, as ITK special-cases.
// Double precision throughout (d + 3 - i is exact).
__device__ inline void lanczos4_weights(double d, double w[8])
{
  if (d == 0.0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) w[i] = (i == 3) ? 1.0 : 0.0;
    return;
  }
  constexpr double h = 0.70710678118654752440;  // sqrt(1/2)
  constexpr double kCos[8] = {-h, 0.0, h, 1.0, h, 0.0, -h, -1.0};  // cos(k pi / 4), k = 3 - i
  constexpr double kSin[8] = {h, 1.0, h, 0.0, -h, -1.0, -h, 0.0};  // sin(k pi / 4)
  constexpr double kFourOverPi2 = 0.40528473456935108578;         // 4 / pi^2
  const double s = sinpi(d), sq = sinpi(0.25 * d), cq = cospi(0.25 * d);
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const double t = d + (double)(3 - i);
    const double st = ((3 - i) & 1) ? -s : s;
    const double sq_t = fma(sq, kCos[i], cq * kSin[i]);
    w[i] = (kFourOverPi2 * st * sq_t) / (t * t);
  }
}

// WindowedSincInterpolateImageFunction<Image<float,3>, 4, LanczosWindowFunction<4>,
// ConstantBoundaryCondition>::EvaluateAtContinuousIndex: the 8x8x8 taps around floor(x) with
// separable weights (zero apron A = 4: tap floor(x) - 3 + i sits at grid index floor(x) + 1 + i).
__device__ inline float sample_sinc(const GridView& g, float x, float y, float z)
{
  const int bx = itk_floor(x), by = itk_floor(y), bz = itk_floor(z);
  double wx[8], wy[8], wz[8];
  lanczos4_weights((double)x - (double)bx, wx);
  lanczos4_weights((double)y - (double)by, wy);
  lanczos4_weights((double)z - (double)bz, wz);
  const uint32_t e0 = (uint32_t)(bx + 1) + g.py * (uint32_t)(by + 1) + g.pz * (uint32_t)(bz + 1);
  double acc = 0.0;
  for (int k = 0; k < 8; ++k) {
    double plane = 0.0;
    for (int j = 0; j < 8; ++j) {
      const uint32_t e = e0 + g.pz * (uint32_t)k + g.py * (uint32_t)j;
      double row = 0.0;
#pragma unroll
      for (int i = 0; i < 8; ++i) row = fma(wx[i], (double)grid_load(g, e + (uint32_t)i), row);
      plane = fma(wy[j], row, plane);
    }
    acc = fma(wz[k], plane, acc);
  }
  return (float)acc;
}

// Uniform cubic B-spline weights of the taps floor(x) - 1 .. floor(x) + 2 at t = x - floor(x).
__device__ inline void bspline3_weights(double t, double w[4])
{
  const double t2 = t * t, t3 = t2 * t, u = 1.0 - t;
  w[0] = (u * u * u) * (1.0 / 6.0);
  w[1] = fma(0.5, t3, 2.0 / 3.0 - t2);
  w[2] = fma(-0.5, t3, fma(0.5, t2, fma(0.5, t, 1.0 / 6.0)));
  w[3] = t3 * (1.0 / 6.0);
}

// BSplineInterpolateImageFunction (order 3)::EvaluateAtContinuousIndex: the 4x4x4 coefficients
// around floor(x) with separable weights (mirrored apron A = 2: coefficient floor(x) - 1 + i
// sits at grid index floor(x) + 1 + i).
__device__ inline float sample_bspline(const GridView& g, float x, float y, float z)
{
  const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
  double wx[4], wy[4], wz[4];
  bspline3_weights((double)(x - fx), wx);
  bspline3_weights((double)(y - fy), wy);
  bspline3_weights((double)(z - fz), wz);
  const uint32_t e0 = (uint32_t)((int)fx + 1) + g.py * (uint32_t)((int)fy + 1) + g.pz * (uint32_t)((int)fz + 1);
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double plane = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t e = e0 + g.pz * (uint32_t)k + g.py * (uint32_t)j;
      double row = 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i) row = fma(wx[i], grid_load_d(g, e + (uint32_t)i), row);
      plane = fma(wy[j], row, plane);
    }
    acc = fma(wz[k], plane, acc);
  }
  return (float)acc;
}

template <int kInterp>
__device__ inline float sample_interp(const GridView& g, float x, float y, float z)
{
  if constexpr (kInterp == XRH_INTERP_LINEAR) return sample_linear(g, x, y, z);
  else if constexpr (kInterp == XRH_INTERP_NN) return sample_nn(g, x, y, z);
  else if constexpr (kInterp == XRH_INTERP_SINC) return sample_sinc(g, x, y, z);
  else return sample_bspline(g, x, y, z);
}

// ---------------------------------------------------------------------------------------
// anti-aliasing jitter

// The SplitMix64 output function (Steele, Lea and Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014).
__device__ inline uint64_t splitmix64(uint64_t state)
{
  uint64_t z = state;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Jitter of ray k of pixel `pixel` (the reference's range index: projection * rows * cols + row
// * cols + col), component 0 (column) or 1 (row): uniform on [-0.5, 0.5) in steps of 2^-24, as
// std::uniform_real_distribution<float>(-0.5, 0.5) draws it (xregRayCastLineIntCPU.cpp:162-174).
// The reference seeds a Mersenne Twister from std::random_device, so its jitter is never
// reproducible; here it is a pure function of (seed, pixel, ray, component): counter-based
// SplitMix64, the same on the device and in any process.
__device__ inline float aa_jitter(uint64_t seed, uint64_t pixel, int k, int comp)
{
  const uint64_t counter = (pixel << 17) | ((uint64_t)k << 1) | (uint64_t)comp;
  const uint64_t h = splitmix64(seed + 0x9E3779B97F4A7C15ull * (counter + 1));
  return (float)(uint32_t)(h >> 40) * 0x1.0p-24f - 0.5f;
}

// One lane = one pixel: `rays` rays (1 without anti-aliasing), each the reference's float set-up
// through the (jittered) detector index, n + 1 samples at positions advanced by repeated float
// addition (:270-277), times the step; the rays' values are averaged as the reference does
// (aa_sum += sum * (1 / rays), :281-282) and stored as store_pixel does (RayCasterCPU::
// pre_compute semantics: ACCUM onto the buffer / background, REPLACE onto fill, :285).
template <int kOp, int kInterp>
__global__ __launch_bounds__(kThreads) void drr_interp_kernel(
    const DrrItem* __restrict__ items, const uint32_t* __restrict__ band_start, int64_t num_items,
    int bands, float* __restrict__ out, int accumulate, float fill, InterpGrid grid, int aa_rays,
    uint64_t aa_seed)
{
  const TileRef tr = decode_tile(band_start, num_items, bands, items);
  if (!tr.valid) return;
  const DrrItem& it = items[tr.item];
  int col, row;
  lane_pixel(tr.tx, tr.ty, 0, col, row);
  if (col >= it.cam.cols || row >= it.cam.rows) return;
  GridView g;
  g.rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(grid.data), (short)0, (int)grid.bytes, 0x00020000);
  g.py = grid.py;
  g.pz = grid.pz;
  g.X = grid.X;
  g.Y = grid.Y;
  g.Z = grid.Z;
  const int64_t o = it.out_off + (int64_t)row * it.cam.cols + col;
  const int rays = aa_rays > 0 ? aa_rays : 1;
  const float inv = 1.0f / (float)rays;
  float aa_sum = 0.0f;
  for (int k = 0; k < rays; ++k) {
    float c = (float)col, r = (float)row;
    if (aa_rays > 0) {
      c = c + aa_jitter(aa_seed, (uint64_t)o, k, 0);
      r = r + aa_jitter(aa_seed, (uint64_t)o, k, 1);
    }
    float acc = (kOp == XRH_LINE_INT_SUM) ? 0.0f : -3.402823466e+38f;
    const Ray ray = setup_ray_f32(it, c, r);
    if (ray.n >= 0) {
      float px = ray.sx, py = ray.sy, pz = ray.sz;
      for (int i = 0; i <= ray.n; ++i) {
        acc = line_op<kOp>(acc, sample_interp<kInterp>(g, px, py, pz));
        px = px + ray.dx;
        py = py + ray.dy;
        pz = pz + ray.dz;
      }
      acc = acc * it.step;
    }
    aa_sum = aa_sum + acc * inv;
  }
  const float prev = accumulate ? out[o] : fill;
  if constexpr (kOp == XRH_LINE_INT_SUM) out[o] = prev + aa_sum; else out[o] = (prev < aa_sum) ? aa_sum : prev;
}

template <int kOp, int kInterp>
void launch_interp_kernel(const DrrLaunch& L, const InterpGrid& g, dim3 grid, hipStream_t s, int accumulate,
                          int aa_rays, uint64_t aa_seed)
{
  hipLaunchKernelGGL((drr_interp_kernel<kOp, kInterp>), grid, dim3(kThreads), 0, s, L.items_dev, L.band_start_dev,
                     L.num_items, L.bands, L.out, accumulate, L.fill, g, aa_rays, aa_seed);
}

template <int kOp>
void launch_interp_op(const DrrLaunch& L, const InterpGrid& g, int32_t interp, dim3 grid, hipStream_t s,
                      int accumulate, int aa_rays, uint64_t aa_seed)
{
  switch (interp) {
    case XRH_INTERP_LINEAR:
      launch_interp_kernel<kOp, XRH_INTERP_LINEAR>(L, g, grid, s, accumulate, aa_rays, aa_seed);
      break;
    case XRH_INTERP_NN:
      launch_interp_kernel<kOp, XRH_INTERP_NN>(L, g, grid, s, accumulate, aa_rays, aa_seed);
      break;
    case XRH_INTERP_SINC:
      launch_interp_kernel<kOp, XRH_INTERP_SINC>(L, g, grid, s, accumulate, aa_rays, aa_seed);
      break;
    default:
      launch_interp_kernel<kOp, XRH_INTERP_BSPLINE>(L, g, grid, s, accumulate, aa_rays, aa_seed);
      break;
  }
}

std::mutex& grid_mutex()
{
  static std::mutex mu;
  return mu;
}

}  // namespace

const InterpGrid& interp_prepare(xrh_volume* v, int32_t interp)
{
  XRH_REQUIRE(v, "null volume");
  XRH_REQUIRE(interp >= XRH_INTERP_LINEAR && interp <= XRH_INTERP_BSPLINE, "bad interpolation method");
  const int k = (interp == XRH_INTERP_BSPLINE) ? kGridBSpline : kGridVoxels;
  std::lock_guard<std::mutex> lk(grid_mutex());
  if (v->grid_mem[k]) return v->grid[k];
  xrh_ctx* ctx = v->ctx;
  const VolDesc& d = v->desc;
  const int A = (k == kGridBSpline) ? 2 : 4;
  const int64_t X = d.X, Y = d.Y, Z = d.Z;
  const int64_t GX = X + 2 * A, GY = Y + 2 * A, GZ = Z + 2 * A;
  const int64_t n = GX * GY * GZ;
  const uint64_t esize = (k == kGridBSpline) ? 8 : 4;
  const uint64_t bytes = (uint64_t)n * esize;
  // grid offsets are 32-bit (buffer loads) and the tap arithmetic is unsigned 32-bit
  XRH_REQUIRE(bytes < 0xFFFFF000ull, "interpolation grid exceeds 4 GiB");
  const int threads = 256;
  void* mem = ctx_alloc(ctx, bytes);
  double* coef = nullptr;
  try {
    if (k == kGridVoxels) {
      hipLaunchKernelGGL(voxel_grid_kernel, dim3((unsigned)((n + threads - 1) / threads)), dim3(threads), 0,
                         ctx->stream, d, A, (float*)mem, n);
      XRH_HIP_CHECK(hipGetLastError());
    } else {
      const int64_t nv = X * Y * Z;
      coef = (double*)ctx_alloc(ctx, (uint64_t)nv * 8);
      hipLaunchKernelGGL(voxels_to_double_kernel, dim3((unsigned)((nv + threads - 1) / threads)), dim3(threads), 0,
                         ctx->stream, d, coef, nv);
      XRH_HIP_CHECK(hipGetLastError());
      const double z = std::sqrt(3.0) - 2.0;
      const int64_t horizon = (int64_t)std::ceil(std::log(1.0e-10) / std::log(std::fabs(z)));
      const int64_t len[3] = {X, Y, Z}, stride[3] = {1, X, X * Y};
      const int64_t na[3] = {Y * Z, X, X * Y}, sa[3] = {X, 1, 1}, sb[3] = {0, X * Y, 0};
      for (int ax = 0; ax < 3; ++ax) {  // x, then y, then z, as ITK filters the dimensions
        const int64_t lines = nv / len[ax];
        const int lt = 64;
        hipLaunchKernelGGL(bspline_lines_kernel, dim3((unsigned)((lines + lt - 1) / lt)), dim3(lt), 0, ctx->stream,
                           coef, len[ax], stride[ax], lines, na[ax], sa[ax], sb[ax], z, horizon,
                           std::pow(z, (double)(len[ax] - 1)));
        XRH_HIP_CHECK(hipGetLastError());
      }
      hipLaunchKernelGGL(mirror_grid_kernel, dim3((unsigned)((n + threads - 1) / threads)), dim3(threads), 0,
                         ctx->stream, (const double*)coef, (int)X, (int)Y, (int)Z, A, (double*)mem, n);
      XRH_HIP_CHECK(hipGetLastError());
    }
    XRH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (coef) ctx_free(ctx, coef);
    coef = nullptr;
  } catch (...) {
    (void)hipStreamSynchronize(ctx->stream);
    if (coef) ctx_free(ctx, coef);
    ctx_free(ctx, mem);
    throw;
  }
  v->grid_mem[k] = mem;
  v->grid_bytes[k] = bytes;
  InterpGrid& g = v->grid[k];
  g.data = mem;
  g.bytes = (uint32_t)bytes;
  g.py = (uint32_t)GX;
  g.pz = (uint32_t)(GX * GY);
  g.apron = A;
  g.X = (int32_t)X;
  g.Y = (int32_t)Y;
  g.Z = (int32_t)Z;
  return g;
}

void launch_drr_interp(const DrrLaunch& L, const InterpGrid& g, int32_t interp, int32_t aa_rays, uint64_t aa_seed,
                       hipStream_t s)
{
  if (L.num_items <= 0 || L.max_band_tiles <= 0) return;
  XRH_REQUIRE(interp >= XRH_INTERP_LINEAR && interp <= XRH_INTERP_BSPLINE, "bad interpolation method");
  XRH_REQUIRE(aa_rays >= 0 && aa_rays <= XRH_MAX_AA_RAYS, "anti-aliasing rays out of range");
  XRH_REQUIRE(g.data && g.apron == ((interp == XRH_INTERP_BSPLINE) ? 2 : 4), "interpolation grid not prepared");
  const int64_t nblocks = L.max_band_tiles * L.bands;
  XRH_REQUIRE(nblocks < (int64_t)INT32_MAX, "too many DRR tiles in one launch");
  const dim3 grid((unsigned)nblocks);
  const int accumulate = (L.store == XRH_STORE_ACCUM) ? 1 : 0;
  if (L.kernel == XRH_LINE_INT_SUM) launch_interp_op<XRH_LINE_INT_SUM>(L, g, interp, grid, s, accumulate, aa_rays, aa_seed);
  else launch_interp_op<XRH_LINE_INT_MAX>(L, g, interp, grid, s, accumulate, aa_rays, aa_seed);
  XRH_HIP_CHECK(hipGetLastError());
}

}  // namespace xrh
