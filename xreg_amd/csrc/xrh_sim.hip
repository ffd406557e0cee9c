// xrh_sim.hip -- NCC, gradient-NCC, patch-NCC and patch-gradient-NCC on gfx950.
//
// Semantics (smaller is better, as in xReg):
//   NCC        (1 - ncc)/2, ncc = sum(f~ m~)/(n sf sm), s = max(1e-6, sqrt(var/(n-1)))
//              (lib/regi/sim_metrics_2d/xregImgSimMetric2DNCCCPU.cpp:52-236)
//   Grad-NCC   0.5 (NCC(gx_f, gx_m) + NCC(gy_f, gy_m)) with OpenCV Sobel ksize 3 after an
//              optional GaussianBlur(ksize, sigma 0), BORDER_REFLECT_101
//              (xregImgSimMetric2DGradImgCPU.cpp:79-102, xregImgSimMetric2DGradNCCCPU.cpp:51-65)
//   Patch-NCC  mean over (2r+1)^2 patches (stride s) of 1 - sum(((m-mu_m)/s_m) (f-mu_f)/(s_f N))
//              (xregImgSimMetric2DPatchNCCCPU.cpp:190-287, 558-619)
//   Patch-Grad-NCC  0.5 (PatchNCC(gx) + PatchNCC(gy)) (xregImgSimMetric2DPatchGradNCCCPU.cpp:104-224)
//
// CDNA4 mapping: one 256-lane workgroup per 16x16 output tile of one moving image; the
// tile plus its halo is staged in LDS (REFLECT_101 applied while staging), per-lane FP64
// partial sums are reduced across the wave with cross-lane shuffles, one FP64 record per
// tile goes to a partial buffer, and a one-wave-per-image finish kernel sums the records in
// a fixed order (deterministic, FP64) and applies the metric's closed form.
#include "xrh_internal.h"
#include "xrh_sim.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace xrh {
namespace {

constexpr int kT = 16;            // output tile edge
constexpr int kThreads = kT * kT;
constexpr int kMaxHalo = 4;       // blur radius 3 + Sobel 1
constexpr int kGradLds = kT + 2 * kMaxHalo;  // 24

__device__ inline int reflect101(int p, int n)
{
  if (n == 1) return 0;
  // |p| < 2n always holds for the halos used here; loop is for safety
  while (p < 0 || p >= n) {
    if (p < 0) p = -p;
    if (p >= n) p = 2 * (n - 1) - p;
  }
  return p;
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Reduces K per-lane doubles over the 256-lane workgroup; lane 0 writes them to dst.
template <int K>
__device__ inline void block_reduce_store(double (&v)[K], double* dst)
{
  __shared__ double red[4][K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) red[w][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) dst[k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
  }
}

__device__ inline float gauss_k(int ksize, int j)
{
  // OpenCV getGaussianKernel fixed table for sigma <= 0, odd ksize <= 7
  if (ksize == 3) {
    const float k[3] = {0.25f, 0.5f, 0.25f};
    return k[j];
  } else if (ksize == 5) {
    const float k[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
    return k[j];
  } else if (ksize == 7) {
    const float k[7] = {0.03125f, 0.109375f, 0.21875f, 0.28125f, 0.21875f, 0.109375f, 0.03125f};
    return k[j];
  }
  return 1.0f;
}

// Computes the Sobel gradients (after optional blur) of the 16x16 tile at (x0, y0) of an
// image. Each lane returns gx, gy of its own pixel. LDS: raw (24x24), blurred-rows, blurred.
__device__ inline void tile_gradients(const float* __restrict__ img, int rows, int cols, int x0,
                                      int y0, int ksize, float& gx, float& gy, bool& inside)
{
#pragma clang fp contract(off)
  __shared__ float raw[kGradLds][kGradLds];
  __shared__ float tmp[kGradLds][kGradLds];
  __shared__ float sm[kGradLds][kGradLds];
  const int rb = (ksize > 1) ? ksize / 2 : 0;
  const int halo = rb + 1;
  const int W = kT + 2 * halo;
  for (int i = threadIdx.x; i < W * W; i += kThreads) {
    const int ly = i / W, lx = i % W;
    const int y = reflect101(y0 - halo + ly, rows);
    const int x = reflect101(x0 - halo + lx, cols);
    raw[ly][lx] = img[(int64_t)y * cols + x];
  }
  __syncthreads();
  // blurred image over the (16+2) x (16+2) region that the Sobel needs
  const int B = kT + 2;
  if (rb > 0) {
    // row pass over W rows x B cols: acc = k[r] s0 + sum_j k[r+j] (s[-j] + s[+j])
    for (int i = threadIdx.x; i < W * B; i += kThreads) {
      const int ly = i / B, bx = i % B;
      const int lx = bx + rb;  // centre in raw coordinates
      float acc = gauss_k(ksize, rb) * raw[ly][lx];
      for (int j = 1; j <= rb; ++j) acc = acc + gauss_k(ksize, rb + j) * (raw[ly][lx - j] + raw[ly][lx + j]);
      tmp[ly][bx] = acc;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < B * B; i += kThreads) {
      const int by = i / B, bx = i % B;
      const int ly = by + rb;
      float acc = gauss_k(ksize, rb) * tmp[ly][bx];
      for (int j = 1; j <= rb; ++j) acc = acc + gauss_k(ksize, rb + j) * (tmp[ly - j][bx] + tmp[ly + j][bx]);
      sm[by][bx] = acc;
    }
  } else {
    for (int i = threadIdx.x; i < B * B; i += kThreads) {
      const int by = i / B, bx = i % B;
      sm[by][bx] = raw[by][bx];
    }
  }
  __syncthreads();
  const int tx = threadIdx.x % kT, ty = threadIdx.x / kT;
  const int x = x0 + tx, y = y0 + ty;
  inside = (x < cols) && (y < rows);
  // sm index (by, bx) = image (y0 - 1 + by, x0 - 1 + bx); the lane's pixel is (ty+1, tx+1).
  // gx: row derivative then column [1 2 1]; gy: row [1 2 1] then column derivative
  const int cy = ty + 1, cx = tx + 1;
  const float dxm = sm[cy - 1][cx + 1] - sm[cy - 1][cx - 1];
  const float dx0 = sm[cy][cx + 1] - sm[cy][cx - 1];
  const float dxp = sm[cy + 1][cx + 1] - sm[cy + 1][cx - 1];
  gx = (dxm + dxp) + dx0 * 2.0f;
  const float sym = (sm[cy - 1][cx - 1] + sm[cy - 1][cx + 1]) + sm[cy - 1][cx] * 2.0f;
  const float syp = (sm[cy + 1][cx - 1] + sm[cy + 1][cx + 1]) + sm[cy + 1][cx] * 2.0f;
  gy = syp - sym;
}

typedef const FixedDev __attribute__((address_space(4)))* ConstFixed;

__device__ inline __amdgpu_buffer_rsrc_t ro_rsrc(const void* p, int64_t bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ inline float ld_f32(__amdgpu_buffer_rsrc_t r, int byte_off)
{
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, byte_off, 0, 0));
}
__device__ inline double ld_f64(__amdgpu_buffer_rsrc_t r, int byte_off)
{
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0);
  return __hiloint2double((int)v[1], (int)v[0]);
}

// tile_gradients with the blur radius kRb (0: no blur) fixed at compile time: the window
// reads are issued together before any is used, and the LDS passes have constant trip counts.
// Same arithmetic, in the same order, as tile_gradients.
template <int kRb>
__device__ inline void tile_gradients_k(const float* __restrict__ img, int rows, int cols, int x0, int y0,
                                        float& gx, float& gy, bool& inside)
{
#pragma clang fp contract(off)
  constexpr int kHalo = kRb + 1;
  constexpr int kW = kT + 2 * kHalo;  // <= 24
  constexpr int kB = kT + 2;
  constexpr int kRows = (kW + 7) / 8;
  constexpr int ksize = 2 * kRb + 1;
  __shared__ float raw[kW][kW];
  __shared__ float tmp[kRb > 0 ? kW : 1][kB];
  __shared__ float smb[kRb > 0 ? kB : 1][kB];
  const int lx = threadIdx.x & 31, ly0 = threadIdx.x >> 5;
  const int xc = reflect101(x0 - kHalo + min(lx, kW - 1), cols);
  const auto rs = ro_rsrc(img, (int64_t)rows * cols * 4);
  float v[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int y = reflect101(y0 - kHalo + min(ly0 + 8 * r, kW - 1), rows);
    v[r] = ld_f32(rs, (y * cols + xc) * 4);
  }
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    if (lx < kW && ly0 + 8 * r < kW) raw[ly0 + 8 * r][lx] = v[r];
  }
  __syncthreads();
  float (*sm)[kB] = nullptr;
  if constexpr (kRb > 0) {
    for (int i = threadIdx.x; i < kW * kB; i += kThreads) {
      const int ly = i / kB, bx = i % kB;
      const int lxx = bx + kRb;
      float acc = gauss_k(ksize, kRb) * raw[ly][lxx];
#pragma unroll
      for (int j = 1; j <= kRb; ++j) acc = acc + gauss_k(ksize, kRb + j) * (raw[ly][lxx - j] + raw[ly][lxx + j]);
      tmp[ly][bx] = acc;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kB * kB; i += kThreads) {
      const int by = i / kB, bx = i % kB;
      const int ly = by + kRb;
      float acc = gauss_k(ksize, kRb) * tmp[ly][bx];
#pragma unroll
      for (int j = 1; j <= kRb; ++j) acc = acc + gauss_k(ksize, kRb + j) * (tmp[ly - j][bx] + tmp[ly + j][bx]);
      smb[by][bx] = acc;
    }
    __syncthreads();
    sm = smb;
  } else {
    sm = raw;  // kW == kB: the Sobel reads the raw window directly
  }
  const int tx = threadIdx.x % kT, ty = threadIdx.x / kT;
  inside = (x0 + tx < cols) && (y0 + ty < rows);
  const int cy = ty + 1, cx = tx + 1;
  const float dxm = sm[cy - 1][cx + 1] - sm[cy - 1][cx - 1];
  const float dx0 = sm[cy][cx + 1] - sm[cy][cx - 1];
  const float dxp = sm[cy + 1][cx + 1] - sm[cy + 1][cx - 1];
  gx = (dxm + dxp) + dx0 * 2.0f;
  const float sym = (sm[cy - 1][cx - 1] + sm[cy - 1][cx + 1]) + sm[cy - 1][cx] * 2.0f;
  const float syp = (sm[cy + 1][cx - 1] + sm[cy + 1][cx + 1]) + sm[cy + 1][cx] * 2.0f;
  gy = syp - sym;
}

// ---- kernels over SimItem tables: blockIdx.x -> tile (xcd_tile), blockIdx.y = item ----

// The tile of this workgroup among an image's T tiles (row-major), or -1 for a padding slot.
// The dispatcher deals workgroups to the 8 XCDs in turn, so with gridDim.x a multiple of 8 the
// workgroups with blockIdx.x = 8 j + k all land on XCD k; giving class k the k-th contiguous
// run of ceil(T / 8) tiles keeps neighbouring tiles, and the window rows they share, in one
// XCD's L2.
__device__ inline int xcd_tile(int T)
{
  const int per = (T + 7) >> 3;
  const int x = (int)blockIdx.x, k = x & 7, j = x >> 3;
  if (j >= per) return -1;
  const int t = k * per + j;
  return t < T ? t : -1;
}

// gridDim.x for up to `tiles` tiles per image under xcd_tile
inline unsigned xcd_grid_x(int tiles) { return (unsigned)(8 * ((tiles + 7) / 8)); }

// NCC sums over raw moving images: per tile (sum m, sum m^2, sum m f~) [masked]
__global__ __launch_bounds__(kThreads) void ncc_sums_kernel(const SimItem* __restrict__ items,
                                                            double* __restrict__ partials, int ptiles)
{
  const SimItem it = items[blockIdx.y];
  const FixedDev& fx = *it.fixed;
  const int TX = (fx.cols + kT - 1) / kT, TY = (fx.rows + kT - 1) / kT;
  const int tile = xcd_tile(TX * TY);
  if (tile < 0) return;
  const int tx0 = (tile % TX) * kT, ty0 = (tile / TX) * kT;
  const int x = tx0 + (threadIdx.x % kT), y = ty0 + (threadIdx.x / kT);
  double v[3] = {0.0, 0.0, 0.0};
  if (x < fx.cols && y < fx.rows) {
    const int64_t o = (int64_t)y * fx.cols + x;
    if (!fx.mask || fx.mask[o]) {
      const double m = (double)it.mov[o];
      v[0] = m;
      v[1] = m * m;
      v[2] = m * (double)fx.zf[o];
    }
  }
  block_reduce_store<3>(v, partials + (it.out_index * ptiles + tile) * 4);
}

// Grad-NCC sums: gradients in LDS, then sums for x and y directions
template <int kRb>
__global__ __launch_bounds__(kThreads) void grad_ncc_sums_kernel(const SimItem* __restrict__ items,
                                                                 double* __restrict__ partials, int ptiles)
{
  const SimItem it = items[blockIdx.y];
  const FixedDev& fx = *it.fixed;
  const int TX = (fx.cols + kT - 1) / kT, TY = (fx.rows + kT - 1) / kT;
  const int tile = xcd_tile(TX * TY);
  if (tile < 0) return;
  const int x0 = (tile % TX) * kT, y0 = (tile / TX) * kT;
  float gx, gy;
  bool inside;
  tile_gradients_k<kRb>(it.mov, fx.rows, fx.cols, x0, y0, gx, gy, inside);
  double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (inside) {
    const int64_t o = (int64_t)(y0 + threadIdx.x / kT) * fx.cols + x0 + (threadIdx.x % kT);
    if (!fx.mask || fx.mask[o]) {
      const double a = (double)gx, b = (double)gy;
      v[0] = a;
      v[1] = a * a;
      v[2] = a * (double)fx.zf[o];
      v[3] = b;
      v[4] = b * b;
      v[5] = b * (double)fx.zfy[o];
    }
  }
  block_reduce_store<6>(v, partials + (it.out_index * ptiles + tile) * 8);
}

// Gradient images of moving images (for the patch metric), written to it.gx / it.gy.
template <int kRb>
__global__ __launch_bounds__(kThreads) void grad_images_kernel(const SimItem* __restrict__ items)
{
  const SimItem it = items[blockIdx.y];
  const FixedDev& fx = *it.fixed;
  const int TX = (fx.cols + kT - 1) / kT, TY = (fx.rows + kT - 1) / kT;
  const int tile = xcd_tile(TX * TY);
  if (tile < 0) return;
  const int x0 = (tile % TX) * kT, y0 = (tile / TX) * kT;
  float gx, gy;
  bool inside;
  tile_gradients_k<kRb>(it.mov, fx.rows, fx.cols, x0, y0, gx, gy, inside);
  if (inside) {
    const int64_t o = (int64_t)(y0 + threadIdx.x / kT) * fx.cols + x0 + (threadIdx.x % kT);
    it.gx[o] = gx;
    it.gy[o] = gy;
  }
}

// Patch NCC over one (or two, for gradients) image(s): each lane owns one patch of a 16x16
// block of patch positions; the image region under the block is staged in LDS.
// LDS edge for a 16x16 block of patches: (16-1)*stride + diameter (26 for r=5, s=1)
__host__ __device__ inline int patch_lds_edge(int stride, int radius) { return (kT - 1) * stride + 2 * radius + 1; }

// Patch NCC of one moving image (ImgSimMetric2DPatchNCCCPU, xregImgSimMetric2DPatchNCCCPU.cpp:
// 212-253) from box sums per patch, in double: S1 = sum m, S2 = sum m^2 over the patch and,
// over the pixels U the correlation uses (all, or the unmasked ones), SfU = sum m f and
// S1U = sum m. mean = S1 / N, var = (S2 - mean S1) / (N - 1), sd = max(1e-6, sqrt(var)), and
// sum_U (m - mean)(f - mean_f) = SfU - mean_f S1U - mean G with G = sum_U f - |U| mean_f
// (fixed, per patch). The kernel adds 1 - cross / (sd * den) per patch, den = sd_f N / c with
// c the patch's coefficient in the image score (its weight, 1, or 0 for a skipped patch,
// :212, 252-253; den = +inf then); the finish adds sum c - P. Without a mask S1U = S1 and
// SfU = Sf. A 16 x 16 block of patches stages its moving and fixed
// (masked) window in LDS as doubles, forms the horizontal d-wide sums of every window row for
// its 16 patch columns, then each lane sums d of those vertically: about 5 d + 3 d operations
// per patch instead of 5 d^2.
template <bool kTwo>
__global__ __launch_bounds__(kThreads) void patch_kernel(const SimItem* __restrict__ items,
                                                         double* __restrict__ partials, int ptiles)
{
  extern __shared__ double patch_lds[];
  const SimItem it = items[blockIdx.y];
  const FixedDev& fx = *it.fixed;
  const int s = fx.patch_stride, r = fx.patch_radius, d = 2 * r + 1;
  const int ppr = fx.ppr, ppc = fx.ppc;
  const int TX = (ppr + kT - 1) / kT, TY = (ppc + kT - 1) / kT;
  const int tile = xcd_tile(TX * TY);
  if (tile < 0) return;
  const int px0 = (tile % TX) * kT, py0 = (tile / TX) * kT;  // first patch index
  const int X0 = px0 * s, Y0 = py0 * s;                                  // first image pixel
  const int W = patch_lds_edge(s, r);
  const bool masked = fx.mask != nullptr;
  const bool mstats = fx.pcount != nullptr;  // statistics over the unmasked pixels (needs a mask)
  double* lm = patch_lds;
  double* lf = lm + W * W;
  double* h1 = lf + W * W;
  double* h2 = h1 + W * kT;
  double* hf = h2 + W * kT;
  double* lk = hf + W * kT;  // masked only: mask window, then sum_U m (and sum_U m^2)
  double* hu = lk + W * W;
  double* h2u = hu + W * kT;
  const int pxl = threadIdx.x % kT, pyl = threadIdx.x / kT;
  const int pi = px0 + pxl, pj = py0 + pyl;
  const bool valid = (pi < ppr) && (pj < ppc);
  const double N = (double)(d * d);
  double acc[2] = {0.0, 0.0};
  for (int dir = 0; dir < (kTwo ? 2 : 1); ++dir) {
    const float* mimg = (dir == 0) ? (kTwo ? it.gx : it.mov) : it.gy;
    const float* fimg = (dir == 0) ? fx.pf : fx.pfy;
    const float* pmu = (dir == 0) ? fx.pmean : fx.pmeany;
    const float* pden = (dir == 0) ? fx.pden : fx.pdeny;
    const double* pg = (dir == 0) ? fx.pg : fx.pgy;
    __syncthreads();
    for (int i = threadIdx.x; i < W * W; i += kThreads) {
      const int ly = i / W, lx = i % W;
      const int y = min(Y0 + ly, fx.rows - 1), x = min(X0 + lx, fx.cols - 1);
      const int64_t o = (int64_t)y * fx.cols + x;
      const double k = masked ? (fx.mask[o] ? 1.0 : 0.0) : 1.0;
      lm[i] = (double)mimg[o];
      lf[i] = (double)fimg[o] * k;
      if (masked) lk[i] = k;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < W * kT; e += kThreads) {
      const int row = e / kT, x0 = (e % kT) * s;
      const double* m = lm + row * W + x0;
      const double* f = lf + row * W + x0;
      double a1 = 0.0, a2 = 0.0, af = 0.0;
      for (int xx = 0; xx < d; ++xx) {
        a1 += m[xx];
        a2 = fma(m[xx], m[xx], a2);
        af = fma(m[xx], f[xx], af);
      }
      h1[e] = a1;
      h2[e] = a2;
      hf[e] = af;
      if (masked) {
        const double* k = lk + row * W + x0;
        double au = 0.0, a2u = 0.0;
        for (int xx = 0; xx < d; ++xx) au = fma(m[xx], k[xx], au);
        hu[e] = au;
        if (mstats) {
          for (int xx = 0; xx < d; ++xx) a2u = fma(m[xx] * m[xx], k[xx], a2u);
          h2u[e] = a2u;
        }
      }
    }
    __syncthreads();
    if (valid) {
      double S1 = 0.0, S2 = 0.0, Sf = 0.0, SU = 0.0, S2U = 0.0;
      for (int yy = 0; yy < d; ++yy) {
        const int e = (pyl * s + yy) * kT + pxl;
        S1 += h1[e];
        S2 += h2[e];
        Sf += hf[e];
        if (masked) SU += hu[e];
        if (mstats) S2U += h2u[e];
      }
      if (!masked) SU = S1;
      const int64_t p = (int64_t)pj * ppr + pi;
      double mean, var;
      if (mstats) {
        // ComputePatchMeanStdDev over the n unmasked pixels: the sum is divided only when n > 1
        const double n = (double)fx.pcount[p];
        mean = (n > 1.0) ? SU / n : SU;
        var = (n > 1.0) ? fmax(S2U - mean * SU, 0.0) / (n - 1.0) : 0.0;
      } else {
        mean = S1 / N;
        var = fmax(S2 - mean * S1, 0.0) / (N - 1.0);
      }
      const double sd = fmax((double)1.0e-6f, sqrt(var));
      const double cross = (Sf - (double)pmu[p] * SU) - mean * pg[p];
      // a patch with no unmasked pixel correlates over nothing: its score is exactly 1
      const bool empty = mstats && fx.pcount[p] == 0.0f;
      acc[dir] = empty ? 1.0 : 1.0 - (cross / sd) / (double)pden[p];
      if (it.patch_scores) {
        // the unweighted patch score (ImgSimMetric2DPatchNCCCPU's sim_vals_for_each_patch_, :247-250);
        // Patch-Grad-NCC keeps 0.5 (x + y) of the float scores (xregImgSimMetric2DPatchGradNCCCPU.cpp:226-240)
        const float* unit = (dir == 0) ? fx.pden_unit : fx.pdeny_unit;
        const float sp = empty ? 1.0f : (float)(1.0 - (cross / sd) / (double)unit[p]);
        if (dir == 0) {
          it.patch_scores[p] = sp;
        } else {
          it.patch_scores[p] = (float)(0.5 * (double)(it.patch_scores[p] + sp));
        }
      }
    }
  }
  block_reduce_store<2>(acc, partials + (it.out_index * ptiles + tile) * 2);
}

// patch_kernel for the common case -- no mask, stride 1, a compile-time patch diameter kD.
// Same per-patch closed form; the box sums are formed with sliding windows so that LDS
// traffic drops about threefold:
//  * each direction's window and per-patch fixed statistics are read before the first one is
//    waited for; the second direction's window is read once the first direction's sums are
//    done, so that its registers are free during them;
//  * the window is staged in LDS as float (the values are floats; products are formed in
//    double);
//  * horizontal pass: 4 lanes per window row, each forming 4 consecutive d-wide sums -- the
//    first directly, the next three by adding the entering and removing the leaving column;
//  * vertical pass: 64 lanes, one per patch column and group of 4 patch rows, slide the same
//    way down the column and evaluate the closed form for their 4 patches. The row sums are
//    kept at an odd pitch (17 doubles) so that the horizontal pass's stores spread over the
//    LDS banks (2-way instead of 8-way conflicts at a pitch of 16).
// Occupancy: the two-direction form runs at 6 waves per SIMD (78 VGPRs; with the second window
// read during the first direction's sums it needed 91, 5 waves, and measured 3% slower), the
// one-direction form at 8 (64 VGPRs).
// The sums are FP64 throughout; the sliding updates differ from direct sums by a few units in
// the last place of a double.
template <bool kTwo, int kD>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kTwo ? 6 : 8)))
void patch_kernel_d(const SimItem* __restrict__ items, double* __restrict__ partials, int ptiles)
{
  constexpr int kW = (kT - 1) + kD;  // window edge at stride 1
  constexpr int kRows = (kW + 7) / 8;
  constexpr int kDirs = kTwo ? 2 : 1;
  constexpr int kSeg = 4;            // sums per lane in the sliding passes
  constexpr int kSegs = kT / kSeg;   // lanes per window row / per patch column
  constexpr int kHP = kT + 1;        // row pitch of the double row sums (odd: spreads the banks)
  static_assert(kW <= 32, "one 32-lane column pass per window row");
  static_assert(kW * kSegs <= kThreads && kT * kSegs <= kThreads, "sliding lanes fit the block");
  extern __shared__ double patch_lds[];
  const SimItem it = items[blockIdx.y];
  const ConstFixed fx = (ConstFixed)it.fixed;  // uniform: scalar loads
  const int rows = fx->rows, cols = fx->cols, ppr = fx->ppr, ppc = fx->ppc;
  const int TX = (ppr + kT - 1) / kT, TY = (ppc + kT - 1) / kT;
  const int tile = xcd_tile(TX * TY);
  if (tile < 0) return;
  const int X0 = (tile % TX) * kT, Y0 = (tile / TX) * kT;  // first patch = first pixel
  float* lm = (float*)patch_lds;
  float* lf = lm + kW * kW;
  double* h1 = (double*)(lf + kW * kW);  // 2 kW^2 floats: a multiple of 8 bytes
  double* h2 = h1 + kW * kHP;
  double* hf = h2 + kW * kHP;
  constexpr double kN = (double)(kD * kD);
  constexpr double kInvN = 1.0 / kN, kInvN1 = 1.0 / (kN - 1.0);
  const int lx = threadIdx.x & 31, ly0 = threadIdx.x >> 5;
  const bool col_ok = lx < kW;
  // lanes outside the window read a clamped in-image pixel and do not store it
  const int xo = min(X0 + min(lx, kW - 1), cols - 1);
  const int64_t img_bytes = (int64_t)rows * cols * 4;
  const int64_t np = (int64_t)ppr * ppc;
  // vertical-pass lanes: patch column vc, patch rows vs * kSeg ... + kSeg - 1
  const bool vlane = threadIdx.x < kT * kSegs;
  const int vc = threadIdx.x & (kT - 1), vs = (threadIdx.x >> 4) & (kSegs - 1);

  // the byte offset of window row r staged by this lane (formed at each read)
  auto wo = [&](int r) { return (min(Y0 + min(ly0 + 8 * r, kW - 1), rows - 1) * cols + xo) * 4; };
  float vm[kRows], vf[kRows];
  auto load_window = [&](int dir) {
    const auto rm = ro_rsrc((dir == 0) ? (kTwo ? it.gx : it.mov) : it.gy, img_bytes);
    const auto rf = ro_rsrc((dir == 0) ? fx->pf : fx->pfy, img_bytes);
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int o = wo(r);
      vm[r] = ld_f32(rm, o);
      vf[r] = ld_f32(rf, o);
    }
  };
  load_window(0);
  float mu[kSeg], den[kSeg];
  double g[kSeg];
  bool pv[kSeg];
#pragma unroll
  for (int k = 0; k < kSeg; ++k) pv[k] = vlane && (X0 + vc < ppr) && (Y0 + vs * kSeg + k < ppc);
  // one direction's per-patch fixed statistics, read when that direction starts (its window
  // staging and horizontal pass cover the reads)
  auto load_stats = [&](int dir) {
    const auto rmu = ro_rsrc((dir == 0) ? fx->pmean : fx->pmeany, np * 4);
    const auto rden = ro_rsrc((dir == 0) ? fx->pden : fx->pdeny, np * 4);
    const auto rg = ro_rsrc((dir == 0) ? fx->pg : fx->pgy, np * 8);
#pragma unroll
    for (int k = 0; k < kSeg; ++k) {
      const int p = pv[k] ? (Y0 + vs * kSeg + k) * ppr + X0 + vc : 0;  // invalid: patch 0, discarded
      mu[k] = ld_f32(rmu, p * 4);
      den[k] = ld_f32(rden, p * 4);
      g[k] = ld_f64(rg, p * 8);
    }
  };

  double acc[2] = {0.0, 0.0};
#pragma unroll
  for (int dir = 0; dir < kDirs; ++dir) {
    load_stats(dir);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int ly = ly0 + 8 * r;
      if (col_ok && ly < kW) {
        lm[ly * kW + lx] = vm[r];
        lf[ly * kW + lx] = vf[r];
      }
    }
    __syncthreads();
    if (threadIdx.x < kW * kSegs) {
      const int row = threadIdx.x / kSegs, c0 = (threadIdx.x % kSegs) * kSeg;
      const float* m = lm + row * kW + c0;
      const float* f = lf + row * kW + c0;
      double* o1 = h1 + row * kHP + c0;
      double* o2 = h2 + row * kHP + c0;
      double* of = hf + row * kHP + c0;
      double a1 = 0.0, a2 = 0.0, af = 0.0;
#pragma unroll
      for (int xx = 0; xx < kD; ++xx) {
        const double dm = (double)m[xx];
        a1 += dm;
        a2 = fma(dm, dm, a2);
        af = fma(dm, (double)f[xx], af);
      }
      o1[0] = a1;
      o2[0] = a2;
      of[0] = af;
#pragma unroll
      for (int k = 1; k < kSeg; ++k) {
        const double mo = (double)m[k - 1], mn = (double)m[k - 1 + kD];
        a1 = (a1 - mo) + mn;
        a2 = fma(mn, mn, fma(-mo, mo, a2));
        af = fma(mn, (double)f[k - 1 + kD], fma(-mo, (double)f[k - 1], af));
        o1[k] = a1;
        o2[k] = a2;
        of[k] = af;
      }
    }
    __syncthreads();
    if (vlane) {
      const int e0 = vs * kSeg * kHP + vc;
      double S1 = 0.0, S2 = 0.0, Sf = 0.0;
#pragma unroll
      for (int yy = 0; yy < kD; ++yy) {
        S1 += h1[e0 + yy * kHP];
        S2 += h2[e0 + yy * kHP];
        Sf += hf[e0 + yy * kHP];
      }
#pragma unroll
      for (int k = 0; k < kSeg; ++k) {
        if (k > 0) {
          const int eo = e0 + (k - 1) * kHP, en = e0 + (k - 1 + kD) * kHP;
          S1 = (S1 - h1[eo]) + h1[en];
          S2 = (S2 - h2[eo]) + h2[en];
          Sf = (Sf - hf[eo]) + hf[en];
        }
        if (pv[k]) {
          const double mean = S1 * kInvN;
          const double var = fmax(S2 - mean * S1, 0.0) * kInvN1;
          const double sd = fmax((double)1.0e-6f, sqrt(var));
          const double cross = (Sf - (double)mu[k] * S1) - mean * g[k];
          acc[dir] += 1.0 - cross / (sd * (double)den[k]);
        }
      }
    }
    if (dir + 1 < kDirs) load_window(dir + 1);
  }
  block_reduce_store<2>(acc, partials + (it.out_index * ptiles + tile) * 2);
}

// fixed-image per-patch statistics (reference float order, ComputePatchMeanStdDev): over the
// whole patch, or over its unmasked pixels when count_out is given (use_mask_for_patch_stats)
__global__ void patch_stats_kernel(const float* __restrict__ img, const uint8_t* __restrict__ mask, int cols,
                                   int ppr, int ppc, int s, int d, float* __restrict__ mean_out,
                                   float* __restrict__ den_out, double* __restrict__ g_out,
                                   float* __restrict__ count_out)
{
#pragma clang fp contract(off)
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)ppr * ppc) return;
  const int r0 = (int)(p / ppr) * s, c0 = (int)(p % ppr) * s;
  const bool mstats = count_out != nullptr;
  float mean = 0.0f;
  int n = 0;
  for (int y = 0; y < d; ++y)
    for (int x = 0; x < d; ++x) {
      const int64_t o = (int64_t)(r0 + y) * cols + c0 + x;
      if (!mstats || mask[o]) {
        mean += img[o];
        ++n;
      }
    }
  float var = 0.0f;
  if (n > 1) {
    mean /= (float)n;
    for (int y = 0; y < d; ++y)
      for (int x = 0; x < d; ++x) {
        const int64_t o = (int64_t)(r0 + y) * cols + c0 + x;
        if (!mstats || mask[o]) {
          const float t = img[o] - mean;
          var += t * t;
        }
      }
    var /= (float)n - 1.0f;
  }
  const float sd = fmaxf(1.0e-6f, sqrtf(var));
  mean_out[p] = mean;
  den_out[p] = sd * (float)n;
  if (mstats) count_out[p] = (float)n;
  // over the pixels the correlation uses
  double f1 = 0.0, nu = 0.0;
  for (int y = 0; y < d; ++y)
    for (int x = 0; x < d; ++x) {
      const int64_t o = (int64_t)(r0 + y) * cols + c0 + x;
      if (!mask || mask[o]) {
        f1 += (double)img[o];
        nu += 1.0;
      }
    }
  g_out[p] = f1 - nu * (double)mean;
}

// fixed gradient images (one image)
__global__ __launch_bounds__(kThreads) void fixed_grad_kernel(const float* __restrict__ img, int rows,
                                                              int cols, int ksize, float* __restrict__ gx_out,
                                                              float* __restrict__ gy_out)
{
  const int TX = (cols + kT - 1) / kT;
  const int x0 = (blockIdx.x % TX) * kT, y0 = (blockIdx.x / TX) * kT;
  float gx, gy;
  bool inside;
  tile_gradients(img, rows, cols, x0, y0, ksize, gx, gy, inside);
  if (inside) {
    const int64_t o = (int64_t)(y0 + threadIdx.x / kT) * cols + x0 + (threadIdx.x % kT);
    gx_out[o] = gx;
    gy_out[o] = gy;
  }
}

// one image: sum and (after mean is known) sum of squared deviations, FP64, single block
__global__ __launch_bounds__(1024) void image_stats_kernel(const float* __restrict__ img, const uint8_t* mask,
                                                           int64_t n, double* __restrict__ out)
{
  __shared__ double red[16];
  __shared__ double redc[16];
  __shared__ double mean_sh;
  double s = 0.0, c = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    if (!mask || mask[i]) {
      s += (double)img[i];
      c += 1.0;
    }
  }
  s = wave_sum(s);
  c = wave_sum(c);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w] = s;
    redc[w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0, tc = 0.0;
    for (int i = 0; i < nw; ++i) {
      t += red[i];
      tc += redc[i];
    }
    out[0] = t;
    out[1] = tc;
    mean_sh = t / tc;
  }
  __syncthreads();
  double v = 0.0;
  const double mu = mean_sh;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    if (!mask || mask[i]) {
      const double t = (double)img[i] - mu;
      v += t * t;
    }
  }
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t += red[i];
    out[2] = t;
  }
}

__global__ void sub_scalar_kernel(const float* __restrict__ src, float mean, int64_t n, float* __restrict__ dst)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i] - mean;
}

__global__ __launch_bounds__(1024) void sum_kernel(const float* __restrict__ img, const uint8_t* mask, int64_t n,
                                                   double* __restrict__ out)
{
  __shared__ double red[16];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    if (!mask || mask[i]) s += (double)img[i];
  }
  s = wave_sum(s);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < nw; ++i) t += red[i];
    out[0] = t;
  }
}

// NCC closed form from FP64 sums (one direction)
__device__ inline double ncc_from_sums(double s1, double s2, double sf, double n, double sum_zf,
                                       double std_f)
{
  const double mu = s1 / n;
  double var = (s2 - s1 * mu) / (n - 1.0);
  if (var < 0.0) var = 0.0;
  const double sd = fmax(1.0e-6, sqrt(var));
  // sum f~ (m - mu) = sum f~ m - mu sum f~
  const double num = sf - mu * sum_zf;
  const double ncc = num / (n * std_f * sd);
  return (1.0 - ncc) * 0.5;
}

// finish: one wave per item sums the tile records in a fixed order
__global__ void finish_kernel(const SimItem* __restrict__ items, const double* __restrict__ partials,
                              int ptiles, int stride_rec, float* __restrict__ sim_out, int64_t n_items)
{
  const int64_t item = blockIdx.x;
  if (item >= n_items) return;
  const SimItem it = items[item];
  const FixedDev& fx = *it.fixed;
  int ntiles;
  if (fx.metric == XRH_SIM_PATCH_NCC || fx.metric == XRH_SIM_PATCH_GRAD_NCC) {
    ntiles = ((fx.ppr + kT - 1) / kT) * ((fx.ppc + kT - 1) / kT);
  } else {
    ntiles = ((fx.cols + kT - 1) / kT) * ((fx.rows + kT - 1) / kT);
  }
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const double* base = partials + it.out_index * ptiles * stride_rec;
  for (int t = threadIdx.x; t < ntiles; t += 64) {
    for (int k = 0; k < 6 && k < stride_rec; ++k) acc[k] += base[(int64_t)t * stride_rec + k];
  }
  for (int k = 0; k < 6; ++k) acc[k] = wave_sum(acc[k]);
  if (threadIdx.x != 0) return;
  double sim = 0.0;
  const double n = (double)fx.count;
  switch (fx.metric) {
    case XRH_SIM_NCC:
      sim = ncc_from_sums(acc[0], acc[1], acc[2], n, fx.sum_zf, fx.std_f);
      break;
    case XRH_SIM_GRAD_NCC: {
      const float sx = (float)ncc_from_sums(acc[0], acc[1], acc[2], n, fx.sum_zf, fx.std_f);
      const float sy = (float)ncc_from_sums(acc[3], acc[4], acc[5], n, fx.sum_zfy, fx.std_fy);
      sim = 0.5 * (double)(sx + sy);
      break;
    }
    case XRH_SIM_PATCH_NCC:
      sim = (acc[0] + fx.patch_offset[0]) / fx.patch_norm[0];
      break;
    default: {
      const float sx = (float)((acc[0] + fx.patch_offset[0]) / fx.patch_norm[0]);
      const float sy = (float)((acc[1] + fx.patch_offset[1]) / fx.patch_norm[1]);
      sim = 0.5 * (double)(sx + sy);
      break;
    }
  }
  sim_out[it.out_index] = (float)sim;
}

}  // namespace

int sim_record_stride(int metric)
{
  switch (metric) {
    case XRH_SIM_NCC: return 4;
    case XRH_SIM_GRAD_NCC: return 8;
    default: return 2;
  }
}

int sim_img_tiles(const FixedDev& f) { return ((f.cols + kT - 1) / kT) * ((f.rows + kT - 1) / kT); }

int sim_group_key(const FixedDev& f)
{
  int d = 0;
  // (the specialised kernel addresses images and patch tables with 32-bit byte offsets)
  const bool small = (int64_t)f.rows * f.cols * 4 < (int64_t)INT32_MAX && f.num_patches * 8 < (int64_t)INT32_MAX;
  if ((f.metric == XRH_SIM_PATCH_NCC || f.metric == XRH_SIM_PATCH_GRAD_NCC) && !f.mask && f.patch_stride == 1 &&
      small) {
    const int dd = 2 * f.patch_radius + 1;
    if (dd == 7 || dd == 9 || dd == 11 || dd == 13 || dd == 15) d = dd;
  }
  const int rb = (f.metric == XRH_SIM_GRAD_NCC || f.metric == XRH_SIM_PATCH_GRAD_NCC) && f.smooth_ksize > 1
                     ? f.smooth_ksize / 2 : 0;
  return f.metric | (d << kSimKeyShift) | (rb << kSimKeyBlurShift);
}

#define XRH_BLUR_DISPATCH(rb, KERNEL, ...)                                            \
  switch (rb) {                                                                       \
    case 0: hipLaunchKernelGGL(KERNEL<0>, __VA_ARGS__); break;                        \
    case 1: hipLaunchKernelGGL(KERNEL<1>, __VA_ARGS__); break;                        \
    case 2: hipLaunchKernelGGL(KERNEL<2>, __VA_ARGS__); break;                        \
    case 3: hipLaunchKernelGGL(KERNEL<3>, __VA_ARGS__); break;                        \
    default: fail(XRH_ERR_INVALID, "blur radius out of range");                       \
  }

int sim_patch_lds_bytes(const FixedDev& f, bool general)
{
  if (f.metric != XRH_SIM_PATCH_NCC && f.metric != XRH_SIM_PATCH_GRAD_NCC) return 0;
  const int W = patch_lds_edge(f.patch_stride, f.patch_radius);
  if (!general && ((sim_group_key(f) >> kSimKeyShift) & ((1 << (kSimKeyBlurShift - kSimKeyShift)) - 1))) {
    // patch_kernel_d: float window pair + three double row-sum arrays
    return 2 * W * W * (int)sizeof(float) + 3 * W * (kT + 1) * (int)sizeof(double);
  }
  return (f.mask ? (3 * W * W + (f.pcount ? 5 : 4) * W * kT) : (2 * W * W + 3 * W * kT)) * (int)sizeof(double);
}

int sim_tiles(const FixedDev& f)
{
  if (f.metric == XRH_SIM_PATCH_NCC || f.metric == XRH_SIM_PATCH_GRAD_NCC) {
    return ((f.ppr + kT - 1) / kT) * ((f.ppc + kT - 1) / kT);
  }
  return sim_img_tiles(f);
}

template <bool kTwo>
static void launch_patch_d(int d, dim3 grid, int lds, hipStream_t s, const SimItem* items, double* partials,
                           int ptiles)
{
  switch (d) {
    case 7: hipLaunchKernelGGL((patch_kernel_d<kTwo, 7>), grid, dim3(kThreads), lds, s, items, partials, ptiles); break;
    case 9: hipLaunchKernelGGL((patch_kernel_d<kTwo, 9>), grid, dim3(kThreads), lds, s, items, partials, ptiles); break;
    case 11: hipLaunchKernelGGL((patch_kernel_d<kTwo, 11>), grid, dim3(kThreads), lds, s, items, partials, ptiles); break;
    case 13: hipLaunchKernelGGL((patch_kernel_d<kTwo, 13>), grid, dim3(kThreads), lds, s, items, partials, ptiles); break;
    case 15: hipLaunchKernelGGL((patch_kernel_d<kTwo, 15>), grid, dim3(kThreads), lds, s, items, partials, ptiles); break;
    default: fail(XRH_ERR_INVALID, "no specialised patch kernel for this diameter");
  }
}

void launch_sim(int key, const SimItem* items_dev, int64_t n_items, int max_img_tiles,
                int max_patch_tiles, double* partials, float* sim_out_dev, bool fused_ncc_done, int patch_lds_bytes,
                hipStream_t s)
{
  if (n_items <= 0) return;
  const int metric = key & ((1 << kSimKeyShift) - 1);
  // compile-time patch diameter (0: the general patch kernel) and blur radius of the group
  const int dfix = (key >> kSimKeyShift) & ((1 << (kSimKeyBlurShift - kSimKeyShift)) - 1);
  const int rb = key >> kSimKeyBlurShift;
  XRH_REQUIRE(n_items < 65536, "too many moving images in one similarity launch");
  const bool patch = (metric == XRH_SIM_PATCH_NCC || metric == XRH_SIM_PATCH_GRAD_NCC);
  const int ptiles = patch ? max_patch_tiles : max_img_tiles;
  const dim3 grid(xcd_grid_x(ptiles), (unsigned)n_items);
  const int rec = sim_record_stride(metric);
  switch (metric) {
    case XRH_SIM_NCC:
      if (!fused_ncc_done) hipLaunchKernelGGL(ncc_sums_kernel, grid, dim3(kThreads), 0, s, items_dev, partials, ptiles);
      break;
    case XRH_SIM_GRAD_NCC:
      XRH_BLUR_DISPATCH(rb, grad_ncc_sums_kernel, grid, dim3(kThreads), 0, s, items_dev, partials, ptiles);
      break;
    case XRH_SIM_PATCH_NCC:
      if (dfix) {
        launch_patch_d<false>(dfix, grid, patch_lds_bytes, s, items_dev, partials, ptiles);
      } else {
        hipLaunchKernelGGL(patch_kernel<false>, grid, dim3(kThreads), patch_lds_bytes, s, items_dev, partials, ptiles);
      }
      break;
    case XRH_SIM_PATCH_GRAD_NCC:
      XRH_BLUR_DISPATCH(rb, grad_images_kernel, dim3(xcd_grid_x(max_img_tiles), (unsigned)n_items), dim3(kThreads), 0,
                        s, items_dev);
      if (dfix) {
        launch_patch_d<true>(dfix, grid, patch_lds_bytes, s, items_dev, partials, ptiles);
      } else {
        hipLaunchKernelGGL(patch_kernel<true>, grid, dim3(kThreads), patch_lds_bytes, s, items_dev, partials, ptiles);
      }
      break;
    default:
      fail(XRH_ERR_INVALID, "unknown similarity metric");
  }
  XRH_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(finish_kernel, dim3((unsigned)n_items), dim3(64), 0, s, items_dev, partials, ptiles, rec,
                     sim_out_dev, n_items);
  XRH_HIP_CHECK(hipGetLastError());
}

// ---- fixed-image processing (ImgSimMetric2D*::allocate_resources / process_mask) ----

static float* fixed_buf(xrh_fixed* f, uint64_t n)
{
  void* p = ctx_alloc(f->ctx, n * sizeof(float));
  f->buffers.push_back(p);
  return (float*)p;
}

// zero-mean copy + FP64 statistics of one fixed image (ComputeZeroMeanImageAndStats,
// xregImgSimMetric2DNCCCPU.cpp:66-75, 211-236)
static void ncc_prepare(xrh_fixed* f, const float* img, int64_t n, const uint8_t* mask_dev, float** zf_out,
                        double* sum_out, double* std_out, int64_t* count_out)
{
  hipStream_t s = f->ctx->stream;
  double* st = (double*)ctx_scratch(f->ctx, 64);
  hipLaunchKernelGGL(image_stats_kernel, dim3(1), dim3(1024), 0, s, img, mask_dev, n, st);
  XRH_HIP_CHECK(hipGetLastError());
  double h[3];
  XRH_HIP_CHECK(hipMemcpyAsync(h, st, sizeof(h), hipMemcpyDeviceToHost, s));
  XRH_HIP_CHECK(hipStreamSynchronize(s));
  const double cnt = h[1];
  XRH_REQUIRE(cnt >= 2.0, "fixed image (or mask) needs at least two pixels");
  const float mean = (float)(h[0] / cnt);
  double sd = std::sqrt(h[2] / (cnt - 1.0));
  if (sd < 1.0e-6) sd = 1.0e-6;
  float* zf = fixed_buf(f, (uint64_t)n);
  hipLaunchKernelGGL(sub_scalar_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, img, mean, n, zf);
  XRH_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(1024), 0, s, zf, mask_dev, n, st);
  XRH_HIP_CHECK(hipGetLastError());
  double z;
  XRH_HIP_CHECK(hipMemcpyAsync(&z, st, sizeof(z), hipMemcpyDeviceToHost, s));
  XRH_HIP_CHECK(hipStreamSynchronize(s));
  *zf_out = zf;
  *sum_out = z;
  *std_out = (double)(float)sd;
  *count_out = (int64_t)cnt;
}

static void patch_prepare(xrh_fixed* f, const float* img, float** mean_out, float** den_out, double** g_out,
                          float* count)
{
  FixedDev& h = f->host;
  float* mean = fixed_buf(f, (uint64_t)h.num_patches);
  float* den = fixed_buf(f, (uint64_t)h.num_patches);
  double* g = (double*)fixed_buf(f, (uint64_t)h.num_patches * 2);
  const int d = 2 * h.patch_radius + 1;
  hipLaunchKernelGGL(patch_stats_kernel, dim3((unsigned)((h.num_patches + 255) / 256)), dim3(256), 0,
                     f->ctx->stream, img, h.mask, h.cols, h.ppr, h.ppc, h.patch_stride, d, mean, den, g, count);
  XRH_HIP_CHECK(hipGetLastError());
  *mean_out = mean;
  *den_out = den;
  *g_out = g;
}

// Coefficient of each patch's score in the image score and the divisor of their sum, after
// ImgSimMetric2DPatchNCCCPU::compute (xregImgSimMetric2DPatchNCCCPU.cpp:186-287): every use of
// patch p in the index list adds (weight_in_combine ? w_p : 1) to c_p, except that with
// weight_in_combine a patch with |w_p| <= 1e-6 is skipped (:212); the sum is divided by the
// number of uses (compute_mean), else by the float sum of the used patches' weights
// (weight_in_combine), else by 1. The coefficients are folded into the per-patch denominators
// (den / c_p, +inf for c_p = 0) and the offset sum c - P, so the patch kernels are the same for
// every weighting. Patch-Grad-NCC's y direction takes its own weights when `wy` is given.
// `upload` = false only at creation, before the device copy exists.
static void patch_coefficients(xrh_fixed* f, const float* w, const float* wy, const int64_t* inds, int64_t n_inds,
                               bool wic, bool mean, bool upload)
{
  FixedDev& h = f->host;
  const int64_t np = h.num_patches;
  const bool grad = h.metric == XRH_SIM_PATCH_GRAD_NCC;
  XRH_REQUIRE(!wy || grad, "separate y-direction weights need Patch-Grad-NCC");
  XRH_REQUIRE(!wy || w, "y-direction weights need x-direction weights");
  std::vector<float> wt[2];
  wt[0] = w ? std::vector<float>(w, w + np) : f->create_weights;
  wt[1] = wy ? std::vector<float>(wy, wy + np) : wt[0];
  for (int k = 0; k < 2; ++k)
    for (float v : wt[k]) XRH_REQUIRE(std::isfinite(v), "patch weights must be finite");
  const int64_t nu = inds ? n_inds : np;
  XRH_REQUIRE(nu >= 1, "need at least one patch index");
  for (int64_t k = 0; k < nu; ++k) {
    const int64_t p = inds ? inds[k] : k;
    XRH_REQUIRE(p >= 0 && p < np, "patch index out of range");
  }
  const int dirs = grad ? 2 : 1;
  std::vector<double> coef[2];
  double norm[2] = {1.0, 1.0};
  bool unit[2] = {true, true};
  for (int k = 0; k < dirs; ++k) {
    coef[k].assign((size_t)np, 0.0);
    float wsum = 0.0f;
    for (int64_t q = 0; q < nu; ++q) {
      const int64_t p = inds ? inds[q] : q;
      const float wp = wt[k][(size_t)p];
      wsum += wp;
      if (!wic) {
        coef[k][(size_t)p] += 1.0;
      } else if (std::fabs(wp) > 1.0e-6f) {
        coef[k][(size_t)p] += (double)wp;
      }
    }
    if (mean) {
      norm[k] = (double)(float)nu;
    } else if (wic) {
      XRH_REQUIRE(wsum != 0.0f, "the weights of the patches used sum to zero");
      norm[k] = (double)wsum;
    }
    unit[k] = std::all_of(coef[k].begin(), coef[k].end(), [](double c) { return c == 1.0; });
  }
  hipStream_t s = f->ctx->stream;
  if (upload) XRH_HIP_CHECK(hipStreamSynchronize(s));  // work queued earlier reads the old tables
  const float* den[2] = {f->base_den_dev[0], f->base_den_dev[1]};
  std::vector<float> scaled((size_t)np);
  for (int k = 0; k < dirs; ++k) {
    double csum = 0.0;
    for (double c : coef[k]) csum += c;
    h.patch_offset[k] = unit[k] ? 0.0 : csum - (double)np;
    h.patch_norm[k] = norm[k];
    if (unit[k]) continue;
    if (!f->scaled_den_dev[k]) f->scaled_den_dev[k] = fixed_buf(f, (uint64_t)np);
    for (int64_t p = 0; p < np; ++p) {
      const double c = coef[k][(size_t)p];
      scaled[(size_t)p] = (c == 0.0) ? INFINITY : (float)((double)f->base_den[k][(size_t)p] / c);
    }
    XRH_HIP_CHECK(hipMemcpyAsync(f->scaled_den_dev[k], scaled.data(), sizeof(float) * np, hipMemcpyHostToDevice, s));
    XRH_HIP_CHECK(hipStreamSynchronize(s));
    den[k] = f->scaled_den_dev[k];
  }
  if (!grad) {
    h.patch_offset[1] = h.patch_offset[0];
    h.patch_norm[1] = h.patch_norm[0];
  }
  h.pden = den[0];
  if (grad) h.pdeny = den[1];
  f->weights = std::move(wt[0]);
  f->weights_y = std::move(wt[1]);
  if (upload) {
    XRH_HIP_CHECK(hipMemcpyAsync(f->dev, &h, sizeof(FixedDev), hipMemcpyHostToDevice, s));
    XRH_HIP_CHECK(hipStreamSynchronize(s));
  }
}

static void fixed_create(xrh_ctx* ctx, const float* fixed_host, int32_t rows, int32_t cols,
                         const xrh_sim_params* p, const uint8_t* mask_host, xrh_fixed** out)
{
  XRH_REQUIRE(ctx && fixed_host && p && out, "null argument");
  XRH_REQUIRE(rows >= 2 && cols >= 2, "fixed image must be at least 2x2");
  // the gradient and patch kernels address an image with 32-bit byte offsets
  XRH_REQUIRE((int64_t)rows * cols * 4 < (int64_t)INT32_MAX, "fixed image has 2^29 pixels or more");
  XRH_REQUIRE(p->metric >= XRH_SIM_NCC && p->metric <= XRH_SIM_PATCH_GRAD_NCC, "unknown metric");
  const bool grad = (p->metric == XRH_SIM_GRAD_NCC || p->metric == XRH_SIM_PATCH_GRAD_NCC);
  const bool patch = (p->metric == XRH_SIM_PATCH_NCC || p->metric == XRH_SIM_PATCH_GRAD_NCC);
  if (grad) {
    XRH_REQUIRE(p->smooth_ksize == 0 || p->smooth_ksize == 1 || p->smooth_ksize == 3 || p->smooth_ksize == 5 ||
                    p->smooth_ksize == 7,
                "smoothing aperture must be 0 (off) or an odd size <= 7");
  }
  if (patch) {
    XRH_REQUIRE(p->patch_radius >= 1 && p->patch_radius <= 10, "patch radius must be in [1, 10]");
    XRH_REQUIRE(p->patch_stride >= 1 && p->patch_stride <= 4, "patch stride must be in [1, 4]");
    XRH_REQUIRE(2 * p->patch_radius + 1 <= rows && 2 * p->patch_radius + 1 <= cols,
                "patch diameter exceeds the image");  // xregImgSimMetric2DPatchCommon.cpp:271-272

  }
  auto* f = new xrh_fixed();
  f->ctx = ctx;
  try {
    FixedDev& h = f->host;
    std::memset(&h, 0, sizeof(h));
    h.metric = p->metric;
    h.rows = rows;
    h.cols = cols;
    h.smooth_ksize = (grad && p->smooth_ksize > 1) ? p->smooth_ksize : 0;
    h.patch_radius = p->patch_radius;
    h.patch_stride = p->patch_stride > 0 ? p->patch_stride : 1;
    const int64_t n = (int64_t)rows * cols;
    hipStream_t s = ctx->stream;
    float* img = fixed_buf(f, (uint64_t)n);
    XRH_HIP_CHECK(hipMemcpyAsync(img, fixed_host, (size_t)n * 4, hipMemcpyHostToDevice, s));
    uint8_t* mask = nullptr;
    if (mask_host) {
      mask = (uint8_t*)ctx_alloc(ctx, (uint64_t)n);
      f->buffers.push_back(mask);
      XRH_HIP_CHECK(hipMemcpyAsync(mask, mask_host, (size_t)n, hipMemcpyHostToDevice, s));
      h.mask = mask;
    }
    const float* gx = img;
    const float* gy = nullptr;
    if (grad) {
      float* bx = fixed_buf(f, (uint64_t)n);
      float* by = fixed_buf(f, (uint64_t)n);
      const int TX = (cols + kT - 1) / kT, TY = (rows + kT - 1) / kT;
      hipLaunchKernelGGL(fixed_grad_kernel, dim3((unsigned)(TX * TY)), dim3(kThreads), 0, s, img, rows, cols,
                         h.smooth_ksize, bx, by);
      XRH_HIP_CHECK(hipGetLastError());
      gx = bx;
      gy = by;
    }
    if (!patch) {
      float* z = nullptr;
      ncc_prepare(f, gx, n, mask, &z, &h.sum_zf, &h.std_f, &h.count);
      h.zf = z;
      if (grad) {
        int64_t c2 = 0;
        ncc_prepare(f, gy, n, mask, &z, &h.sum_zfy, &h.std_fy, &c2);
        h.zfy = z;
      }
    } else {
      const int d = 2 * h.patch_radius + 1;
      h.ppr = (cols - d) / h.patch_stride + 1;
      h.ppc = (rows - d) / h.patch_stride + 1;
      h.num_patches = (int64_t)h.ppr * h.ppc;
      h.count = n;
      h.pf = gx;
      h.patch_norm[0] = h.patch_norm[1] = (double)h.num_patches;
      h.patch_offset[0] = h.patch_offset[1] = 0.0;
      f->create_weights.assign((size_t)h.num_patches, 1.0f);
      if (mask_host) {
        // ImgSimMetric2DPatchCommon::compute_weights with a mask and the default options
        // (xregImgSimMetric2DPatchCommon.cpp:306-400): unmasked fraction of each patch,
        // normalised to sum to one, in float
        std::vector<float>& w = f->create_weights;
        float wsum = 0.0f;
        for (int64_t q = 0; q < h.num_patches; ++q) {
          const int r0 = (int)(q / h.ppr) * h.patch_stride, c0 = (int)(q % h.ppr) * h.patch_stride;
          int64_t cnt = 0;
          for (int y = 0; y < d; ++y)
            for (int x = 0; x < d; ++x) cnt += mask_host[(int64_t)(r0 + y) * cols + c0 + x] ? 1 : 0;
          w[(size_t)q] = (float)cnt / (float)(d * d);
        }
        for (float v : w) wsum += v;
        XRH_REQUIRE(wsum > 0.0f, "the mask leaves no unmasked pixel in any patch");
        for (float& v : w) v /= wsum;
      }
      float *mu = nullptr, *den = nullptr;
      double* g = nullptr;
      // use_mask_for_patch_stats: only with a mask (without one it changes nothing)
      float* count = (p->patch_mask_stats && mask) ? fixed_buf(f, (uint64_t)h.num_patches) : nullptr;
      h.pcount = count;
      patch_prepare(f, gx, &mu, &den, &g, count);
      h.pmean = mu;
      h.pden = den;
      h.pg = g;
      f->base_den_dev[0] = den;
      h.pden_unit = den;
      if (grad) {
        h.pfy = gy;
        patch_prepare(f, gy, &mu, &den, &g, count);
        h.pmeany = mu;
        h.pdeny = den;
        h.pgy = g;
        f->base_den_dev[1] = den;
        h.pdeny_unit = den;
      }
      for (int k = 0; k < (grad ? 2 : 1); ++k) {
        f->base_den[k].resize((size_t)h.num_patches);
        XRH_HIP_CHECK(hipMemcpyAsync(f->base_den[k].data(), f->base_den_dev[k], sizeof(float) * h.num_patches,
                                     hipMemcpyDeviceToHost, s));
      }
      XRH_HIP_CHECK(hipStreamSynchronize(s));
      f->weights = f->weights_y = f->create_weights;
      // the default combine: weighted by the patch weights, divided by their sum
      // (xregImgSimMetric2DPatchNCCCPU.cpp:252-253, 272-285); unit weights need no scaling
      if (mask_host) patch_coefficients(f, nullptr, nullptr, nullptr, 0, true, false, false);
    }
    if (patch) {
      int max_lds = 0;
      XRH_HIP_CHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
      if (sim_patch_lds_bytes(h) > max_lds) {
        fail(XRH_ERR_UNSUPPORTED, "patch radius / stride need " + std::to_string(sim_patch_lds_bytes(h)) +
                                      " bytes of LDS per workgroup; the device has " + std::to_string(max_lds));
      }
    }
    f->dev = (FixedDev*)ctx_alloc(ctx, sizeof(FixedDev));
    f->buffers.push_back(f->dev);
    XRH_HIP_CHECK(hipMemcpyAsync(f->dev, &h, sizeof(FixedDev), hipMemcpyHostToDevice, s));
    XRH_HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipStreamSynchronize(ctx->stream);
    for (void* b : f->buffers) ctx_free(ctx, b);
    delete f;
    throw;
  }
  *out = f;
}

static void sim_compute(xrh_ctx* ctx, const xrh_fixed* fixed, const float* mov_dev, int64_t num_mov,
                        float* sim_dev_out, float* sim_host_out, float* scores_host_out = nullptr)
{
  XRH_REQUIRE(ctx && fixed && mov_dev, "null argument");
  XRH_REQUIRE(fixed->ctx == ctx, "fixed image belongs to another context");
  XRH_REQUIRE(num_mov >= 1, "need at least one moving image");
  XRH_REQUIRE(sim_dev_out || sim_host_out, "no output for similarity values");
  const FixedDev& h = fixed->host;
  const int64_t npix = (int64_t)h.rows * h.cols;
  ctx_check_range(ctx, mov_dev, (uint64_t)(npix * num_mov) * 4, "moving images");
  if (sim_dev_out) ctx_check_range(ctx, sim_dev_out, (uint64_t)num_mov * 4, "similarity output");
  const int img_tiles = sim_img_tiles(h);
  const int ptiles = sim_tiles(h);
  const int rec = sim_record_stride(h.metric);
  auto align = [](uint64_t x) { return (x + 255) & ~uint64_t(255); };
  const bool need_grad = (h.metric == XRH_SIM_PATCH_GRAD_NCC);
  const uint64_t off_items = 0;
  const uint64_t off_part = align(sizeof(SimItem) * num_mov);
  const uint64_t off_sim = align(off_part + sizeof(double) * rec * ptiles * num_mov);
  const uint64_t off_grad = align(off_sim + sizeof(float) * num_mov);
  const uint64_t off_scores = align(off_grad + (need_grad ? sizeof(float) * 2 * npix * num_mov : 0));
  const uint64_t total = align(off_scores + (scores_host_out ? sizeof(float) * h.num_patches * num_mov : 0));
  XRH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  char* dev = (char*)ctx_scratch(ctx, total);
  SimItem* hi = (SimItem*)ctx_pinned(ctx, sizeof(SimItem) * num_mov);
  for (int64_t i = 0; i < num_mov; ++i) {
    hi[i].mov = mov_dev + i * npix;
    hi[i].fixed = fixed->dev;
    hi[i].gx = need_grad ? (float*)(dev + off_grad) + 2 * i * npix : nullptr;
    hi[i].gy = need_grad ? (float*)(dev + off_grad) + (2 * i + 1) * npix : nullptr;
    hi[i].out_index = i;
    hi[i].patch_scores = scores_host_out ? (float*)(dev + off_scores) + i * h.num_patches : nullptr;
  }
  XRH_HIP_CHECK(hipMemcpyAsync(dev + off_items, hi, sizeof(SimItem) * num_mov, hipMemcpyHostToDevice, ctx->stream));
  float* sims = sim_dev_out ? sim_dev_out : (float*)(dev + off_sim);
  int key = sim_group_key(h);
  if (scores_host_out) {
    // the general patch kernel writes the per-patch scores: clear the group key's diameter
    const int dmask = ((1 << (kSimKeyBlurShift - kSimKeyShift)) - 1) << kSimKeyShift;
    key &= ~dmask;
  }
  launch_sim(key, (const SimItem*)(dev + off_items), num_mov, img_tiles, ptiles, (double*)(dev + off_part), sims,
             false, sim_patch_lds_bytes(h, scores_host_out != nullptr), ctx->stream);
  if (scores_host_out) {
    XRH_HIP_CHECK(hipMemcpyAsync(scores_host_out, dev + off_scores, sizeof(float) * h.num_patches * num_mov,
                                 hipMemcpyDeviceToHost, ctx->stream));
  }
  if (sim_host_out) {
    XRH_HIP_CHECK(hipMemcpyAsync(sim_host_out, sims, sizeof(float) * num_mov, hipMemcpyDeviceToHost, ctx->stream));
  }
  if (sim_host_out || scores_host_out) XRH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
}

}  // namespace xrh

using namespace xrh;

extern "C" {

xrh_status xrh_fixed_create(xrh_ctx* ctx, const float* fixed_host, int32_t rows, int32_t cols,
                            const xrh_sim_params* params, const uint8_t* mask_host, xrh_fixed** out)
{
  return guarded([&] { fixed_create(ctx, fixed_host, rows, cols, params, mask_host, out); });
}

xrh_status xrh_fixed_destroy(xrh_fixed* fixed)
{
  return guarded([&] {
    if (!fixed) return;
    XRH_HIP_CHECK(hipStreamSynchronize(fixed->ctx->stream));
    for (void* b : fixed->buffers) ctx_free(fixed->ctx, b);
    delete fixed;
  });
}

xrh_status xrh_fixed_patches(const xrh_fixed* fixed, int64_t capacity, int32_t* start_row, int32_t* start_col,
                             int32_t* stop_row, int32_t* stop_col, float* weight, int64_t* num_patches)
{
  return guarded([&] {
    XRH_REQUIRE(fixed && num_patches, "null argument");
    const FixedDev& h = fixed->host;
    XRH_REQUIRE(h.metric == XRH_SIM_PATCH_NCC || h.metric == XRH_SIM_PATCH_GRAD_NCC, "not a patch metric");
    *num_patches = h.num_patches;
    if (capacity <= 0) return;
    XRH_REQUIRE(capacity >= h.num_patches, "patch arrays shorter than the patch count");
    // ImgSimMetric2DPatchCommon::setup_patches (xregImgSimMetric2DPatchCommon.cpp:274-292):
    // centres r, r + stride, ... in each axis, rows outer
    const int r = h.patch_radius;
    for (int64_t q = 0; q < h.num_patches; ++q) {
      const int cr = r + (int)(q / h.ppr) * h.patch_stride, cc = r + (int)(q % h.ppr) * h.patch_stride;
      if (start_row) start_row[q] = cr - r;
      if (start_col) start_col[q] = cc - r;
      if (stop_row) stop_row[q] = cr + r;
      if (stop_col) stop_col[q] = cc + r;
    }
    if (weight) std::copy(fixed->weights.begin(), fixed->weights.end(), weight);
  });
}

xrh_status xrh_fixed_set_patch_weights(xrh_fixed* fixed, const float* weights, const float* weights_y,
                                       const int64_t* patch_inds, int64_t num_inds, int32_t weight_in_combine,
                                       int32_t compute_mean)
{
  return guarded([&] {
    XRH_REQUIRE(fixed, "null argument");
    const FixedDev& h = fixed->host;
    XRH_REQUIRE(h.metric == XRH_SIM_PATCH_NCC || h.metric == XRH_SIM_PATCH_GRAD_NCC, "not a patch metric");
    patch_coefficients(fixed, weights, weights_y, patch_inds, num_inds, weight_in_combine != 0, compute_mean != 0,
                       true);
  });
}

xrh_status xrh_fixed_image(const xrh_fixed* fixed, int32_t which, float* host_out)
{
  return guarded([&] {
    XRH_REQUIRE(fixed && host_out, "null argument");
    const FixedDev& h = fixed->host;
    const bool grad = h.metric == XRH_SIM_GRAD_NCC || h.metric == XRH_SIM_PATCH_GRAD_NCC;
    const bool patch = h.metric == XRH_SIM_PATCH_NCC || h.metric == XRH_SIM_PATCH_GRAD_NCC;
    XRH_REQUIRE(which == 0 || (which == 1 && grad), "which must be 0, or 1 for a gradient metric");
    XRH_REQUIRE(patch, "processed images are kept for the patch metrics");
    const float* src = (which == 0) ? h.pf : h.pfy;
    XRH_HIP_CHECK(hipMemcpyAsync(host_out, src, sizeof(float) * (size_t)h.rows * h.cols, hipMemcpyDeviceToHost,
                                 fixed->ctx->stream));
    XRH_HIP_CHECK(hipStreamSynchronize(fixed->ctx->stream));
  });
}

xrh_status xrh_sim_compute(xrh_ctx* ctx, const xrh_fixed* fixed, const float* mov_dev, int64_t num_mov,
                           float* sim_dev_out, float* sim_host_out)
{
  return guarded([&] { sim_compute(ctx, fixed, mov_dev, num_mov, sim_dev_out, sim_host_out); });
}

xrh_status xrh_sim_patch_scores(xrh_ctx* ctx, const xrh_fixed* fixed, const float* mov_dev, int64_t num_mov,
                                float* sim_host_out, float* scores_host_out)
{
  return guarded([&] {
    XRH_REQUIRE(fixed && sim_host_out && scores_host_out, "null argument");
    const FixedDev& h = fixed->host;
    XRH_REQUIRE(h.metric == XRH_SIM_PATCH_NCC || h.metric == XRH_SIM_PATCH_GRAD_NCC, "not a patch metric");
    sim_compute(ctx, fixed, mov_dev, num_mov, nullptr, sim_host_out, scores_host_out);
  });
}

}  // extern "C"
