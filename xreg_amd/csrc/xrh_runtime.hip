// xrh_runtime.hip -- context, budgeted device memory, resident volumes and xrh_line_int.
#include "xrh_internal.h"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdlib>

namespace xrh {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

[[noreturn]] void fail(xrh_status s, const std::string& msg) { throw Error{s, msg}; }

void* ctx_alloc(xrh_ctx* ctx, uint64_t bytes)
{
  if (bytes == 0) bytes = 256;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->budget && ctx->in_use + bytes > ctx->budget) {
      fail(XRH_ERR_BUDGET, "HBM budget exceeded: " + std::to_string(ctx->in_use) + " + " +
                               std::to_string(bytes) + " > " + std::to_string(ctx->budget));
    }
    ctx->in_use += bytes;
  }
  void* p = nullptr;
  const hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->in_use -= bytes;
    (void)hipGetLastError();
    fail(XRH_ERR_OOM, "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->allocs[p] = bytes;
  return p;
}

void ctx_free(xrh_ctx* ctx, void* p)
{
  if (!p) return;
  uint64_t bytes = 0;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->allocs.find(p);
    XRH_REQUIRE(it != ctx->allocs.end(), "device pointer was not allocated by this context");
    bytes = it->second;
    ctx->allocs.erase(it);
    ctx->in_use -= bytes;
  }
  XRH_HIP_CHECK(hipFree(p));
}

void ctx_check_range(xrh_ctx* ctx, const void* p, uint64_t bytes, const char* what)
{
  if (!p || bytes == 0) return;
  const uintptr_t a = (uintptr_t)p;
  std::lock_guard<std::mutex> lk(ctx->mu);
  for (const auto& kv : ctx->allocs) {
    const uintptr_t b = (uintptr_t)kv.first;
    if (a >= b && a < b + kv.second) {
      if (bytes > (b + kv.second) - a) {
        fail(XRH_ERR_INVALID, std::string(what) + ": " + std::to_string(bytes) + " bytes from offset " +
                                  std::to_string(a - b) + " run past the end of a " +
                                  std::to_string(kv.second) + "-byte allocation");
      }
      return;
    }
  }
  // not one of this context's allocations: the caller vouches for its size
}

void* ctx_scratch(xrh_ctx* ctx, uint64_t bytes)
{
  if (ctx->scratch_bytes >= bytes) return ctx->scratch;
  XRH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  if (ctx->scratch) ctx_free(ctx, ctx->scratch);
  ctx->scratch = nullptr;
  ctx->scratch_bytes = 0;
  const uint64_t want = std::max<uint64_t>(bytes, 1ull << 20);
  ctx->scratch = ctx_alloc(ctx, want);
  ctx->scratch_bytes = want;
  return ctx->scratch;
}

void* ctx_pinned(xrh_ctx* ctx, uint64_t bytes)
{
  // the caller has synchronised the stream before reusing the staging area
  if (ctx->pinned_bytes >= bytes) return ctx->pinned;
  if (ctx->pinned) XRH_HIP_CHECK(hipHostFree(ctx->pinned));
  ctx->pinned = nullptr;
  ctx->pinned_bytes = 0;
  const uint64_t want = std::max<uint64_t>(bytes, 1ull << 20);
  XRH_HIP_CHECK(hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault));
  ctx->pinned_bytes = want;
  return ctx->pinned;
}

void check_camera(const xrh_camera& c)
{
  XRH_REQUIRE(c.num_det_rows > 0 && c.num_det_cols > 0, "camera detector size must be positive");
  XRH_REQUIRE(c.num_det_rows <= 32768 && c.num_det_cols <= 32768, "camera detector too large");
  XRH_REQUIRE(c.coord_frame >= 0 && c.coord_frame <= 2, "camera coord_frame must be 0, 1 or 2");
  XRH_REQUIRE(c.focal_len > 1.0e-8f, "camera focal length must be positive");
}

CamDev make_cam_dev(const xrh_camera& c)
{
  check_camera(c);
  CamDev d;
  std::memset(&d, 0, sizeof(d));
  for (int i = 0; i < 9; ++i) {
    d.kinv[i] = c.intrins_inv[i];
    d.kinv_d[i] = (double)c.intrins_inv[i];
  }
  for (int i = 0; i < 12; ++i) {
    d.einv[i] = c.extrins_inv[i];
    d.einv_d[i] = (double)c.extrins_inv[i];
  }
  for (int i = 0; i < 3; ++i) {
    d.pinhole[i] = c.pinhole[i];
    d.pinhole_d[i] = (double)c.pinhole[i];
  }
  // det_z_val = ((frame == NEG_Z) ? -1 : 1) * focal_len (xregPerspectiveXform.cpp:411)
  d.det_z = ((c.coord_frame == 1) ? -1.0f : 1.0f) * c.focal_len;
  d.det_z_d = (double)d.det_z;
  d.z_adj = (c.coord_frame == 2) ? -c.focal_len : 0.0f;
  d.z_adj_d = (double)d.z_adj;
  d.rows = c.num_det_rows;
  d.cols = c.num_det_cols;
  return d;
}

// float 3x3 inverse in Eigen's cofactor order (Eigen/src/LU/InverseImpl.h), used for the
// phys->idx transform exactly as RayCasterLineIntCPU::compute obtains it via
// FrameTransform::inverse() (xregRayCastLineIntCPU.cpp:307-309).
static void affine_inverse_f32(const float m[16], float out[12])
{
#define M3(r, c) lin[3 * (r) + (c)]
#define COF(i, j) (M3(((i) + 1) % 3, ((j) + 1) % 3) * M3(((i) + 2) % 3, ((j) + 2) % 3) - \
                   M3(((i) + 1) % 3, ((j) + 2) % 3) * M3(((i) + 2) % 3, ((j) + 1) % 3))
  const float lin[9] = {m[0], m[1], m[2], m[4], m[5], m[6], m[8], m[9], m[10]};
  float c00 = COF(0, 0), c10 = COF(1, 0), c20 = COF(2, 0);
  const float det = (c00 * M3(0, 0) + c10 * M3(1, 0)) + c20 * M3(2, 0);
  const float invdet = 1.0f / det;
  const float inv[9] = {c00 * invdet,       c10 * invdet,       c20 * invdet,
                        COF(0, 1) * invdet, COF(1, 1) * invdet, COF(2, 1) * invdet,
                        COF(0, 2) * invdet, COF(1, 2) * invdet, COF(2, 2) * invdet};
#undef COF
#undef M3
  for (int i = 0; i < 3; ++i) {
    out[4 * i + 0] = inv[3 * i + 0];
    out[4 * i + 1] = inv[3 * i + 1];
    out[4 * i + 2] = inv[3 * i + 2];
    const float t = (inv[3 * i] * m[3] + inv[3 * i + 1] * m[7]) + inv[3 * i + 2] * m[11];
    out[4 * i + 3] = -t;
  }
}

static void affine_inverse_f64(const float mf[16], double out[12])
{
  double m[16];
  for (int i = 0; i < 16; ++i) m[i] = (double)mf[i];
  const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], k = m[10];
  const double c00 = e * k - f * h, c01 = f * g - d * k, c02 = d * h - e * g;
  const double det = a * c00 + b * c01 + c * c02;
  const double inv[9] = {c00 / det, (c * h - b * k) / det, (b * f - c * e) / det,
                         c01 / det, (a * k - c * g) / det, (c * d - a * f) / det,
                         c02 / det, (b * g - a * h) / det, (a * e - b * d) / det};
  for (int i = 0; i < 3; ++i) {
    out[4 * i + 0] = inv[3 * i + 0];
    out[4 * i + 1] = inv[3 * i + 1];
    out[4 * i + 2] = inv[3 * i + 2];
    out[4 * i + 3] = -(inv[3 * i] * m[3] + inv[3 * i + 1] * m[7] + inv[3 * i + 2] * m[11]);
  }
}

namespace {

__device__ inline uint16_t f2h_bits(float v) { return __half_as_ushort(__float2half_rn(v)); }

__device__ inline uint16_t f2bf_bits(float v)
{
  const uint32_t u = __float_as_uint(v);
  const uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite input)
  return (uint16_t)(r >> 16);
}

// Builds the padded layout from the raw x-fastest volume. One thread per padded voxel.
__global__ void build_layout_kernel(const float* __restrict__ src, int X, int Y, int Z,
                                    void* __restrict__ dst, int storage, int layout)
{
  const int64_t Xp = X + 2, Yp = Y + 2, Zp = Z + 2;
  const int64_t n = Xp * Yp * Zp;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int a = (int)(i % Xp);
  const int b = (int)((i / Xp) % Yp);
  const int c = (int)(i / (Xp * Yp));
  auto V = [&](int aa, int bb, int cc) -> float {
    // padded index -> source index with edge replication
    int x = min(max(aa - 1, 0), X - 1);
    int y = min(max(bb - 1, 0), Y - 1);
    int z = min(max(cc - 1, 0), Z - 1);
    return src[(int64_t)x + (int64_t)X * ((int64_t)y + (int64_t)Y * z)];
  };
  if (layout == kLayoutLinear) {
    const float v = V(a, b, c);
    if (storage == XRH_F32) ((float*)dst)[i] = v;
    else if (storage == XRH_F16) ((uint16_t*)dst)[i] = f2h_bits(v);
    else ((uint16_t*)dst)[i] = f2bf_bits(v);
  } else if (layout == kLayoutQuad) {
    float4 q;
    q.x = V(a, b, c);
    q.y = V(a + 1, b, c);
    q.z = V(a, b + 1, c);
    q.w = V(a + 1, b + 1, c);
    ((float4*)dst)[i] = q;
  } else if (layout == kLayoutFace) {
    // face normal to x at (a, b, c): corners (b, c), (b+1, c), (b, c+1), (b+1, c+1) over y, z;
    // bilinear coefficients f(v, w) = c0 + c1 v + c2 w + c3 v w formed in double, rounded once
    const double v00 = V(a, b, c), v10 = V(a, b + 1, c), v01 = V(a, b, c + 1), v11 = V(a, b + 1, c + 1);
    ((float4*)dst)[i] = make_float4((float)v00, (float)(v10 - v00), (float)(v01 - v00),
                                    (float)((v11 - v01) - (v10 - v00)));
  } else if (layout == kLayoutBrick) {
    // the face of (a, b, c) as in kLayoutFace, written to its brick and, when it is the first
    // x-face of a brick, also as the last x-face of the previous brick in x
    const double v00 = V(a, b, c), v10 = V(a, b + 1, c), v01 = V(a, b, c + 1), v11 = V(a, b + 1, c + 1);
    const float4 f = make_float4((float)v00, (float)(v10 - v00), (float)(v01 - v00),
                                 (float)((v11 - v01) - (v10 - v00)));
    const int nbx = (int)((Xp - 1 + kBrickEdge - 1) / kBrickEdge);   // bricks of cells 0..X
    const int nby = (int)((Yp + kBrickEdge - 1) / kBrickEdge);
    const int by = b / kBrickEdge, bz = c / kBrickEdge;
    const int local_yz = (c % kBrickEdge) * (kBrickEdge + 1) * kBrickEdge + (b % kBrickEdge) * (kBrickEdge + 1);
    const int bx = a / kBrickEdge;
    if (bx < nbx) {
      const int64_t brick = ((int64_t)bz * nby + by) * nbx + bx;
      ((float4*)dst)[brick * kBrickFaces + local_yz + (a % kBrickEdge)] = f;
    }
    if (a % kBrickEdge == 0 && a > 0) {
      const int64_t brick = ((int64_t)bz * nby + by) * nbx + (bx - 1);
      ((float4*)dst)[brick * kBrickFaces + local_yz + kBrickEdge] = f;
    }
  } else if (layout == kLayoutCoef) {
    // f(u, v, w) = c0 + c1 u + c2 v + c3 u v + w (c4 + c5 u + c6 v + c7 u v) over the cell's
    // corners, differences taken in double and rounded once
    const double v000 = V(a, b, c), v100 = V(a + 1, b, c), v010 = V(a, b + 1, c), v110 = V(a + 1, b + 1, c);
    const double v001 = V(a, b, c + 1), v101 = V(a + 1, b, c + 1), v011 = V(a, b + 1, c + 1),
                 v111 = V(a + 1, b + 1, c + 1);
    const double c3 = (v110 - v010) - (v100 - v000);
    const double c7 = ((v111 - v011) - (v101 - v001)) - c3;
    float4* o = (float4*)dst + 2 * i;
    o[0] = make_float4((float)v000, (float)(v100 - v000), (float)(v010 - v000), (float)c3);
    o[1] = make_float4((float)(v001 - v000), (float)((v101 - v001) - (v100 - v000)),
                       (float)((v011 - v001) - (v010 - v000)), (float)c7);
  } else {
    const float f[8] = {V(a, b, c),     V(a + 1, b, c),     V(a, b + 1, c),     V(a + 1, b + 1, c),
                        V(a, b, c + 1), V(a + 1, b, c + 1), V(a, b + 1, c + 1), V(a + 1, b + 1, c + 1)};
    uint32_t w[4];
    for (int k = 0; k < 4; ++k) {
      const uint32_t lo = (storage == XRH_F16) ? f2h_bits(f[2 * k]) : f2bf_bits(f[2 * k]);
      const uint32_t hi = (storage == XRH_F16) ? f2h_bits(f[2 * k + 1]) : f2bf_bits(f[2 * k + 1]);
      w[k] = lo | (hi << 16);
    }
    ((uint4*)dst)[i] = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

// One thread per brick of B^3 voxels: 1 if any voxel within one voxel of the brick's
// trilinear support [b*B - 1, b*B + B + 1] (clamped to the volume) is non-zero.
__global__ void brick_occupancy_kernel(const float* __restrict__ src, int X, int Y, int Z, int B, int gx, int gy,
                                       int gz, uint8_t* __restrict__ occ)
{
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= (int64_t)gx * gy * gz) return;
  const int bx = (int)(b % gx), by = (int)((b / gx) % gy), bz = (int)(b / ((int64_t)gx * gy));
  const int x0 = max(bx * B - 1, 0), x1 = min(bx * B + B + 1, X - 1);
  const int y0 = max(by * B - 1, 0), y1 = min(by * B + B + 1, Y - 1);
  const int z0 = max(bz * B - 1, 0), z1 = min(bz * B + B + 1, Z - 1);
  uint8_t any = 0;
  for (int z = z0; z <= z1 && !any; ++z)
    for (int y = y0; y <= y1 && !any; ++y)
      for (int x = x0; x <= x1; ++x)
        if (src[(int64_t)x + (int64_t)X * ((int64_t)y + (int64_t)Y * z)] != 0.0f) {
          any = 1;
          break;
        }
  occ[b] = any;
}

__global__ void copy_bg_kernel(const float* __restrict__ bg, const int64_t* __restrict__ cam_for_proj,
                               int64_t npx, int64_t num_projs, float* __restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npx * num_projs) return;
  const int64_t p = i / npx;
  out[i] = bg[cam_for_proj[p] * npx + (i - p * npx)];
}

}  // namespace

// Exact Chebyshev (chessboard) distance to the nearest occupied brick: two raster passes of
// the 26-neighbour chamfer mask with unit weights. Values are capped at 255.
static void chebyshev_distance(std::vector<uint8_t>& d, int gx, int gy, int gz)
{
  auto at = [&](int x, int y, int z) -> uint8_t& { return d[(size_t)x + (size_t)gx * ((size_t)y + (size_t)gy * z)]; };
  for (auto& v : d) v = v ? 0 : 255;
  for (int pass = 0; pass < 2; ++pass) {
    const int s = pass == 0 ? 1 : -1;
    for (int zi = 0; zi < gz; ++zi) {
      const int z = pass == 0 ? zi : gz - 1 - zi;
      for (int yi = 0; yi < gy; ++yi) {
        const int y = pass == 0 ? yi : gy - 1 - yi;
        for (int xi = 0; xi < gx; ++xi) {
          const int x = pass == 0 ? xi : gx - 1 - xi;
          int best = at(x, y, z);
          // neighbours already visited in this pass: earlier in raster order
          for (int dz = -1; dz <= 0; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
              for (int dx = -1; dx <= 1; ++dx) {
                if (dz == 0 && (dy > 0 || (dy == 0 && dx >= 0))) continue;
                const int nx = x + s * dx, ny = y + s * dy, nz = z + s * dz;
                if (nx < 0 || ny < 0 || nz < 0 || nx >= gx || ny >= gy || nz >= gz) continue;
                best = std::min(best, (int)at(nx, ny, nz) + 1);
              }
          at(x, y, z) = (uint8_t)std::min(best, 255);
        }
      }
    }
  }
}

// log2 of the empty-space brick edge (XRH_SKIP_BRICK_SHIFT = 1..4). Default 2 (4^3 voxels):
// at C2 (256x256x200, 0.5 mm steps) 8^3 bricks evaluate 5.62e9 samples per lambda = 100 launch
// (5.68 ms), 4^3 bricks 5.38e9 (5.55 ms), 2^3 bricks 5.25e9 (5.54 ms: the longer walks eat
// the gain).
static int skip_brick_shift()
{
  const char* env = std::getenv("XRH_SKIP_BRICK_SHIFT");
  if (env && env[0] >= '1' && env[0] <= '4' && env[1] == 0) return env[0] - '0';
  return 2;
}

static bool skip_enabled()
{
  const char* env = std::getenv("XRH_SKIP_EMPTY");
  return !(env && std::strcmp(env, "0") == 0);
}

static int default_layout(int storage)
{
  const char* env = std::getenv("XRH_VOLUME_LAYOUT");
  if (env && std::strcmp(env, "linear") == 0) return kLayoutLinear;
  if (env && std::strcmp(env, "quad") == 0 && storage == XRH_F32) return kLayoutQuad;
  if (env && std::strcmp(env, "coef") == 0 && storage == XRH_F32) return kLayoutCoef;
  if (env && std::strcmp(env, "brick") == 0 && storage == XRH_F32) return kLayoutBrick;
  // FP32: x-face bilinear coefficients; 16-bit: all corners in one 16-byte element
  return (storage == XRH_F32) ? kLayoutFace : kLayoutOct;
}

// HUToLinAttFilter::GenerateData (lib/image/xregHUToLinAtt.cpp:45-68): per voxel, in double,
// max((hu * hu_scale + mu_water) - mu_lower, 0) rounded to float; std::max keeps a NaN.
__global__ void hu_to_lin_att_kernel(float* __restrict__ v, int64_t n, double hu_scale, double mu_water,
                                     double mu_lower)
{
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = ((double)v[i] * hu_scale + mu_water) - mu_lower;
  v[i] = (float)((a < 0.0) ? 0.0 : a);
}

struct HuParams {
  double hu_lower, mu_water, mu_air;
};

static void volume_create(xrh_ctx* ctx, const float* vox, const int64_t size[3],
                          const float idx_to_phys[16], int32_t storage, int32_t layout,
                          xrh_volume** out, const HuParams* hu = nullptr)
{
  XRH_REQUIRE(vox && size && idx_to_phys && out, "null argument");
  XRH_REQUIRE(storage == XRH_F32 || storage == XRH_F16 || storage == XRH_BF16, "bad storage type");
  for (int k = 0; k < 3; ++k) XRH_REQUIRE(size[k] >= 1 && size[k] < (1 << 20), "bad volume size");
  // the DRR kernels form element indices with signed 24-bit multiply-adds (iz * pitch_z and
  // iy * pitch_y), exact only while both padded pitches stay below 2^23 elements
  if (!((size[0] + 2) * (size[1] + 2) < (int64_t(1) << 23))) {
    fail(XRH_ERR_UNSUPPORTED, "volume slice too large: (size_x + 2) * (size_y + 2) must stay below 2^23 voxels");
  }
  XRH_REQUIRE(ctx, "null context");
  const bool auto_layout = layout < 0;
  if (auto_layout) layout = default_layout(storage);
  if (layout == kLayoutQuad || layout == kLayoutCoef || layout == kLayoutFace || layout == kLayoutBrick) {
    XRH_REQUIRE(storage == XRH_F32, "layout needs FP32 storage");
  }
  if (layout == kLayoutOct) XRH_REQUIRE(storage != XRH_F32, "oct layout needs 16-bit storage");
  const int64_t X = size[0], Y = size[1], Z = size[2];
  const int64_t padded = (X + 2) * (Y + 2) * (Z + 2);
  if (layout == kLayoutCoef && (uint64_t)padded * 32ull >= 0xFFFFF000ull) {
    XRH_REQUIRE(auto_layout, "coefficient layout exceeds 4 GiB; use the quad layout");
    layout = kLayoutQuad;
  }
  // buffer loads address with 32-bit byte offsets: with the automatic choice, a volume whose
  // packed layout would pass 4 GiB is kept in the padded linear layout (one scalar per voxel)
  const uint64_t packed_esize = (layout == kLayoutCoef) ? 32 : 16;
  if (auto_layout && layout != kLayoutLinear && (uint64_t)padded * packed_esize >= 0xFFFFF000ull) {
    layout = kLayoutLinear;
  }
  const int64_t esize = (layout == kLayoutLinear) ? (storage == XRH_F32 ? 4 : 2) : (layout == kLayoutCoef ? 32 : 16);
  // bricks: cells 0..X in x, padded faces 0..Y+1 and 0..Z+1 in y and z, 5x4x4 faces each
  const int64_t nbx = (X + 1 + kBrickEdge - 1) / kBrickEdge, nby = (Y + 2 + kBrickEdge - 1) / kBrickEdge,
                nbz = (Z + 2 + kBrickEdge - 1) / kBrickEdge;
  const uint64_t bytes = (layout == kLayoutBrick) ? (uint64_t)(nbx * nby * nbz * kBrickFaces * 16)
                                                  : (uint64_t)(padded * esize);
  XRH_REQUIRE(bytes < 0xFFFFF000ull, "volume layout exceeds 4 GiB; use the linear layout or split the volume");
  auto* v = new xrh_volume();
  v->ctx = ctx;
  v->size[0] = X;
  v->size[1] = Y;
  v->size[2] = Z;
  void* raw = nullptr;
  int skip_dims[3] = {0, 0, 0};
  const int kSkipShift = skip_brick_shift();
  try {
    v->data = ctx_alloc(ctx, bytes);
    v->bytes = bytes;
    raw = ctx_alloc(ctx, (uint64_t)(X * Y * Z) * 4);
    XRH_HIP_CHECK(hipMemcpyAsync(raw, vox, (size_t)(X * Y * Z) * 4, hipMemcpyHostToDevice, ctx->stream));
    const int threads = 256;
    if (hu) {
      const double hu_scale = (hu->mu_water - hu->mu_air) * 1.0e-3;
      const double mu_lower = (hu->hu_lower * hu_scale) + hu->mu_water;
      const int64_t nv = X * Y * Z;
      hipLaunchKernelGGL(hu_to_lin_att_kernel, dim3((unsigned)((nv + threads - 1) / threads)), dim3(threads), 0,
                         ctx->stream, (float*)raw, nv, hu_scale, hu->mu_water, mu_lower);
      XRH_HIP_CHECK(hipGetLastError());
    }
    const int64_t blocks = (padded + threads - 1) / threads;
    // brick entries past the volume's faces are never sampled; zero them all the same
    if (layout == kLayoutBrick) XRH_HIP_CHECK(hipMemsetAsync(v->data, 0, bytes, ctx->stream));
    hipLaunchKernelGGL(build_layout_kernel, dim3((unsigned)blocks), dim3(threads), 0, ctx->stream,
                       (const float*)raw, (int)X, (int)Y, (int)Z, v->data, storage, layout);
    XRH_HIP_CHECK(hipGetLastError());
    const int SB = 1 << kSkipShift;
    const int64_t sgx = (X + SB - 1) / SB, sgy = (Y + SB - 1) / SB, sgz = (Z + SB - 1) / SB;
    // the walk indexes the map with 24-bit multiply-adds: larger maps go without skipping
    if (skip_enabled() && sgx * sgy * sgz < (int64_t(1) << 24) && sgy * sgz < (int64_t(1) << 24)) {
      const int B = SB;
      skip_dims[0] = (int)sgx;
      skip_dims[1] = (int)sgy;
      skip_dims[2] = (int)sgz;
      const int64_t nb = (int64_t)skip_dims[0] * skip_dims[1] * skip_dims[2];
      v->skip = ctx_alloc(ctx, (uint64_t)nb);
      hipLaunchKernelGGL(brick_occupancy_kernel, dim3((unsigned)((nb + 127) / 128)), dim3(128), 0, ctx->stream,
                         (const float*)raw, (int)X, (int)Y, (int)Z, B, skip_dims[0], skip_dims[1], skip_dims[2],
                         (uint8_t*)v->skip);
      XRH_HIP_CHECK(hipGetLastError());
      std::vector<uint8_t> occ((size_t)nb);
      XRH_HIP_CHECK(hipMemcpyAsync(occ.data(), v->skip, (size_t)nb, hipMemcpyDeviceToHost, ctx->stream));
      XRH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
      chebyshev_distance(occ, skip_dims[0], skip_dims[1], skip_dims[2]);
      XRH_HIP_CHECK(hipMemcpyAsync(v->skip, occ.data(), (size_t)nb, hipMemcpyHostToDevice, ctx->stream));
    }
    XRH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx_free(ctx, raw);
    raw = nullptr;
  } catch (...) {
    (void)hipStreamSynchronize(ctx->stream);
    if (raw) ctx_free(ctx, raw);
    if (v->data) ctx_free(ctx, v->data);
    if (v->skip) ctx_free(ctx, v->skip);
    delete v;
    throw;
  }
  VolDesc& d = v->desc;
  std::memset(&d, 0, sizeof(d));
  d.skip = (const uint8_t*)v->skip;
  d.skip_gx = skip_dims[0];
  d.skip_gy = skip_dims[1];
  d.skip_gz = skip_dims[2];
  d.skip_shift = kSkipShift;
  affine_inverse_f32(idx_to_phys, d.p2i);
  affine_inverse_f64(idx_to_phys, d.p2i_d);
  d.data = v->data;
  d.bytes = (uint32_t)bytes;
  d.X = (int32_t)X;
  d.Y = (int32_t)Y;
  d.Z = (int32_t)Z;
  if (layout == kLayoutBrick) {
    // brick pitches (bricks per row, per slice)
    d.pitch_y = (uint32_t)nbx;
    d.pitch_z = (uint32_t)(nbx * nby);
  } else {
    d.pitch_y = (uint32_t)(X + 2);
    d.pitch_z = (uint32_t)((X + 2) * (Y + 2));
  }
  d.storage = storage;
  d.layout = layout;
  float ms = INFINITY;
  double chord = 0.0;
  const int64_t dims[3] = {X, Y, Z};
  for (int c = 0; c < 3; ++c) {
    const float n = std::sqrt(idx_to_phys[c] * idx_to_phys[c] + idx_to_phys[4 + c] * idx_to_phys[4 + c] +
                              idx_to_phys[8 + c] * idx_to_phys[8 + c]);
    ms = std::min(ms, n);
    chord += (double)n * (double)(dims[c] + 1);  // +1: the nudge and edge samples
  }
  XRH_REQUIRE(ms > 0.0f && std::isfinite(ms), "idx_to_phys has a zero or non-finite column");
  d.min_spacing = ms;
  d.max_chord = (float)chord;
  *out = v;
}

// Band-major tile table of num_projs equal detectors into band_start ([bands][num_projs + 1],
// the layout build_tiles produces); returns the largest band's tile count. Writes only into
// the given array.
static int64_t build_tiles_uniform(int32_t rows, int32_t cols, int64_t num_projs, int32_t bands,
                                   uint32_t* band_start)
{
  const int TX = (cols + 15) / 16, TY = (rows + 15) / 16;
  int64_t max_tiles = 0;
  for (int b = 0; b < bands; ++b) {
    const int r0 = (int)(((int64_t)b * TY) / bands);
    const int r1 = (int)(((int64_t)(b + 1) * TY) / bands);
    const uint32_t per = (uint32_t)((r1 - r0) * TX);
    for (int64_t i = 0; i <= num_projs; ++i) band_start[(size_t)b * (num_projs + 1) + i] = (uint32_t)(i * per);
    max_tiles = std::max<int64_t>(max_tiles, (int64_t)per * num_projs);
  }
  return max_tiles;
}

// A caller-provided staging area (xrh_line_int_noalloc): `host` for the packed tables, `dev`
// for their device copy and the set-up kernel's items; both `bytes` long.
struct FixedStaging {
  unsigned char* host;
  uint64_t host_bytes;
  unsigned char* dev;
  uint64_t dev_bytes;
};

// Packs PoseIn/CamDev/VolDesc/band tables into one blob, uploads it and launches the set-up +
// DRR kernels. Shared by xrh_line_int, its sample counter and xrh_line_int_noalloc. Without
// `fixed` the blob goes through the context's pinned staging and device scratch (grown on
// demand) and a tuned launch variant applies; with `fixed` nothing is allocated: the blob
// must fit the given area and the default variant is used.
static void line_int_launch(xrh_ctx* ctx, const xrh_volume* vol, const xrh_camera* cams,
                            int32_t num_cams, const float* xforms, const int64_t* cam_for_proj,
                            int64_t num_projs, float step, int32_t kernel, int32_t store,
                            const float* bg_dev, float default_bg, int32_t geom, float* drr_dev,
                            uint64_t* samples_out, const FixedStaging* fixed = nullptr,
                            const InterpGrid* igrid = nullptr, int32_t interp = XRH_INTERP_LINEAR,
                            int32_t aa_rays = 0, uint64_t aa_seed = 0)
{
  XRH_REQUIRE(ctx && vol && cams && xforms && cam_for_proj, "null argument");
  XRH_REQUIRE(num_cams >= 1 && num_projs >= 1, "need at least one camera and one projection");
  XRH_REQUIRE(kernel == XRH_LINE_INT_SUM || kernel == XRH_LINE_INT_MAX, "bad line-integral kernel");
  XRH_REQUIRE(store == XRH_STORE_REPLACE || store == XRH_STORE_ACCUM, "bad store method");
  XRH_REQUIRE(geom == XRH_GEOM_FP32 || geom == XRH_GEOM_FP64, "bad geometry mode");
  XRH_REQUIRE(std::isfinite(step) && step > 0.0f, "ray step must be positive");
  check_step(step, vol->desc);
  XRH_REQUIRE(samples_out || drr_dev, "null DRR buffer");
  const int32_t rows = cams[0].num_det_rows, cols = cams[0].num_det_cols;
  for (int32_t c = 0; c < num_cams; ++c) {
    check_camera(cams[c]);
    // xregRayCastBaseCPU.cpp:54-70
    XRH_REQUIRE(cams[c].num_det_rows == rows, "Inconsistent number rows in camera model images!");
    XRH_REQUIRE(cams[c].num_det_cols == cols, "Inconsistent number columns in camera model images!");
  }
  for (int64_t p = 0; p < num_projs; ++p) {
    XRH_REQUIRE(cam_for_proj[p] >= 0 && cam_for_proj[p] < num_cams, "camera index out of range");
  }
  const int64_t npx = (int64_t)rows * cols;
  XRH_REQUIRE(npx * num_projs < (1ll << 40), "DRR buffer too large");
  ctx_check_range(ctx, drr_dev, (uint64_t)(npx * num_projs) * 4, "DRR buffer");
  if (bg_dev) ctx_check_range(ctx, bg_dev, (uint64_t)(npx * num_cams) * 4, "background images");

  // a variant tuned for this launch shape (xrh_batch_autotune), else the defaults
  DrrVariant var;
  if (!fixed) {
    const int64_t wgs = (int64_t)((rows + 15) / 16) * ((cols + 15) / 16) * num_projs;
    const std::string key = drr_tune_key(vol->desc.storage, vol->desc.layout, geom, kernel, rows, cols, wgs);
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->tuned.find(key);
    if (it != ctx->tuned.end()) var = it->second;
  }
  const int32_t bands = std::min(std::max(var.bands, 1), kMaxDrrBands);

  auto align = [](uint64_t x) { return (x + 255) & ~uint64_t(255); };
  const uint64_t off_pose = 0;
  const uint64_t off_cam = align(off_pose + sizeof(PoseIn) * num_projs);
  const uint64_t off_vol = align(off_cam + sizeof(CamDev) * num_cams);
  const uint64_t off_band = align(off_vol + sizeof(VolDesc));
  const uint64_t off_c4p = align(off_band + sizeof(uint32_t) * (uint64_t)bands * (num_projs + 1));
  const uint64_t off_cnt = align(off_c4p + sizeof(int64_t) * num_projs);
  const uint64_t blob = align(off_cnt + 16);
  const uint64_t off_items = blob;
  const uint64_t total = align(off_items + sizeof(DrrItem) * num_projs);

  XRH_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // staging reuse
  char* host;
  char* dev;
  if (fixed) {
    XRH_REQUIRE(blob <= fixed->host_bytes && total <= fixed->dev_bytes,
                "projections and cameras exceed the fixed staging area");
    host = (char*)fixed->host;
    dev = (char*)fixed->dev;
  } else {
    host = (char*)ctx_pinned(ctx, blob);
    dev = (char*)ctx_scratch(ctx, total);
  }
  const int64_t max_band_tiles = build_tiles_uniform(rows, cols, num_projs, bands, (uint32_t*)(host + off_band));
  PoseIn* pin = (PoseIn*)(host + off_pose);
  for (int64_t p = 0; p < num_projs; ++p) {
    std::memset(&pin[p], 0, sizeof(PoseIn));
    for (int k = 0; k < 16; ++k) pin[p].pose[k] = (double)xforms[16 * p + k];
    pin[p].cam = (int32_t)cam_for_proj[p];
    pin[p].vol = 0;
    pin[p].out_off = p * npx;
    pin[p].ncc_zf = nullptr;
    pin[p].step = step;
  }
  CamDev* cd = (CamDev*)(host + off_cam);
  for (int32_t c = 0; c < num_cams; ++c) cd[c] = make_cam_dev(cams[c]);
  std::memcpy(host + off_vol, &vol->desc, sizeof(VolDesc));
  std::memcpy(host + off_c4p, cam_for_proj, sizeof(int64_t) * num_projs);
  std::memset(host + off_cnt, 0, 16);
  XRH_HIP_CHECK(hipMemcpyAsync(dev, host, blob, hipMemcpyHostToDevice, ctx->stream));

  if (!samples_out) {
    // RayCasterCPU::pre_compute (xregRayCastBaseCPU.cpp:128-158): background images replace the
    // buffer's contents under either store method; the cast then accumulates onto them
    if (bg_dev) {
      const int64_t n = npx * num_projs;
      hipLaunchKernelGGL(copy_bg_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                         bg_dev, (const int64_t*)(dev + off_c4p), npx, num_projs, drr_dev);
      XRH_HIP_CHECK(hipGetLastError());
    }
  }
  DrrLaunch L;
  std::memset(&L, 0, sizeof(L));
  L.poses_dev = (const PoseIn*)(dev + off_pose);
  L.vols_dev = (const VolDesc*)(dev + off_vol);
  L.cams_dev = (const CamDev*)(dev + off_cam);
  L.items_dev = (DrrItem*)(dev + off_items);
  L.band_start_dev = (const uint32_t*)(dev + off_band);
  L.num_items = num_projs;
  L.bands = bands;
  L.max_band_tiles = max_band_tiles;
  L.geom = geom;
  L.kernel = kernel;
  L.store = (store == XRH_STORE_ACCUM || bg_dev) ? XRH_STORE_ACCUM : XRH_STORE_REPLACE;
  L.fill = default_bg;
  L.storage = vol->desc.storage;
  L.layout = vol->desc.layout;
  L.out = drr_dev;
  L.sample_counter = samples_out ? (uint64_t*)(dev + off_cnt) : nullptr;
  L.reuse = var.reuse;
  L.wave_shape = var.shape;
  launch_drr_setup(L, ctx->stream);
  if (igrid && !samples_out) {
    launch_drr_interp(L, *igrid, interp, aa_rays, aa_seed, ctx->stream);
  } else {
    launch_drr(L, ctx->stream);
  }
  if (samples_out) {
    XRH_HIP_CHECK(hipMemcpyAsync(samples_out, dev + off_cnt, 16, hipMemcpyDeviceToHost, ctx->stream));
    XRH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
}

// ---------------------------------------------------------------------------------------
// Fixed staging of xrh_line_int_noalloc. The device half is the file-scope array `scratch`
// (every GPU the library runs on holds its own copy until the process ends); host code reaches
// it only through the address hipGetSymbolAddress returns on that device, kept per device in
// g_scratch_dev at context creation. The host half is a static array. One mutex serialises
// the entry point within the process and each call waits for its work before releasing it, so
// two calls never use a device's scratch at once.

constexpr uint64_t na_align(uint64_t x) { return (x + 255) & ~uint64_t(255); }
// the packed tables for the largest call (the same offsets line_int_launch computes)
constexpr uint64_t kNoAllocHostBytes =
    na_align(na_align(na_align(na_align(na_align(sizeof(PoseIn) * XRH_NOALLOC_MAX_PROJS) +
                                        sizeof(CamDev) * XRH_NOALLOC_MAX_CAMS) +
                               sizeof(VolDesc)) +
                      sizeof(uint32_t) * kMaxDrrBands * (XRH_NOALLOC_MAX_PROJS + 1)) +
             sizeof(int64_t) * XRH_NOALLOC_MAX_PROJS) +
    na_align(16);
// ... followed on the device by the set-up kernel's per-projection items
constexpr uint64_t kNoAllocDevBytes = na_align(kNoAllocHostBytes + sizeof(DrrItem) * XRH_NOALLOC_MAX_PROJS);
static_assert(kNoAllocDevBytes <= (2u << 20), "fixed staging exceeds its stated 2 MiB bound");

static __device__ __align__(256) unsigned char scratch[kNoAllocDevBytes];

constexpr int kNoAllocMaxDevices = 64;
static unsigned char* g_scratch_dev[kNoAllocMaxDevices];
static unsigned char g_noalloc_host[kNoAllocHostBytes];
static std::mutex g_noalloc_mu;

// Called with `device` current (context creation): records the device address of `scratch`.
static void noalloc_init_device(int device)
{
  XRH_REQUIRE(device >= 0 && device < kNoAllocMaxDevices, "device index beyond the fixed staging table");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_scratch_dev[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    fail(XRH_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  }
  g_scratch_dev[device] = (unsigned char*)p;
}

}  // namespace xrh

using namespace xrh;

extern "C" {

const char* xrh_last_error(void) { return g_last_error.c_str(); }

int32_t xrh_abi_version(void) { return XRH_ABI_VERSION; }

xrh_status xrh_ctx_create(int32_t device, uint64_t hbm_budget_bytes, xrh_ctx** out)
{
  return guarded([&] {
    XRH_REQUIRE(out, "null output pointer");
    int count = 0;
    XRH_HIP_CHECK(hipGetDeviceCount(&count));
    XRH_REQUIRE(device >= 0 && device < count, "HIP device index out of range");
    XRH_HIP_CHECK(hipSetDevice(device));
    auto* ctx = new xrh_ctx();
    ctx->device = device;
    ctx->budget = hbm_budget_bytes;
    hipDeviceProp_t prop;
    XRH_HIP_CHECK(hipGetDeviceProperties(&prop, device));
    ctx->num_cus = prop.multiProcessorCount;
    try {
      noalloc_init_device(device);
    } catch (...) {
      delete ctx;
      throw;
    }
    const hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete ctx;
      fail(XRH_ERR_HIP, std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
    }
    *out = ctx;
  });
}

xrh_status xrh_ctx_destroy(xrh_ctx* ctx)
{
  return guarded([&] {
    if (!ctx) return;
    XRH_HIP_CHECK(hipSetDevice(ctx->device));
    XRH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (ctx->pinned) XRH_HIP_CHECK(hipHostFree(ctx->pinned));
    for (auto& kv : ctx->allocs) (void)hipFree(kv.first);
    ctx->allocs.clear();
    XRH_HIP_CHECK(hipStreamDestroy(ctx->stream));
    delete ctx;
  });
}

xrh_status xrh_ctx_synchronize(xrh_ctx* ctx)
{
  return guarded([&] {
    XRH_REQUIRE(ctx, "null context");
    XRH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

void* xrh_ctx_stream(xrh_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

uint64_t xrh_ctx_bytes_in_use(const xrh_ctx* ctx) { return ctx ? ctx->in_use : 0; }

xrh_status xrh_device_alloc(xrh_ctx* ctx, uint64_t bytes, void** dptr)
{
  return guarded([&] {
    XRH_REQUIRE(ctx && dptr, "null argument");
    *dptr = ctx_alloc(ctx, bytes);
  });
}

xrh_status xrh_device_free(xrh_ctx* ctx, void* dptr)
{
  return guarded([&] {
    XRH_REQUIRE(ctx, "null context");
    XRH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    ctx_free(ctx, dptr);
  });
}

xrh_status xrh_copy_h2d(xrh_ctx* ctx, void* dst_dev, const void* src_host, uint64_t bytes)
{
  return guarded([&] {
    XRH_REQUIRE(ctx && dst_dev && src_host, "null argument");
    ctx_check_range(ctx, dst_dev, bytes, "copy destination");
    XRH_HIP_CHECK(hipMemcpyAsync(dst_dev, src_host, bytes, hipMemcpyHostToDevice, ctx->stream));
    XRH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

xrh_status xrh_copy_d2h(xrh_ctx* ctx, void* dst_host, const void* src_dev, uint64_t bytes)
{
  return guarded([&] {
    XRH_REQUIRE(ctx && dst_host && src_dev, "null argument");
    ctx_check_range(ctx, src_dev, bytes, "copy source");
    XRH_HIP_CHECK(hipMemcpyAsync(dst_host, src_dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    XRH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

// Streaming copy: 16-byte vectors when aligned (n16 of them), the tail (or an unaligned
// copy) byte by byte. Grid-stride, so any grid size covers the buffer.
__global__ void xrh_copy_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n16,
                                unsigned char* __restrict__ dst_b, const unsigned char* __restrict__ src_b,
                                uint64_t tail_begin, uint64_t tail_end)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (uint64_t i = tid; i < n16; i += stride) dst[i] = src[i];
  for (uint64_t i = tail_begin + tid; i < tail_end; i += stride) dst_b[i] = src_b[i];
}

xrh_status xrh_copy_d2d(xrh_ctx* ctx, void* dst_dev, const void* src_dev, uint64_t bytes)
{
  return guarded([&] {
    XRH_REQUIRE(ctx && dst_dev && src_dev, "null argument");
    if (bytes == 0) return;
    ctx_check_range(ctx, dst_dev, bytes, "copy destination");
    ctx_check_range(ctx, src_dev, bytes, "copy source");
    const bool aligned = ((uintptr_t)dst_dev % 16 == 0) && ((uintptr_t)src_dev % 16 == 0);
    const uint64_t n16 = aligned ? bytes / 16 : 0;
    const uint64_t work = n16 + (bytes - 16 * n16);
    const uint64_t blocks = std::min<uint64_t>(std::max<uint64_t>((work + 255) / 256, 1), 8192);
    hipLaunchKernelGGL(xrh_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, (uint4*)dst_dev,
                       (const uint4*)src_dev, n16, (unsigned char*)dst_dev, (const unsigned char*)src_dev,
                       16 * n16, bytes);
    XRH_HIP_CHECK(hipGetLastError());
    XRH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

xrh_status xrh_volume_create(xrh_ctx* ctx, const float* vox, const int64_t size[3],
                             const float idx_to_phys[16], int32_t storage, xrh_volume** out)
{
  return guarded([&] { volume_create(ctx, vox, size, idx_to_phys, storage, -1, out); });
}

xrh_status xrh_volume_create_hu(xrh_ctx* ctx, const float* vox_hu, const int64_t size[3],
                                const float idx_to_phys[16], int32_t storage, int32_t layout, double hu_lower,
                                double mu_water, double mu_air, xrh_volume** out)
{
  return guarded([&] {
    XRH_REQUIRE(layout >= -1 && layout <= 5, "bad layout");
    const HuParams hu{hu_lower, mu_water, mu_air};
    volume_create(ctx, vox_hu, size, idx_to_phys, storage, layout, out, &hu);
  });
}

xrh_status xrh_volume_create_layout(xrh_ctx* ctx, const float* vox, const int64_t size[3],
                                    const float idx_to_phys[16], int32_t storage, int32_t layout,
                                    xrh_volume** out)
{
  return guarded([&] {
    XRH_REQUIRE(layout >= 0 && layout <= 5, "bad layout");
    volume_create(ctx, vox, size, idx_to_phys, storage, layout, out);
  });
}

xrh_status xrh_volume_destroy(xrh_volume* vol)
{
  return guarded([&] {
    if (!vol) return;
    XRH_HIP_CHECK(hipStreamSynchronize(vol->ctx->stream));
    ctx_free(vol->ctx, vol->data);
    if (vol->skip) ctx_free(vol->ctx, vol->skip);
    for (void* g : vol->grid_mem) {
      if (g) ctx_free(vol->ctx, g);
    }
    delete vol;
  });
}

uint64_t xrh_volume_bytes(const xrh_volume* vol)
{
  return vol ? vol->bytes + vol->grid_bytes[0] + vol->grid_bytes[1] : 0;
}

xrh_status xrh_volume_info(const xrh_volume* vol, int32_t* storage, int32_t* layout)
{
  return guarded([&] {
    XRH_REQUIRE(vol && storage && layout, "null argument");
    *storage = vol->desc.storage;
    *layout = vol->desc.layout;
  });
}

xrh_status xrh_line_int(xrh_ctx* ctx, const xrh_volume* vol, const xrh_camera* cams,
                        int32_t num_cams, const float* xforms_cam_to_phys,
                        const int64_t* cam_for_proj, int64_t num_projs, float step,
                        int32_t kernel, int32_t store, const float* bg_dev, float default_bg,
                        int32_t geom, float* drr_dev)
{
  return guarded([&] {
    line_int_launch(ctx, vol, cams, num_cams, xforms_cam_to_phys, cam_for_proj, num_projs, step,
                    kernel, store, bg_dev, default_bg, geom, drr_dev, nullptr);
  });
}

xrh_status xrh_line_int_noalloc(xrh_ctx* ctx, const xrh_volume* vol, const xrh_camera* cams,
                                int32_t num_cams, const float* xforms_cam_to_phys,
                                const int64_t* cam_for_proj, int64_t num_projs, float step,
                                int32_t kernel, int32_t store, const float* bg_dev, float default_bg,
                                int32_t geom, float* drr_dev)
{
  return guarded([&] {
    XRH_REQUIRE(ctx, "null context");
    XRH_REQUIRE(num_projs <= XRH_NOALLOC_MAX_PROJS,
                "xrh_line_int_noalloc: " + std::to_string(num_projs) + " projections exceed the fixed staging "
                "limit of " + std::to_string(XRH_NOALLOC_MAX_PROJS));
    XRH_REQUIRE(num_cams <= XRH_NOALLOC_MAX_CAMS,
                "xrh_line_int_noalloc: " + std::to_string(num_cams) + " cameras exceed the fixed staging "
                "limit of " + std::to_string(XRH_NOALLOC_MAX_CAMS));
    XRH_REQUIRE(ctx->device >= 0 && ctx->device < kNoAllocMaxDevices, "device index beyond the staging table");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    unsigned char* dev = g_scratch_dev[ctx->device];
    XRH_REQUIRE(dev, "fixed staging of this device was not set up");
    const FixedStaging st{g_noalloc_host, kNoAllocHostBytes, dev, kNoAllocDevBytes};
    line_int_launch(ctx, vol, cams, num_cams, xforms_cam_to_phys, cam_for_proj, num_projs, step, kernel, store,
                    bg_dev, default_bg, geom, drr_dev, nullptr, &st);
    // the staging is free for the next call (of any context on any device) once this work ends
    XRH_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  });
}

xrh_status xrh_line_int_interp(xrh_ctx* ctx, xrh_volume* vol, const xrh_camera* cams,
                               int32_t num_cams, const float* xforms_cam_to_phys,
                               const int64_t* cam_for_proj, int64_t num_projs, float step,
                               int32_t kernel, int32_t store, const float* bg_dev, float default_bg,
                               int32_t interp, int32_t aa_rays, uint64_t aa_seed, float* drr_dev)
{
  return guarded([&] {
    XRH_REQUIRE(interp >= XRH_INTERP_LINEAR && interp <= XRH_INTERP_BSPLINE, "bad interpolation method");
    XRH_REQUIRE(aa_rays >= 0 && aa_rays <= XRH_MAX_AA_RAYS, "anti-aliasing rays out of range");
    XRH_REQUIRE(ctx && vol, "null argument");
    XRH_REQUIRE(vol->ctx == ctx, "volume belongs to another context");
    // linear without anti-aliasing is the trilinear kernels' cast
    const bool plain = (interp == XRH_INTERP_LINEAR && aa_rays == 0);
    const InterpGrid* g = plain ? nullptr : &interp_prepare(vol, interp);
    line_int_launch(ctx, vol, cams, num_cams, xforms_cam_to_phys, cam_for_proj, num_projs, step, kernel, store,
                    bg_dev, default_bg, XRH_GEOM_FP32, drr_dev, nullptr, nullptr, g, interp, aa_rays, aa_seed);
  });
}

xrh_status xrh_line_int_sample_count(xrh_ctx* ctx, const xrh_volume* vol,
                                     const xrh_camera* cams, int32_t num_cams,
                                     const float* xforms_cam_to_phys,
                                     const int64_t* cam_for_proj, int64_t num_projs,
                                     float step, int32_t geom, uint64_t* samples_out,
                                     uint64_t* evaluated_out)
{
  return guarded([&] {
    XRH_REQUIRE(samples_out, "null output");
    uint64_t both[2] = {0, 0};
    line_int_launch(ctx, vol, cams, num_cams, xforms_cam_to_phys, cam_for_proj, num_projs, step,
                    XRH_LINE_INT_SUM, XRH_STORE_REPLACE, nullptr, 0.0f, geom, nullptr, both);
    *samples_out = both[0];
    if (evaluated_out) *evaluated_out = both[1];
  });
}

}  // extern "C"
