// xrh_ray.h -- per-ray pieces shared by the DRR kernels (xrh_drr.hip) and the
// interpolating casters (xrh_interp.hip): detector points, ray/box intersection and the
// per-ray set-up in float (the reference's arithmetic) and in double, the line kernels, and
// the band-major tile decode with its lane -> pixel footprints. Device code only.
#pragma once

#include "xrh_internal.h"

#include <hip/hip_fp16.h>

namespace xrh {
namespace {

// 16-bit storage bits -> float
__device__ inline float h2f(uint32_t bits16) { return __half2float(__ushort_as_half((unsigned short)bits16)); }
__device__ inline float bf2f(uint32_t bits16) { return __uint_as_float(bits16 << 16); }

constexpr int kTile = 16;          // detector tile edge (pixels)
constexpr int kThreads = kTile * kTile;

// ---------------------------------------------------------------------------------------
// float geometry in the reference's operation order, no contraction (XRH_GEOM_FP32)

__device__ inline float dot3(const float* r, const float v0, const float v1, const float v2)
{
#pragma clang fp contract(off)
  return (r[0] * v0 + r[1] * v1) + r[2] * v2;
}

__device__ inline void affine_pt(const float* m, const float p[3], float out[3])
{
#pragma clang fp contract(off)
  const float a = dot3(&m[0], p[0], p[1], p[2]) + m[3];
  const float b = dot3(&m[4], p[0], p[1], p[2]) + m[7];
  const float c = dot3(&m[8], p[0], p[1], p[2]) + m[11];
  out[0] = a;
  out[1] = b;
  out[2] = c;
}

__device__ inline void affine_pt_d(const double* m, const double p[3], double out[3])
{
  const double a = (m[0] * p[0] + m[1] * p[1]) + m[2] * p[2] + m[3];
  const double b = (m[4] * p[0] + m[5] * p[1]) + m[6] * p[2] + m[7];
  const double c = (m[8] * p[0] + m[9] * p[1]) + m[10] * p[2] + m[11];
  out[0] = a;
  out[1] = b;
  out[2] = c;
}

// CameraModel::ind_pt_to_phys_det_pt (xregPerspectiveXform.cpp:401-414)
__device__ inline void ind_to_det(const CamDev& c, float col, float row, float out[3])
{
#pragma clang fp contract(off)
  const float q0 = c.det_z * col;
  const float q1 = c.det_z * row;
  const float q2 = c.det_z * 1.0f;
  float r[3];
  r[0] = dot3(&c.kinv[0], q0, q1, q2) + 0.0f;
  r[1] = dot3(&c.kinv[3], q0, q1, q2) + 0.0f;
  r[2] = dot3(&c.kinv[6], q0, q1, q2) + c.z_adj;
  affine_pt(c.einv, r, out);
}

__device__ inline void ind_to_det_d(const CamDev& c, double col, double row, double out[3])
{
  const double q0 = c.det_z_d * col, q1 = c.det_z_d * row, q2 = c.det_z_d;
  double r[3];
  for (int i = 0; i < 3; ++i) {
    r[i] = (c.kinv_d[3 * i] * q0 + c.kinv_d[3 * i + 1] * q1) + c.kinv_d[3 * i + 2] * q2;
  }
  r[2] += c.z_adj_d;
  affine_pt_d(c.einv_d, r, out);
}

// IEEE-correctly rounded float division and square root, obtained from the double result:
// for these operations rounding a 53-bit result to 24 bits equals rounding the exact value
// (53 >= 2*24 + 2), so the FP32 ray set-up matches the host's SSE arithmetic bit for bit.
__device__ inline float div_rn(float a, float b) { return (float)((double)a / (double)b); }
__device__ inline double div_rn(double a, double b) { return a / b; }
__device__ inline float sqrt_rn(float x) { return (float)sqrt((double)x); }

// RayRectIntersect with limit_to_segment (lib/spatial/xregSpatialPrimitives.cpp:175-222)
template <class T>
__device__ inline bool ray_rect(const T mx[3], const T p0[3], const T d[3], T& t0, T& t1)
{
#pragma clang fp contract(off)
  T ts = T(0), te = T(1);
  for (int k = 0; k < 3; ++k) {
    const T dk = d[k];
    // the reference compares the float component against the double literal 1.0e-8
    if ((double)fabs(dk) > 1.0e-8) {
      const T inv = div_rn(T(1), dk);
      T a = (T(0) - p0[k]) * inv;
      T b = (mx[k] - p0[k]) * inv;
      if (b < a) {
        const T t = a;
        a = b;
        b = t;
      }
      ts = (ts > a) ? ts : a;
      te = (b < te) ? b : te;
      if (ts > te) {
        t0 = ts;
        t1 = te;
        return false;
      }
    } else if ((p0[k] < T(0)) || (p0[k] > mx[k])) {
      t0 = ts;
      t1 = te;
      return false;
    }
  }
  t0 = ts;
  t1 = te;
  return true;
}

struct Ray {
  float sx, sy, sz;  // first sample (continuous index; FP64 set-up: padded-grid index, +1 per axis)
  float dx, dy, dz;  // step vector (index units)
  int n;             // last sample index (n+1 samples), -1 = ray misses
};

// Per-ray set-up in float, reference order (xregRayCastLineIntCPU.cpp:204-268).
// (col, row) is the detector index (a pixel's integer index, or a jittered one).
__device__ inline Ray setup_ray_f32(const DrrItem& it, float col, float row)
{
#pragma clang fp contract(off)
  Ray ray;
  ray.n = -1;
  float det[3];
  ind_to_det(it.cam, col, row, det);
  float ph[3], dp[3];
  affine_pt(it.M, it.cam.pinhole, ph);
  affine_pt(it.M, det, dp);
  const float d[3] = {dp[0] - ph[0], dp[1] - ph[1], dp[2] - ph[2]};
  const float mx[3] = {(float)(it.X - 1), (float)(it.Y - 1), (float)(it.Z - 1)};
  float t0, t1;
  const bool hit = ray_rect<float>(mx, ph, d, t0, t1);
  const float kTol = 1.0e-3f;
  if (hit && ((t1 - t0) > 2.0f * kTol)) {
    t0 = t0 + kTol;
    t1 = t1 - kTol;
    ray.sx = ph[0] + t0 * d[0];
    ray.sy = ph[1] + t0 * d[1];
    ray.sz = ph[2] + t0 * d[2];
    const float dlen = sqrt_rn((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    const float L = (t1 - t0) * dlen;
    float v0 = det[0] - it.cam.pinhole[0];
    float v1 = det[1] - it.cam.pinhole[1];
    float v2 = det[2] - it.cam.pinhole[2];
    const float vz = (v0 * v0 + v1 * v1) + v2 * v2;
    if (vz > 0.0f) {
      const float vn = sqrt_rn(vz);
      v0 = div_rn(v0, vn);
      v1 = div_rn(v1, vn);
      v2 = div_rn(v2, vn);
    }
    v0 = v0 * it.step;
    v1 = v1 * it.step;
    v2 = v2 * it.step;
    const float s0 = dot3(&it.M[0], v0, v1, v2);
    const float s1 = dot3(&it.M[4], v0, v1, v2);
    const float s2 = dot3(&it.M[8], v0, v1, v2);
    const float s = sqrt_rn((s0 * s0 + s1 * s1) + s2 * s2);
    ray.n = (int)div_rn(L, s);
    const float scale = div_rn(s, dlen);
    ray.dx = d[0] * scale;
    ray.dy = d[1] * scale;
    ray.dz = d[2] * scale;
  }
  return ray;
}

// Per-ray set-up in double; samples later placed at start + i*step in float.
__device__ inline Ray setup_ray_f64(const DrrItem& it, int col, int row)
{
  Ray ray;
  ray.n = -1;
  double det[3];
  ind_to_det_d(it.cam, (double)col, (double)row, det);
  double ph[3], dp[3];
  affine_pt_d(it.Md, it.cam.pinhole_d, ph);
  affine_pt_d(it.Md, det, dp);
  const double d[3] = {dp[0] - ph[0], dp[1] - ph[1], dp[2] - ph[2]};
  const double mx[3] = {(double)(it.X - 1), (double)(it.Y - 1), (double)(it.Z - 1)};
  double t0, t1;
  const bool hit = ray_rect<double>(mx, ph, d, t0, t1);
  if (hit && ((t1 - t0) > 2.0e-3)) {
    t0 += 1.0e-3;
    t1 -= 1.0e-3;
    const double dlen = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double L = (t1 - t0) * dlen;
    double v0 = det[0] - it.cam.pinhole_d[0];
    double v1 = det[1] - it.cam.pinhole_d[1];
    double v2 = det[2] - it.cam.pinhole_d[2];
    const double vn = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    if (vn > 0.0) {
      v0 /= vn;
      v1 /= vn;
      v2 /= vn;
    }
    const double st = (double)it.step;
    v0 *= st;
    v1 *= st;
    v2 *= st;
    const double s0 = it.Md[0] * v0 + it.Md[1] * v1 + it.Md[2] * v2;
    const double s1 = it.Md[4] * v0 + it.Md[5] * v1 + it.Md[6] * v2;
    const double s2 = it.Md[8] * v0 + it.Md[9] * v1 + it.Md[10] * v2;
    const double s = sqrt(s0 * s0 + s1 * s1 + s2 * s2);
    ray.n = (int)(L / s);
    const double scale = s / dlen;
    // start in the padded grid's index frame (+1 per axis, the apron), rounded once
    ray.sx = (float)((ph[0] + t0 * d[0]) + 1.0);
    ray.sy = (float)((ph[1] + t0 * d[1]) + 1.0);
    ray.sz = (float)((ph[2] + t0 * d[2]) + 1.0);
    ray.dx = (float)(d[0] * scale);
    ray.dy = (float)(d[1] * scale);
    ray.dz = (float)(d[2] * scale);
  }
  return ray;
}

// AccumLineIntKernel / MaxLineIntKernel (xregRayCastLineIntCPU.cpp:45-75)
template <int kOp>
__device__ inline float line_op(float acc, float s)
{
  if constexpr (kOp == XRH_LINE_INT_SUM) return acc + s;
  else return (acc < s) ? s : acc;
}

// ---------------------------------------------------------------------------------------
// tile decode: blockIdx -> (item, tile_x, tile_y) through the band plan.

struct TileRef {
  int64_t item;
  int tx, ty;
  bool valid;
};

__device__ inline TileRef decode_tile(const uint32_t* band_start, int64_t num_items, int bands,
                                      const DrrItem* items, int stride = 1)
{
  TileRef t;
  const uint32_t b = blockIdx.x;
  const int band = (int)(b % (uint32_t)bands);
  const uint32_t j = b / (uint32_t)bands;
  const uint32_t* bs = band_start + (int64_t)band * (num_items + 1);
  t.valid = j < bs[num_items];
  if (!t.valid) return t;
  // largest item with bs[item] <= j (uniform binary search on scalar registers)
  int64_t lo = 0, hi = num_items - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (bs[mid] <= j) lo = mid; else hi = mid - 1;
  }
  t.item = lo;
  const uint32_t local = j - bs[lo];
  const int TX = items[lo * stride].tiles_x, TY = items[lo * stride].tiles_y;
  const int row0 = (int)(((int64_t)band * TY) / bands);
  t.ty = row0 + (int)(local / (uint32_t)TX);
  t.tx = (int)(local % (uint32_t)TX);
  return t;
}

// The pixel of this lane in tile (tx, ty). shape (uniform) picks each wave's footprint in
// the 16x16 tile: 0 = 8x8 quadrants, 1 = 16x4 row strips, 2 = 4x16 column strips; every
// shape covers the tile's 256 pixels once.
__device__ inline void lane_pixel(int tx, int ty, int shape, int& col, int& row)
{
  const int t = threadIdx.x;
  const int w = t >> 6, l = t & 63;
  if (shape == 1) {
    col = tx * kTile + (l & 15);
    row = ty * kTile + w * 4 + (l >> 4);
  } else if (shape == 2) {
    col = tx * kTile + w * 4 + (l & 3);
    row = ty * kTile + (l >> 2);
  } else {
    col = tx * kTile + (w & 1) * 8 + (l & 7);
    row = ty * kTile + (w >> 1) * 8 + (l >> 3);
  }
}

}  // namespace
}  // namespace xrh
