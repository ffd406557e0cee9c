// xrh_drr.hip -- line-integral DRR kernels for gfx950.
//
// Semantics follow RayCasterLineIntCPU::compute / ComputeLineInts
// (lib/ray_cast/xregRayCastLineIntCPU.cpp:105-349):
//   per ray: detector point (CameraModel::ind_pt_to_phys_det_pt), pinhole and direction in
//   volume index space, slab intersection against [0, N-1] limited to the source->detector
//   segment, 1e-3 parametric nudge, n = floor(chord / step_len), n+1 trilinear samples
//   accumulated with SUM or MAX, times the physical step; stored with the same operator.
//
// Mapping onto CDNA4:
//   - one lane = one ray; a 256-lane workgroup owns a 16x16 detector tile and each of its
//     four wave64s an 8x8 sub-tile, so neighbouring lanes march neighbouring pixels and
//     their gathers hit the same cache lines;
//   - workgroups are ordered band-major and swizzled so that the workgroups that the
//     dispatcher sends to one XCD (blockIdx mod 8) all work on the same band of detector
//     rows across every projection: each XCD's L2 then holds one slab of the volume;
//   - the volume sits in HBM in an edge-padded layout read through buffer loads
//     (hardware range check: a stray offset returns 0, it can never fault); the quad (f32)
//     and oct (f16/bf16) layouts pre-pack interpolation corners so a sample costs 2 or 1
//     16-byte loads instead of 8 scalar gathers;
//   - trilinear weights and lerps are FP32 in software: HIP's image/texture API is not
//     available for gfx950 (hipcc rejects tex3D for this target), so there is no hardware
//     filtering to fall back on.
#include "xrh_internal.h"
#include "xrh_ray.h"

#include <hip/hip_fp16.h>

#include <climits>
#include <cstdlib>
#include <cstring>

namespace xrh {
namespace {

// ---------------------------------------------------------------------------------------
// trilinear sampling per layout. Formats: 0 f32 linear, 1 f32 quad, 2 f16 linear, 3 f16 oct,
// 4 bf16 linear, 5 bf16 oct.

struct VolView {
  __amdgpu_buffer_rsrc_t rsrc;
  uint32_t py;     // row pitch (elements)
  uint32_t pz;     // slice pitch (elements)
  uint32_t apron;  // 1 + py + pz: element index of voxel (0, 0, 0)
  // last cell fetched (coefficient layout with cell reuse): its index and 32 bytes
  uint32_t ce;
  uint4 ca, cb;
};

// a (1 - w) + b w as two fused multiply-adds: as cheap as a + (b - a) w but without its
// cancellation when the result is small next to a (tissue -> air edges), so the FP32 sample
// stays within about an ulp of ITK's double-precision lerp.
__device__ inline float lerp(float a, float b, float w) { return fmaf(w, b, fmaf(-w, a, a)); }

__device__ inline float tri(float v000, float v100, float v010, float v110, float v001,
                            float v101, float v011, float v111, float wx, float wy, float wz)
{
  // x, then y, then z -- the order of ITK's LinearInterpolateImageFunction
  const float x00 = lerp(v000, v100, wx);
  const float x10 = lerp(v010, v110, wx);
  const float x01 = lerp(v001, v101, wx);
  const float x11 = lerp(v011, v111, wx);
  const float xx0 = lerp(x00, x10, wy);
  const float xx1 = lerp(x01, x11, wy);
  return lerp(xx0, xx1, wz);
}


// floor(x) converted to int in one instruction (V_CVT_FLR_I32_F32). Inline asm counts as
// convergent, so the sample loops below are unrolled by hand.
__device__ inline int cvt_floor(float x)
{
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// a * b + c for 24-bit unsigned a, b as one V_MAD_U32_U24 (the compiler otherwise forms two
// multiplies and a three-way add)
__device__ inline uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c)
{
  uint32_t r;
  asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// the same for signed 24-bit a and c (V_MAD_I32_I24)
__device__ inline int32_t mad_i24(int32_t a, int32_t b, int32_t c)
{
  int32_t r;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// kPadded: the position is already in the padded grid's index frame (FP64 ray set-up), so the
// element index needs no apron offset: two 24-bit multiply-adds.
template <int kFmt, bool kPadded>
__device__ inline float sample(VolView& v, float x, float y, float z)
{
  // x - floor(x) (V_FRACT_F32) and floor(x) as int: two instructions per axis
  const float wx = __builtin_amdgcn_fractf(x), wy = __builtin_amdgcn_fractf(y), wz = __builtin_amdgcn_fractf(z);
  const int ix = cvt_floor(x), iy = cvt_floor(y), iz = cvt_floor(z);
  // element index of the base corner; ix, iy, iz >= -1 and the pitches < 2^23, so the
  // signed 24-bit multiply-adds are exact; v.apron = 1 + py + pz shifts past the apron
  uint32_t e;
  if constexpr (kFmt == 12 || kFmt == 13) {
    // bricked x-faces: padded cell (ix+1, iy+1, iz+1), brick (./4) in brick order with the
    // brick pitches in py / pz, 5x4x4 faces per brick (x fastest, the last x-face repeated)
    const uint32_t ax = (uint32_t)(kPadded ? ix : ix + 1), ay = (uint32_t)(kPadded ? iy : iy + 1),
                   az = (uint32_t)(kPadded ? iz : iz + 1);
    const uint32_t brick = __umul24(az >> 2, v.pz) + __umul24(ay >> 2, v.py) + (ax >> 2);
    const uint32_t local = __umul24(az & 3u, 20u) + __umul24(ay & 3u, 5u) + (ax & 3u);
    e = __umul24(brick, (uint32_t)kBrickFaces) + local;
  } else if constexpr (kPadded) {
    // padded indices are >= 0 for every position within a voxel of the index box
    e = mad_u24((uint32_t)iz, v.pz, mad_u24((uint32_t)iy, v.py, (uint32_t)ix));
  } else {
    e = (uint32_t)mad_i24(iz, (int32_t)v.pz, mad_i24(iy, (int32_t)v.py, ix)) + v.apron;
  }
  if constexpr (kFmt == 0) {
    const int o = (int)(e * 4u);
    const int oy = (int)(v.py * 4u);
    const int sz = (int)(v.pz * 4u);
    const float v000 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(v.rsrc, o, 0, 0));
    const float v100 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(v.rsrc, o + 4, 0, 0));
    const float v010 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(v.rsrc, o + oy, 0, 0));
    const float v110 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(v.rsrc, o + oy + 4, 0, 0));
    const float v001 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(v.rsrc, o, sz, 0));
    const float v101 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(v.rsrc, o + 4, sz, 0));
    const float v011 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(v.rsrc, o + oy, sz, 0));
    const float v111 = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(v.rsrc, o + oy + 4, sz, 0));
    return tri(v000, v100, v010, v110, v001, v101, v011, v111, wx, wy, wz);
  } else if constexpr (kFmt == 1) {
    const int o = (int)(e * 16u);
    const int sz = (int)(v.pz * 16u);
    const auto q0 = __builtin_amdgcn_raw_buffer_load_b128(v.rsrc, o, 0, 0);
    const auto q1 = __builtin_amdgcn_raw_buffer_load_b128(v.rsrc, o, sz, 0);
    return tri(__uint_as_float(q0[0]), __uint_as_float(q0[1]), __uint_as_float(q0[2]),
               __uint_as_float(q0[3]), __uint_as_float(q1[0]), __uint_as_float(q1[1]),
               __uint_as_float(q1[2]), __uint_as_float(q1[3]), wx, wy, wz);
  } else if constexpr (kFmt == 2 || kFmt == 4) {
    const int o = (int)(e * 2u);
    const int oy = (int)(v.py * 2u);
    const int sz = (int)(v.pz * 2u);
    uint32_t r[8];
    r[0] = __builtin_amdgcn_raw_buffer_load_b16(v.rsrc, o, 0, 0);
    r[1] = __builtin_amdgcn_raw_buffer_load_b16(v.rsrc, o + 2, 0, 0);
    r[2] = __builtin_amdgcn_raw_buffer_load_b16(v.rsrc, o + oy, 0, 0);
    r[3] = __builtin_amdgcn_raw_buffer_load_b16(v.rsrc, o + oy + 2, 0, 0);
    r[4] = __builtin_amdgcn_raw_buffer_load_b16(v.rsrc, o, sz, 0);
    r[5] = __builtin_amdgcn_raw_buffer_load_b16(v.rsrc, o + 2, sz, 0);
    r[6] = __builtin_amdgcn_raw_buffer_load_b16(v.rsrc, o + oy, sz, 0);
    r[7] = __builtin_amdgcn_raw_buffer_load_b16(v.rsrc, o + oy + 2, sz, 0);
    float f[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) f[i] = (kFmt == 2) ? h2f(r[i]) : bf2f(r[i]);
    return tri(f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], wx, wy, wz);
  } else if constexpr (kFmt == 10 || kFmt == 11 || kFmt == 12 || kFmt == 13) {
    // face layout: faces normal to x at x and x+1 (adjacent 16-byte elements), bilinear in y, z
    // then linear along x. 11 and 13 keep the faces while the sample stays in its cell; 12 and
    // 13 read the bricked arrangement.
    uint4 f0, f1;
    if constexpr (kFmt == 10 || kFmt == 12) {
      const auto a = __builtin_amdgcn_raw_buffer_load_b128(v.rsrc, (int)(e * 16u), 0, 0);
      const auto b = __builtin_amdgcn_raw_buffer_load_b128(v.rsrc, (int)(e * 16u + 16u), 0, 0);
      f0 = make_uint4(a[0], a[1], a[2], a[3]);
      f1 = make_uint4(b[0], b[1], b[2], b[3]);
    } else {
      if (e != v.ce) {
        const auto a = __builtin_amdgcn_raw_buffer_load_b128(v.rsrc, (int)(e * 16u), 0, 0);
        const auto b = __builtin_amdgcn_raw_buffer_load_b128(v.rsrc, (int)(e * 16u + 16u), 0, 0);
        v.ca = make_uint4(a[0], a[1], a[2], a[3]);
        v.cb = make_uint4(b[0], b[1], b[2], b[3]);
        v.ce = e;
      }
      f0 = v.ca;
      f1 = v.cb;
    }
    const float p0 = fmaf(wz, fmaf(wy, __uint_as_float(f0.w), __uint_as_float(f0.z)),
                          fmaf(wy, __uint_as_float(f0.y), __uint_as_float(f0.x)));
    const float p1 = fmaf(wz, fmaf(wy, __uint_as_float(f1.w), __uint_as_float(f1.z)),
                          fmaf(wy, __uint_as_float(f1.y), __uint_as_float(f1.x)));
    return lerp(p0, p1, wx);
  } else if constexpr (kFmt == 7) {
    // coefficient layout, fetched only when the sample leaves the previous sample's cell
    // (about half the samples at half-voxel steps); lanes that stay keep their registers
    if (e != v.ce) {
      const uint32_t ob = e * 32u;
      const auto a = __builtin_amdgcn_raw_buffer_load_b128(v.rsrc, (int)ob, 0, 0);
      const auto b = __builtin_amdgcn_raw_buffer_load_b128(v.rsrc, (int)(ob + 16u), 0, 0);
      v.ca = make_uint4(a[0], a[1], a[2], a[3]);
      v.cb = make_uint4(b[0], b[1], b[2], b[3]);
      v.ce = e;
    }
    const float p0 = fmaf(wy, fmaf(wx, __uint_as_float(v.ca.w), __uint_as_float(v.ca.z)),
                          fmaf(wx, __uint_as_float(v.ca.y), __uint_as_float(v.ca.x)));
    const float p1 = fmaf(wy, fmaf(wx, __uint_as_float(v.cb.w), __uint_as_float(v.cb.z)),
                          fmaf(wx, __uint_as_float(v.cb.y), __uint_as_float(v.cb.x)));
    return fmaf(wz, p1, p0);
  } else if constexpr (kFmt == 6) {
    // trilinear polynomial coefficients of the cell: 2 adjacent 16-byte loads, 7 FMAs
    const uint32_t ob = e * 32u;
    const auto a = __builtin_amdgcn_raw_buffer_load_b128(v.rsrc, (int)ob, 0, 0);
    const auto b = __builtin_amdgcn_raw_buffer_load_b128(v.rsrc, (int)(ob + 16u), 0, 0);
    const float p0 = fmaf(wy, fmaf(wx, __uint_as_float(a[3]), __uint_as_float(a[2])),
                          fmaf(wx, __uint_as_float(a[1]), __uint_as_float(a[0])));
    const float p1 = fmaf(wy, fmaf(wx, __uint_as_float(b[3]), __uint_as_float(b[2])),
                          fmaf(wx, __uint_as_float(b[1]), __uint_as_float(b[0])));
    return fmaf(wz, p1, p0);
  } else {
    const int o = (int)(e * 16u);
    const auto q = __builtin_amdgcn_raw_buffer_load_b128(v.rsrc, o, 0, 0);
    float f[8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t lo = q[i] & 0xffffu, hi = q[i] >> 16;
      f[2 * i] = (kFmt == 3) ? h2f(lo) : bf2f(lo);
      f[2 * i + 1] = (kFmt == 3) ? h2f(hi) : bf2f(hi);
    }
    return tri(f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], wx, wy, wz);
  }
}

// Empty-space skipping. Walks from sample i towards sample `end` (dir = +1 or -1) over samples
// that lie in bricks whose one-voxel-dilated trilinear support holds only zero voxels, so
// their interpolated value is exactly 0 and leaving them out of a SUM changes nothing. At a
// brick at Chebyshev distance d >= 1 from the nearest occupied brick, the cube of bricks
// within d-1 is empty: every sample up to the cube's exit face is skipped and the walk
// continues with the first sample past it (at least one sample per iteration, so at most
// |end - i| + 1 iterations). The walk runs in brick units (positions scaled by 2^-shift,
// exact) with branch-free exit faces: b + d moving up an axis, b + 1 - d moving down, none
// along an axis the ray does not move on. Its positions, start + i*step in float, differ from
// the sample loop's by far less than the one-voxel dilation. Returns the first sample (in walk
// order) that may be non-zero.
__device__ inline int skip_empty(const DrrItem& it, const Ray& r, int i, int end, int dir, float org)
{
  const float B = (float)(1 << it.skip_shift), invB = __builtin_ldexpf(1.0f, -it.skip_shift);
  // ray in brick units
  const float bx0 = (r.sx - org) * invB, by0 = (r.sy - org) * invB, bz0 = (r.sz - org) * invB;
  const float dbx = r.dx * invB, dby = r.dy * invB, dbz = r.dz * invB;
  // per axis: exit face b + o + s (d - 1), samples to reach it (face - p) * k (+ h: infinite
  // when the walk does not move along the axis)
  const float vx = dir > 0 ? r.dx : -r.dx, vy = dir > 0 ? r.dy : -r.dy, vz = dir > 0 ? r.dz : -r.dz;
  const float sx = vx > 0.0f ? 1.0f : -1.0f, sy = vy > 0.0f ? 1.0f : -1.0f, sz = vz > 0.0f ? 1.0f : -1.0f;
  const float ox = vx > 0.0f ? 1.0f : 0.0f, oy = vy > 0.0f ? 1.0f : 0.0f, oz = vz > 0.0f ? 1.0f : 0.0f;
  const float kx = vx != 0.0f ? B * __builtin_amdgcn_rcpf(vx) : 0.0f;
  const float ky = vy != 0.0f ? B * __builtin_amdgcn_rcpf(vy) : 0.0f;
  const float kz = vz != 0.0f ? B * __builtin_amdgcn_rcpf(vz) : 0.0f;
  const float hx = vx != 0.0f ? 0.0f : 3.0e38f, hy = vy != 0.0f ? 0.0f : 3.0e38f, hz = vz != 0.0f ? 0.0f : 3.0e38f;
  const float gx1 = (float)(it.skip_gx - 1), gy1 = (float)(it.skip_gy - 1), gz1 = (float)(it.skip_gz - 1);
  // the map is below 2^24 bytes and gy * gz below 2^24 (checked when it is built), so the
  // 24-bit multiply-adds below are exact and a stray index reads 0 from the buffer range check
  const auto map = __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(it.skip), (short)0,
                                                     it.skip_gx * it.skip_gy * it.skip_gz, 0x00020000);
  while (dir > 0 ? (i <= end) : (i >= end)) {
    const float fi = (float)i;
    const float px = fmaf(fi, dbx, bx0), py = fmaf(fi, dby, by0), pz = fmaf(fi, dbz, bz0);
    const float fx = __builtin_amdgcn_fmed3f(floorf(px), 0.0f, gx1);
    const float fy = __builtin_amdgcn_fmed3f(floorf(py), 0.0f, gy1);
    const float fz = __builtin_amdgcn_fmed3f(floorf(pz), 0.0f, gz1);
    const uint32_t e = mad_u24(mad_u24((uint32_t)fz, (uint32_t)it.skip_gy, (uint32_t)fy), (uint32_t)it.skip_gx,
                               (uint32_t)fx);
    const int d = __builtin_amdgcn_raw_buffer_load_b8(map, (int)e, 0, 0);
    if (d == 0) break;
    const float dm1 = (float)(d - 1);
    const float tx = fmaf((fmaf(sx, dm1, ox) + fx) - px, kx, hx);
    const float ty = fmaf((fmaf(sy, dm1, oy) + fy) - py, ky, hy);
    const float tz = fmaf((fmaf(sz, dm1, oz) + fz) - pz, kz, hz);
    const float s = fminf(fminf(tx, ty), tz);
    const int jump = (s >= 1.0f) ? (int)fminf(s, 1.0e9f) + 1 : 1;
    i += dir * jump;
  }
  return i;
}

// Sample range [first, last] that can hold non-zero values (SUM with an empty-space map), or
// the whole ray [0, n]. FP64 set-up rays start in the padded frame (org 1).
template <int kOp, int kGeom>
__device__ inline void sample_range(const DrrItem& it, const Ray& r, int& first, int& last)
{
  first = 0;
  last = r.n;
  if (kOp == XRH_LINE_INT_SUM && it.skip) {
    const float org = (kGeom == XRH_GEOM_FP64) ? 1.0f : 0.0f;
    first = skip_empty(it, r, 0, r.n, 1, org);
    last = (first <= r.n) ? skip_empty(it, r, r.n, first, -1, org) : first - 1;
  }
}

__device__ inline double wave_sum_d(double v)
{
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ---------------------------------------------------------------------------------------
// kernels

__global__ void drr_setup_kernel(const PoseIn* __restrict__ poses, const VolDesc* __restrict__ vols,
                                 const CamDev* __restrict__ cams, DrrItem* __restrict__ items,
                                 int64_t n)
{
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PoseIn& p = poses[i];
  const VolDesc& v = vols[p.vol];
  DrrItem it;
  // phys->idx * cam->phys: Affine*Affine product (xregRayCastLineIntCPU.cpp:207-208), float
  float T[12];
  for (int k = 0; k < 12; ++k) T[k] = (float)p.pose[k];
  for (int r = 0; r < 3; ++r) {
    const float* a = &v.p2i[4 * r];
    for (int c = 0; c < 3; ++c) it.M[4 * r + c] = (a[0] * T[c] + a[1] * T[4 + c]) + a[2] * T[8 + c];
    it.M[4 * r + 3] = ((a[0] * T[3] + a[1] * T[7]) + a[2] * T[11]) + a[3];
  }
  // the same in double from the FP64 pose
  for (int r = 0; r < 3; ++r) {
    const double* a = &v.p2i_d[4 * r];
    for (int c = 0; c < 3; ++c) {
      it.Md[4 * r + c] = a[0] * p.pose[c] + a[1] * p.pose[4 + c] + a[2] * p.pose[8 + c];
    }
    it.Md[4 * r + 3] = a[0] * p.pose[3] + a[1] * p.pose[7] + a[2] * p.pose[11] + a[3];
  }
  it.cam = cams[p.cam];
  it.vol = v.data;
  it.vol_bytes = v.bytes;
  it.X = v.X;
  it.Y = v.Y;
  it.Z = v.Z;
  it.pitch_y = v.pitch_y;
  it.pitch_z = v.pitch_z;
  it.step = p.step;
  it.out_off = p.out_off;
  it.ncc_zf = p.ncc_zf;
  it.bg = p.bg;
  it.tiles_x = (it.cam.cols + kTile - 1) / kTile;
  it.tiles_y = (it.cam.rows + kTile - 1) / kTile;
  it.skip = v.skip;
  it.skip_gx = v.skip_gx;
  it.skip_gy = v.skip_gy;
  it.skip_gz = v.skip_gz;
  it.skip_shift = v.skip_shift;
  items[i] = it;
}

// Samples [i, last] of one ray into acc, in ray order.
template <int kGeom, int kOp, int kFmt>
__device__ inline float march(const Ray& ray, VolView& vv, int i, int last, float acc)
{
  if constexpr (kGeom == XRH_GEOM_FP32) {
    // positions advance by repeated float addition, as the reference does (:276),
    // including over the leading samples that are skipped
    float px = ray.sx, py = ray.sy, pz = ray.sz;
    if (i <= last) {
      for (int k = 0; k < i; ++k) {
        px = px + ray.dx;
        py = py + ray.dy;
        pz = pz + ray.dz;
      }
    }
    if constexpr (kFmt == 7 || kFmt == 11 || kFmt == 13) {
      // one sample per iteration: the cell cache lives in fixed registers; the count runs
      // in float (exact below 2^24) so the loop test is one float compare
      const float flast = (float)last;
      for (float fi = (float)i; fi <= flast; fi = fi + 1.0f) {
        acc = line_op<kOp>(acc, sample<kFmt, false>(vv, px, py, pz));
        px = px + ray.dx;
        py = py + ray.dy;
        pz = pz + ray.dz;
      }
    } else {
      for (; i < last; i += 2) {
        const float qx = px + ray.dx, qy = py + ray.dy, qz = pz + ray.dz;
        const float s0 = sample<kFmt, false>(vv, px, py, pz);
        const float s1 = sample<kFmt, false>(vv, qx, qy, qz);
        acc = line_op<kOp>(line_op<kOp>(acc, s0), s1);
        px = qx + ray.dx;
        py = qy + ray.dy;
        pz = qz + ray.dz;
      }
      if (i == last) acc = line_op<kOp>(acc, sample<kFmt, false>(vv, px, py, pz));
    }
  } else if constexpr (kFmt == 7 || kFmt == 11 || kFmt == 13) {
    const float flast = (float)last;
    for (float f0 = (float)i; f0 <= flast; f0 = f0 + 1.0f) {
      acc = line_op<kOp>(acc, sample<kFmt, true>(vv, fmaf(f0, ray.dx, ray.sx), fmaf(f0, ray.dy, ray.sy),
                                              fmaf(f0, ray.dz, ray.sz)));
    }
  } else {
    // sample indices as floats (exact: n < 2^24), advanced by float additions
    float f0 = (float)i;
    for (; i < last; i += 2) {
      const float f1 = f0 + 1.0f;
      const float s0 = sample<kFmt, true>(vv, fmaf(f0, ray.dx, ray.sx), fmaf(f0, ray.dy, ray.sy),
                                    fmaf(f0, ray.dz, ray.sz));
      const float s1 = sample<kFmt, true>(vv, fmaf(f1, ray.dx, ray.sx), fmaf(f1, ray.dy, ray.sy),
                                    fmaf(f1, ray.dz, ray.sz));
      acc = line_op<kOp>(line_op<kOp>(acc, s0), s1);
      f0 = f0 + 2.0f;
    }
    if (i == last) {
      const float f0 = (float)i;
      acc = line_op<kOp>(acc, sample<kFmt, true>(vv, fmaf(f0, ray.dx, ray.sx), fmaf(f0, ray.dy, ray.sy),
                                              fmaf(f0, ray.dz, ray.sz)));
    }
  }
  return acc;
}

// The fused NCC epilogue: FP64 sums of m, m^2 and m * zero-mean fixed over the tile, one
// record per tile. Every thread of the block calls it (it synchronises).
__device__ inline void ncc_epilogue(const DrrItem& it, const TileRef& tr, bool inside, int col, int row, float pix,
                                    double* __restrict__ ncc_partials, int ncc_tiles)
{
  if (ncc_partials && it.ncc_zf) {
    // NCC epilogue: sums of m, m^2 and m * zero-mean fixed, FP64, one record per tile
    double s1 = 0.0, s2 = 0.0, sf = 0.0;
    if (inside) {
      const double m = (double)pix;
      const double f = (double)it.ncc_zf[(int64_t)row * it.cam.cols + col];
      s1 = m;
      s2 = m * m;
      sf = m * f;
    }
    s1 = wave_sum_d(s1);
    s2 = wave_sum_d(s2);
    sf = wave_sum_d(sf);
    __shared__ double red[4][3];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      red[w][0] = s1;
      red[w][1] = s2;
      red[w][2] = sf;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double* dst = ncc_partials + ((int64_t)tr.item * ncc_tiles + (int64_t)tr.ty * it.tiles_x + tr.tx) * 4;
      dst[0] = (red[0][0] + red[1][0]) + (red[2][0] + red[3][0]);
      dst[1] = (red[0][1] + red[1][1]) + (red[2][1] + red[3][1]);
      dst[2] = (red[0][2] + red[1][2]) + (red[2][2] + red[3][2]);
      dst[3] = 0.0;
    }
  }
}

// Pixel store (REPLACE / ACCUM with the line kernel, RayCasterCPU::pre_compute semantics) and
// the fused NCC epilogue. Every thread of the block calls it (it synchronises).
template <int kOp>
__device__ inline void store_pixel(const DrrItem& it, const TileRef& tr, bool inside, int col, int row, float acc,
                                   float* __restrict__ out, int accumulate, float fill,
                                   double* __restrict__ ncc_partials, int ncc_tiles)
{
  float pix = 0.0f;
  if (inside) {
    const int64_t o = it.out_off + (int64_t)row * it.cam.cols + col;
    const float prev = accumulate ? out[o] : fill;
    const float aa = 0.0f + acc * 1.0f;
    if constexpr (kOp == XRH_LINE_INT_SUM) pix = prev + aa; else pix = (prev < aa) ? aa : prev;
    out[o] = pix;
  }
  ncc_epilogue(it, tr, inside, col, row, pix, ncc_partials, ncc_tiles);
}

// One ray of one item: the line integral (SUM) or maximum (MAX) of its samples times the step,
// or the kernel's initial value when the ray misses the volume. Called by every lane of the
// wave (`inside` false for lanes past the detector edge), so that the lanes can agree on a
// start: with an empty-space map each lane marches from the wave's earliest first sample
// instead of its own. The samples before a lane's own first lie in empty bricks and are exactly
// 0, so its sum is unchanged bit for bit, and neighbouring rays, which enter the volume at
// nearly the same depth, then fetch the same cells in the same iteration (C2: L1 tag accesses
// -7.5%, L1 tag-conflict stalls -40%, DRR -2%).
template <int kGeom, int kOp, int kFmt>
__device__ inline float cast_ray(const DrrItem& it, int col, int row, bool inside)
{
  float acc = (kOp == XRH_LINE_INT_SUM) ? 0.0f : -3.402823466e+38f;
  Ray ray;
  ray.n = -1;
  int i = 0, last = -1;
  if (inside) {
    ray = (kGeom == XRH_GEOM_FP32) ? setup_ray_f32(it, (float)col, (float)row) : setup_ray_f64(it, col, row);
    if (ray.n >= 0) sample_range<kOp, kGeom>(it, ray, i, last);
  }
  // wave minimum of the first samples of the lanes with samples (all lanes take part)
  int start = (i <= last) ? i : 0x7fffffff;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) start = min(start, __shfl_xor(start, off, 64));
  if (ray.n >= 0) {
    if (i <= last) {
      VolView vv;
      vv.rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(it.vol), (short)0,
                                                  (int)it.vol_bytes, 0x00020000);
      vv.py = it.pitch_y;
      vv.pz = it.pitch_z;
      vv.apron = 1u + it.pitch_y + it.pitch_z;
      vv.ce = 0xFFFFFFFFu;
      acc = march<kGeom, kOp, kFmt>(ray, vv, start, last, acc);
    }
    acc = acc * it.step;
  }
  return acc;
}

template <int kGeom, int kOp, int kFmt>
__global__ __launch_bounds__(kThreads) void drr_kernel(
    const DrrItem* __restrict__ items, const uint32_t* __restrict__ band_start, int64_t num_items,
    int bands, float* __restrict__ out, int accumulate, float fill,
    double* __restrict__ ncc_partials, int ncc_tiles, int shape)
{
  const TileRef tr = decode_tile(band_start, num_items, bands, items);
  if (!tr.valid) return;
  const DrrItem& it = items[tr.item];
  int col, row;
  lane_pixel(tr.tx, tr.ty, shape, col, row);
  const bool inside = (col < it.cam.cols) && (row < it.cam.rows);
  const float acc = cast_ray<kGeom, kOp, kFmt>(it, col, row, inside);
  store_pixel<kOp>(it, tr, inside, col, row, acc, out, accumulate, fill, ncc_partials, ncc_tiles);
}

// Several objects per item (Intensity2D3DRegi::obj_fn with more than one volume,
// xregIntensity2D3DRegi.cpp:598-629): entries item * objs + k, k = 0..objs-1, are cast in
// object order into the item's pixels, the first onto its background image (or `fill`), the
// others added with the line kernel -- per pixel the same float operations as one REPLACE
// launch followed by ACCUM launches. The item's first entry carries the output offset, the
// detector and the fused NCC reference.
template <int kGeom, int kOp, int kFmt>
__global__ __launch_bounds__(kThreads) void drr_objects_kernel(
    const DrrItem* __restrict__ items, const uint32_t* __restrict__ band_start, int64_t num_items,
    int bands, float* __restrict__ out, float fill, double* __restrict__ ncc_partials, int ncc_tiles,
    int shape, int objs)
{
  const TileRef tr = decode_tile(band_start, num_items, bands, items, objs);
  if (!tr.valid) return;
  const DrrItem& it0 = items[tr.item * objs];
  int col, row;
  lane_pixel(tr.tx, tr.ty, shape, col, row);
  const bool inside = (col < it0.cam.cols) && (row < it0.cam.rows);
  const int64_t px = (int64_t)row * it0.cam.cols + col;
  float pix = 0.0f;
  if (inside) pix = it0.bg ? it0.bg[px] : fill;
  for (int k = 0; k < objs; ++k) {
    const float aa = 0.0f + cast_ray<kGeom, kOp, kFmt>(items[tr.item * objs + k], col, row, inside) * 1.0f;
    if (inside) {
      if constexpr (kOp == XRH_LINE_INT_SUM) pix = pix + aa; else pix = (pix < aa) ? aa : pix;
    }
  }
  if (inside) out[it0.out_off + px] = pix;
  ncc_epilogue(it0, tr, inside, col, row, pix, ncc_partials, ncc_tiles);
}

template <int kGeom>
__global__ __launch_bounds__(kThreads) void drr_count_kernel(
    const DrrItem* __restrict__ items, const uint32_t* __restrict__ band_start, int64_t num_items,
    int bands, unsigned long long* __restrict__ counter, int32_t* __restrict__ steps)
{
  const TileRef tr = decode_tile(band_start, num_items, bands, items);
  if (!tr.valid) return;
  const DrrItem& it = items[tr.item];
  int col, row;
  lane_pixel(tr.tx, tr.ty, 0, col, row);
  double cnt = 0.0, evaluated = 0.0;
  if ((col < it.cam.cols) && (row < it.cam.rows)) {
    const Ray ray = (kGeom == XRH_GEOM_FP32) ? setup_ray_f32(it, (float)col, (float)row) : setup_ray_f64(it, col, row);
    if (steps) steps[it.out_off + (int64_t)row * it.cam.cols + col] = ray.n;
    cnt = (double)(ray.n + 1);
    if (ray.n >= 0) {
      int first, last;
      sample_range<XRH_LINE_INT_SUM, kGeom>(it, ray, first, last);
      evaluated = (double)(last - first + 1);
    }
  }
  cnt = wave_sum_d(cnt);
  evaluated = wave_sum_d(evaluated);
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&counter[0], (unsigned long long)cnt);
    atomicAdd(&counter[1], (unsigned long long)evaluated);
  }
}

// Compute units of the current device (cached per device).
int device_cus()
{
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cache[dev] == 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cache[dev] = n;
  }
  return cache[dev];
}

template <int kGeom, int kOp>
void launch_fmt(const DrrLaunch& L, dim3 grid, hipStream_t s, int accumulate, float fill)
{
#define XRH_LAUNCH_FMT(F)                                                                      \
  do {                                                                                         \
    if (L.objs > 1) {                                                                          \
      hipLaunchKernelGGL((drr_objects_kernel<kGeom, kOp, F>), grid, dim3(kThreads), 0, s,      \
                         L.items_dev, L.band_start_dev, L.num_items, L.bands, L.out, fill,     \
                         L.ncc_partials, L.ncc_tiles_per_item, shape, L.objs);                 \
    } else {                                                                                   \
      hipLaunchKernelGGL((drr_kernel<kGeom, kOp, F>), grid, dim3(kThreads), 0, s, L.items_dev, \
                         L.band_start_dev, L.num_items, L.bands, L.out, accumulate, fill,      \
                         L.ncc_partials, L.ncc_tiles_per_item, shape);                         \
    }                                                                                          \
  } while (0)
  // Cell reuse (one sample per iteration, loads only on cell change) wins when the launch
  // fills every CU many times over and the kernel is bound by load issue; small launches are
  // latency-bound and prefer the two-samples-per-iteration kernel. XRH_CELL_REUSE=0 / =1
  // force either (read per launch, so a process can switch between calls); otherwise a tuned
  // variant decides, else the launch size.
  const char* reuse_env = std::getenv("XRH_CELL_REUSE");
  const bool reuse = reuse_env ? (std::strcmp(reuse_env, "0") != 0)
                     : (L.reuse >= 0) ? (L.reuse != 0)
                                      : ((int64_t)grid.x * grid.y >= 16ll * device_cus());
  const int shape = (L.wave_shape >= 0 && L.wave_shape <= 2) ? L.wave_shape : 0;
  const int fmt = (L.storage == XRH_F32)
                      ? (L.layout == kLayoutCoef ? (reuse ? 7 : 6)
                                                 : (L.layout == kLayoutFace ? (reuse ? 11 : 10)
                                                    : (L.layout == kLayoutBrick ? (reuse ? 13 : 12)
                                                                                : (L.layout == kLayoutQuad ? 1 : 0))))
                : (L.storage == XRH_F16) ? (L.layout == kLayoutOct ? 3 : 2)
                                         : (L.layout == kLayoutOct ? 5 : 4);
  switch (fmt) {
    case 0: XRH_LAUNCH_FMT(0); break;
    case 1: XRH_LAUNCH_FMT(1); break;
    case 2: XRH_LAUNCH_FMT(2); break;
    case 3: XRH_LAUNCH_FMT(3); break;
    case 4: XRH_LAUNCH_FMT(4); break;
    case 5: XRH_LAUNCH_FMT(5); break;
    case 6: XRH_LAUNCH_FMT(6); break;
    case 7: XRH_LAUNCH_FMT(7); break;
    case 10: XRH_LAUNCH_FMT(10); break;
    case 12: XRH_LAUNCH_FMT(12); break;
    case 13: XRH_LAUNCH_FMT(13); break;
    default: XRH_LAUNCH_FMT(11); break;
  }
#undef XRH_LAUNCH_FMT
}

}  // namespace

bool drr_default_reuse(int64_t workgroups) { return workgroups >= 16ll * device_cus(); }

std::string drr_tune_key(int32_t storage, int32_t layout, int32_t geom, int32_t kernel, int32_t max_rows,
                         int32_t max_cols, int64_t workgroups)
{
  const int64_t per_cu = workgroups / (int64_t)device_cus();
  int lg = 0;
  while (lg < 40 && (2ll << lg) <= per_cu) ++lg;
  char buf[160];
  std::snprintf(buf, sizeof(buf), "s%d.l%d.g%d.k%d.%dx%d.wg%d", storage, layout, geom, kernel, max_rows, max_cols,
                per_cu > 0 ? lg : -1);
  return buf;
}

TilePlan build_tiles(const std::vector<int32_t>& rows, const std::vector<int32_t>& cols, int32_t bands)
{
  TilePlan p;
  p.bands = bands < 1 ? 1 : bands;
  const int64_t n = (int64_t)rows.size();
  p.band_start.assign((size_t)p.bands * (n + 1), 0);
  p.max_band_tiles = 0;
  for (int b = 0; b < p.bands; ++b) {
    uint32_t acc = 0;
    for (int64_t i = 0; i < n; ++i) {
      const int TX = (cols[i] + kTile - 1) / kTile;
      const int TY = (rows[i] + kTile - 1) / kTile;
      const int r0 = (int)(((int64_t)b * TY) / p.bands);
      const int r1 = (int)(((int64_t)(b + 1) * TY) / p.bands);
      p.band_start[(size_t)b * (n + 1) + i] = acc;
      acc += (uint32_t)((r1 - r0) * TX);
    }
    p.band_start[(size_t)b * (n + 1) + n] = acc;
    if ((int64_t)acc > p.max_band_tiles) p.max_band_tiles = acc;
  }
  return p;
}

void launch_drr_setup(const DrrLaunch& L, hipStream_t s)
{
  if (L.num_items <= 0) return;
  const int64_t entries = L.num_items * (L.objs > 1 ? L.objs : 1);
  const int threads = 64;
  const int blocks = (int)((entries + threads - 1) / threads);
  hipLaunchKernelGGL(drr_setup_kernel, dim3(blocks), dim3(threads), 0, s, L.poses_dev, L.vols_dev,
                     L.cams_dev, L.items_dev, entries);
  XRH_HIP_CHECK(hipGetLastError());
}

void launch_drr(const DrrLaunch& L, hipStream_t s)
{
  if (L.num_items <= 0 || L.max_band_tiles <= 0) return;
  const int64_t nblocks = L.max_band_tiles * L.bands;
  XRH_REQUIRE(nblocks < (int64_t)INT32_MAX, "too many DRR tiles in one launch");
  const dim3 grid((unsigned)nblocks);
  if (L.sample_counter) {
    XRH_REQUIRE(L.objs <= 1, "sample counts are taken per single-object item");
    if (L.geom == XRH_GEOM_FP32) {
      hipLaunchKernelGGL((drr_count_kernel<XRH_GEOM_FP32>), grid, dim3(kThreads), 0, s, L.items_dev,
                         L.band_start_dev, L.num_items, L.bands,
                         (unsigned long long*)L.sample_counter, L.ray_steps);
    } else {
      hipLaunchKernelGGL((drr_count_kernel<XRH_GEOM_FP64>), grid, dim3(kThreads), 0, s, L.items_dev,
                         L.band_start_dev, L.num_items, L.bands,
                         (unsigned long long*)L.sample_counter, L.ray_steps);
    }
    XRH_HIP_CHECK(hipGetLastError());
    return;
  }
  const int accumulate = (L.store == XRH_STORE_ACCUM) ? 1 : 0;
  const float fill = L.fill;
  if (L.geom == XRH_GEOM_FP32) {
    if (L.kernel == XRH_LINE_INT_SUM) launch_fmt<XRH_GEOM_FP32, XRH_LINE_INT_SUM>(L, grid, s, accumulate, fill);
    else launch_fmt<XRH_GEOM_FP32, XRH_LINE_INT_MAX>(L, grid, s, accumulate, fill);
  } else {
    if (L.kernel == XRH_LINE_INT_SUM) launch_fmt<XRH_GEOM_FP64, XRH_LINE_INT_SUM>(L, grid, s, accumulate, fill);
    else launch_fmt<XRH_GEOM_FP64, XRH_LINE_INT_MAX>(L, grid, s, accumulate, fill);
  }
  XRH_HIP_CHECK(hipGetLastError());
}

}  // namespace xrh
