/*
 * xrh_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * CPU restatement of the reference's DRR + 2-D similarity hot path. Every function
 * cites the reference file:line it follows. Build: see oracle/Makefile (-O2
 * -ffp-contract=off so that no multiply-add is fused, matching the reference's
 * x86-64 SSE2 build where Eigen emits separate mul/add instructions).
 *
 * Where the reference's arithmetic lives in third-party code that is absent here
 * (ITK 5.1.1, Eigen 3.3.4, OpenCV 3.4.12), the published algorithm of that pinned
 * version is restated and the spot is marked "parity unpinned" -- its exact
 * rounding cannot be checked against the real library in this environment.
 */
#include "xrh_oracle.h"

#include <math.h>
#include <float.h>
#include <pthread.h>
#include <sched.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ------------------------------------------------------------------------------------ */
/* threading: blocked ranges handed out by an atomic counter (TBB parallel_for analogue,
 * lib/common/xregTBBUtils.h:93-101). Per-item arithmetic is sequential, so results do not
 * depend on the partition. */

typedef void (*orc_range_fn)(void* ctx, int64_t begin, int64_t end);

typedef struct {
  orc_range_fn fn;
  void* ctx;
  int64_t n;
  int64_t grain;
  int64_t next;
} orc_pool_job;

static void* orc_pool_worker(void* arg)
{
  orc_pool_job* job = (orc_pool_job*)arg;
  for (;;) {
    const int64_t b = __atomic_fetch_add(&job->next, job->grain, __ATOMIC_RELAXED);
    if (b >= job->n) break;
    const int64_t e = (b + job->grain < job->n) ? (b + job->grain) : job->n;
    job->fn(job->ctx, b, e);
  }
  return NULL;
}

static void orc_parallel_for(orc_range_fn fn, void* ctx, int64_t n, int64_t grain, int32_t threads)
{
  if (n <= 0) return;
  if (threads <= 0) {
    /* the CPUs this process may run on, at most 16 */
    cpu_set_t set;
    long hw = (sched_getaffinity(0, sizeof(set), &set) == 0) ? (long)CPU_COUNT(&set)
                                                             : sysconf(_SC_NPROCESSORS_ONLN);
    threads = (hw > 0) ? (int32_t)(hw < 16 ? hw : 16) : 1;
  }
  if (grain < 1) grain = 1;
  orc_pool_job job = {fn, ctx, n, grain, 0};
  if (threads == 1 || n <= grain) {
    orc_pool_worker(&job);
    return;
  }
  pthread_t* tids = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
  int32_t started = 0;
  for (int32_t t = 1; t < threads; ++t) {
    if (pthread_create(&tids[t], NULL, orc_pool_worker, &job) == 0) {
      ++started;
    } else {
      tids[t] = 0;
    }
  }
  orc_pool_worker(&job);
  for (int32_t t = 1; t < threads; ++t) {
    if (tids[t]) pthread_join(tids[t], NULL);
  }
  (void)started;
  free(tids);
}

/* ------------------------------------------------------------------------------------ */
/* Eigen 3.3.4 small fixed-size arithmetic as the reference's float types use it
 * (lib/common/xregCommon.h:44-65: CoordScalar = float, FrameTransform = Affine3f).
 * Fixed-size 3x3 products are coefficient-based: res(i) = (a_i0*b0 + a_i1*b1) + a_i2*b2.
 * parity unpinned: Eigen is absent, the order is the library's documented lazy product. */

static inline float dot3f(const float* row, int stride, const float v[3])
{
  return (row[0] * v[0] + row[stride] * v[1]) + row[2 * stride] * v[2];
}

void orc_affine_pt_f32(const float m[16], const float p[3], float out[3])
{
  /* Transform<float,3,Affine> * Vector3f = linear()*p + translation()
   * (Eigen/src/Geometry/Transform.h, transform_right_product_impl). */
  float r[3];
  for (int i = 0; i < 3; ++i) {
    r[i] = dot3f(&m[4 * i], 1, p) + m[4 * i + 3];
  }
  out[0] = r[0];
  out[1] = r[1];
  out[2] = r[2];
}

static inline void affine_pt_f64(const double m[16], const double p[3], double out[3])
{
  double r[3];
  for (int i = 0; i < 3; ++i) {
    r[i] = (m[4 * i] * p[0] + m[4 * i + 1] * p[1]) + m[4 * i + 2] * p[2] + m[4 * i + 3];
  }
  out[0] = r[0];
  out[1] = r[1];
  out[2] = r[2];
}

static inline void linear_vec_f32(const float m[16], const float v[3], float out[3])
{
  float r[3];
  for (int i = 0; i < 3; ++i) r[i] = dot3f(&m[4 * i], 1, v);
  out[0] = r[0];
  out[1] = r[1];
  out[2] = r[2];
}

void orc_affine_mul_f32(const float a[16], const float b[16], float out[16])
{
  /* Affine * Affine: linear = La*Lb, translation = La*tb + ta, last row [0 0 0 1]
   * (Eigen transform_transform_product_impl). */
  float r[16];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      r[4 * i + j] = (a[4 * i] * b[j] + a[4 * i + 1] * b[4 + j]) + a[4 * i + 2] * b[8 + j];
    }
    r[4 * i + 3] = ((a[4 * i] * b[3] + a[4 * i + 1] * b[7]) + a[4 * i + 2] * b[11]) + a[4 * i + 3];
  }
  r[12] = 0.0f;
  r[13] = 0.0f;
  r[14] = 0.0f;
  r[15] = 1.0f;
  memcpy(out, r, sizeof(r));
}

static void affine_mul_f64(const double a[16], const double b[16], double out[16])
{
  double r[16];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      r[4 * i + j] = (a[4 * i] * b[j] + a[4 * i + 1] * b[4 + j]) + a[4 * i + 2] * b[8 + j];
    }
    r[4 * i + 3] = ((a[4 * i] * b[3] + a[4 * i + 1] * b[7]) + a[4 * i + 2] * b[11]) + a[4 * i + 3];
  }
  r[12] = 0.0;
  r[13] = 0.0;
  r[14] = 0.0;
  r[15] = 1.0;
  memcpy(out, r, sizeof(r));
}

void orc_mat3_inverse_f32(const float m[9], float out[9])
{
  /* Eigen compute_inverse<...,3>: cofactors of column 0, det = sum(cof .* col0),
   * invdet = 1/det, every entry = cofactor * invdet (Eigen/src/LU/InverseImpl.h).
   * cofactor(i,j) uses rows i+1, i+2 and cols j+1, j+2 (mod 3); entry (j,i) = cof(i,j). */
#define M3(r, c) m[3 * (r) + (c)]
#define COF(i, j) (M3(((i) + 1) % 3, ((j) + 1) % 3) * M3(((i) + 2) % 3, ((j) + 2) % 3) - \
                   M3(((i) + 1) % 3, ((j) + 2) % 3) * M3(((i) + 2) % 3, ((j) + 1) % 3))
  const float c00 = COF(0, 0);
  const float c10 = COF(1, 0);
  const float c20 = COF(2, 0);
  /* (cofactors_col0.cwiseProduct(matrix.col(0))).sum(): 3 elements, linear redux */
  const float det = (c00 * M3(0, 0) + c10 * M3(1, 0)) + c20 * M3(2, 0);
  const float invdet = 1.0f / det;
  float r[9];
  r[0] = c00 * invdet;
  r[1] = c10 * invdet;
  r[2] = c20 * invdet;
  r[3] = COF(0, 1) * invdet;
  r[4] = COF(1, 1) * invdet;
  r[5] = COF(2, 1) * invdet;
  r[6] = COF(0, 2) * invdet;
  r[7] = COF(1, 2) * invdet;
  r[8] = COF(2, 2) * invdet;
#undef COF
#undef M3
  memcpy(out, r, sizeof(r));
}

static void mat3_inverse_f64(const double m[9], double out[9])
{
#define M3(r, c) m[3 * (r) + (c)]
#define COF(i, j) (M3(((i) + 1) % 3, ((j) + 1) % 3) * M3(((i) + 2) % 3, ((j) + 2) % 3) - \
                   M3(((i) + 1) % 3, ((j) + 2) % 3) * M3(((i) + 2) % 3, ((j) + 1) % 3))
  const double c00 = COF(0, 0), c10 = COF(1, 0), c20 = COF(2, 0);
  const double det = (c00 * M3(0, 0) + c10 * M3(1, 0)) + c20 * M3(2, 0);
  const double invdet = 1.0 / det;
  double r[9] = {c00 * invdet, c10 * invdet, c20 * invdet,
                 COF(0, 1) * invdet, COF(1, 1) * invdet, COF(2, 1) * invdet,
                 COF(0, 2) * invdet, COF(1, 2) * invdet, COF(2, 2) * invdet};
#undef COF
#undef M3
  memcpy(out, r, sizeof(r));
}

void orc_affine_inverse_f64(const double m[16], double out[16])
{
  double lin[9] = {m[0], m[1], m[2], m[4], m[5], m[6], m[8], m[9], m[10]};
  double inv[9];
  mat3_inverse_f64(lin, inv);
  double r[16];
  for (int i = 0; i < 3; ++i) {
    r[4 * i + 0] = inv[3 * i + 0];
    r[4 * i + 1] = inv[3 * i + 1];
    r[4 * i + 2] = inv[3 * i + 2];
    r[4 * i + 3] = -((inv[3 * i] * m[3] + inv[3 * i + 1] * m[7]) + inv[3 * i + 2] * m[11]);
  }
  r[12] = 0.0;
  r[13] = 0.0;
  r[14] = 0.0;
  r[15] = 1.0;
  memcpy(out, r, sizeof(r));
}

void orc_affine_inverse_f32(const float m[16], float out[16])
{
  /* Transform<float,3,Affine>::inverse(): linear part via the 3x3 inverse, translation
   * = -(Linv * t) (Eigen/src/Geometry/Transform.h, inverse()). */
  float lin[9] = {m[0], m[1], m[2], m[4], m[5], m[6], m[8], m[9], m[10]};
  float inv[9];
  orc_mat3_inverse_f32(lin, inv);
  float r[16];
  for (int i = 0; i < 3; ++i) {
    r[4 * i + 0] = inv[3 * i + 0];
    r[4 * i + 1] = inv[3 * i + 1];
    r[4 * i + 2] = inv[3 * i + 2];
    const float t = (inv[3 * i] * m[3] + inv[3 * i + 1] * m[7]) + inv[3 * i + 2] * m[11];
    r[4 * i + 3] = -t;
  }
  r[12] = 0.0f;
  r[13] = 0.0f;
  r[14] = 0.0f;
  r[15] = 1.0f;
  memcpy(out, r, sizeof(r));
}

void orc_cam_naive_f32(float focal, int32_t rows, int32_t cols, float row_sp, float col_sp,
                       orc_cam_f* out)
{
  /* CameraModel::setup(f, nr, nc, rs, cs) (lib/transforms/xregPerspectiveXform.cpp:230-254)
   * with MakeNaiveIntrins (:200-228), NEG_Z frame (default, xregPerspectiveXform.h:173). */
  float K[9] = {0};
  K[0] = focal / col_sp;
  K[4] = focal / row_sp;
  K[8] = 1.0f;
  K[0] *= -1.0f;
  K[4] *= -1.0f;
  K[2] = (float)((double)(cols - 1) * 0.5);
  K[5] = (float)((double)(rows - 1) * 0.5);
  memset(out, 0, sizeof(*out));
  out->rows = rows;
  out->cols = cols;
  out->frame = 1;
  out->focal = focal;
  orc_mat3_inverse_f32(K, out->kinv);
  out->einv[0] = out->einv[5] = out->einv[10] = out->einv[15] = 1.0f;
  out->pinhole[0] = out->pinhole[1] = out->pinhole[2] = 0.0f;
}

void orc_cam_ind_to_det_f32(const orc_cam_f* cam, float col, float row, float out[3])
{
  /* CameraModel::ind_pt_to_phys_det_pt (xregPerspectiveXform.cpp:391-414):
   *   extrins_inv * ((intrins_inv * (det_z * [c, r, 1])) + z_adjust) */
  const float det_z = ((cam->frame == 1) ? -1.0f : 1.0f) * cam->focal;
  const float q[3] = {det_z * col, det_z * row, det_z * 1.0f};
  float r[3];
  for (int i = 0; i < 3; ++i) r[i] = dot3f(&cam->kinv[3 * i], 1, q);
  const float zadj = (cam->frame == 2) ? -cam->focal : 0.0f;
  r[0] = r[0] + 0.0f;
  r[1] = r[1] + 0.0f;
  r[2] = r[2] + zadj;
  orc_affine_pt_f32(cam->einv, r, out);
}

void orc_cam_ind_to_det_f64(const orc_cam_d* cam, double col, double row, double out[3])
{
  const double det_z = ((cam->frame == 1) ? -1.0 : 1.0) * cam->focal;
  const double q[3] = {det_z * col, det_z * row, det_z};
  double r[3];
  for (int i = 0; i < 3; ++i) {
    r[i] = (cam->kinv[3 * i] * q[0] + cam->kinv[3 * i + 1] * q[1]) + cam->kinv[3 * i + 2] * q[2];
  }
  if (cam->frame == 2) r[2] -= cam->focal;
  affine_pt_f64(cam->einv, r, out);
}

int32_t orc_ray_rect_f32(const float mn[3], const float mx[3], const float p0[3],
                         const float d[3], int32_t limit_to_segment, float* t0_out, float* t1_out)
{
  /* RayRectIntersect (lib/spatial/xregSpatialPrimitives.cpp:175-222), slab method. */
  int32_t hit = 1;
  float t_start = 0.0f;
  float t_stop = limit_to_segment ? 1.0f : INFINITY;
  for (int k = 0; k < 3; ++k) {
    if ((double)fabsf(d[k]) > 1.0e-8) { /* float component vs the double literal 1.0e-8 */
      const float inv = 1.0f / d[k];
      float ta = (mn[k] - p0[k]) * inv;
      float tb = (mx[k] - p0[k]) * inv;
      if (tb < ta) {
        const float tmp = ta;
        ta = tb;
        tb = tmp;
      }
      t_start = (t_start > ta) ? t_start : ta; /* std::max(t_start, t[0]) */
      t_stop = (tb < t_stop) ? tb : t_stop;    /* std::min(t_stop, t[1]) */
      if (t_start > t_stop) {
        hit = 0;
        break;
      }
    } else if ((p0[k] < mn[k]) || (p0[k] > mx[k])) {
      hit = 0;
      break;
    }
  }
  *t0_out = t_start;
  *t1_out = t_stop;
  return hit;
}

int32_t orc_ray_rect_f64(const double mn[3], const double mx[3], const double p0[3],
                         const double d[3], int32_t limit_to_segment, double* t0_out, double* t1_out)
{
  int32_t hit = 1;
  double t_start = 0.0;
  double t_stop = limit_to_segment ? 1.0 : INFINITY;
  for (int k = 0; k < 3; ++k) {
    if (fabs(d[k]) > 1.0e-8) {
      const double inv = 1.0 / d[k];
      double ta = (mn[k] - p0[k]) * inv;
      double tb = (mx[k] - p0[k]) * inv;
      if (tb < ta) {
        const double tmp = ta;
        ta = tb;
        tb = tmp;
      }
      t_start = (t_start > ta) ? t_start : ta;
      t_stop = (tb < t_stop) ? tb : t_stop;
      if (t_start > t_stop) {
        hit = 0;
        break;
      }
    } else if ((p0[k] < mn[k]) || (p0[k] > mx[k])) {
      hit = 0;
      break;
    }
  }
  *t0_out = t_start;
  *t1_out = t_stop;
  return hit;
}

/* ------------------------------------------------------------------------------------ */
/* Trilinear interpolation as itk::LinearInterpolateImageFunction<Image<float,3>, float>
 * evaluates it (called at lib/ray_cast/xregRayCastLineIntCPU.cpp:273-274):
 * base = floor(index) clamped below at 0, distance = index - base in float (the
 * coordinate type), values and lerps in double (RealType of float), x then y then z;
 * a dimension whose distance is <= 0, or whose upper neighbour is past the end, is not
 * interpolated (which equals clamping the neighbour to the base voxel). The result is
 * rounded to float when the reference passes it to the line-integral kernel.
 * parity unpinned: ITK is absent; this follows the published 5.1 EvaluateOptimized. */
double orc_itk_linear_interp(const float* vol, const int64_t dims[3], const float idx[3])
{
  int64_t b[3], u[3];
  double w[3];
  for (int k = 0; k < 3; ++k) {
    int64_t bk = (int64_t)floorf(idx[k]);
    if (bk < 0) bk = 0;
    if (bk > dims[k] - 1) bk = dims[k] - 1; /* guard; the 1e-3 nudge keeps this unreached */
    const float dk = idx[k] - (float)bk;
    b[k] = bk;
    u[k] = (bk + 1 > dims[k] - 1) ? bk : bk + 1;
    w[k] = (dk > 0.0f) ? (double)dk : 0.0;
  }
  const int64_t X = dims[0], XY = dims[0] * dims[1];
#define V(i, j, k) ((double)vol[(i) + X * (j) + XY * (k)])
  const double v000 = V(b[0], b[1], b[2]), v100 = V(u[0], b[1], b[2]);
  const double v010 = V(b[0], u[1], b[2]), v110 = V(u[0], u[1], b[2]);
  const double v001 = V(b[0], b[1], u[2]), v101 = V(u[0], b[1], u[2]);
  const double v011 = V(b[0], u[1], u[2]), v111 = V(u[0], u[1], u[2]);
#undef V
  const double x00 = v000 + (v100 - v000) * w[0];
  const double x10 = v010 + (v110 - v010) * w[0];
  const double x01 = v001 + (v101 - v001) * w[0];
  const double x11 = v011 + (v111 - v011) * w[0];
  const double xx0 = x00 + (x10 - x00) * w[1];
  const double xx1 = x01 + (x11 - x01) * w[1];
  return xx0 + (xx1 - xx0) * w[2];
}

double orc_linear_interp_f64(const double* vol, const int64_t dims[3], const double idx[3])
{
  int64_t b[3], u[3];
  double w[3];
  for (int k = 0; k < 3; ++k) {
    int64_t bk = (int64_t)floor(idx[k]);
    if (bk < 0) bk = 0;
    if (bk > dims[k] - 1) bk = dims[k] - 1;
    const double dk = idx[k] - (double)bk;
    b[k] = bk;
    u[k] = (bk + 1 > dims[k] - 1) ? bk : bk + 1;
    w[k] = (dk > 0.0) ? dk : 0.0;
  }
  const int64_t X = dims[0], XY = dims[0] * dims[1];
#define V(i, j, k) (vol[(i) + X * (j) + XY * (k)])
  const double x00 = V(b[0], b[1], b[2]) + (V(u[0], b[1], b[2]) - V(b[0], b[1], b[2])) * w[0];
  const double x10 = V(b[0], u[1], b[2]) + (V(u[0], u[1], b[2]) - V(b[0], u[1], b[2])) * w[0];
  const double x01 = V(b[0], b[1], u[2]) + (V(u[0], b[1], u[2]) - V(b[0], b[1], u[2])) * w[0];
  const double x11 = V(b[0], u[1], u[2]) + (V(u[0], u[1], u[2]) - V(b[0], u[1], u[2])) * w[0];
#undef V
  const double xx0 = x00 + (x10 - x00) * w[1];
  const double xx1 = x01 + (x11 - x01) * w[1];
  return xx0 + (xx1 - xx0) * w[2];
}

/* ------------------------------------------------------------------------------------ */
/* The CPU caster's other interpolators (xregRayCastLineIntCPU.cpp:115-147), each evaluated
 * at a float continuous index and returning double, as ITK 5.1's EvaluateAtContinuousIndex
 * does for Image<float,3> with CoordRep float.
 * parity unpinned against ITK itself (absent here); pinned by the closed forms and library
 * routines in tests/test_oracle_interp.py. */

/* itk::Math::RoundHalfIntegerUp<int64>(float) and itk::Math::Floor<int64>(float) as ITK 5.1
 * compiles them on x86-64 with SSE2 (itkMathDetail.h): cvtss2si, i.e. round-to-nearest-even
 * under the default MXCSR, of 2x + 0.5f (resp. 2x - 0.5f) computed in float, shifted right. */
static int64_t itk_round_half_up(float x) { return ((int64_t)nearbyintf(x + x + 0.5f)) >> 1; }
static int64_t itk_floor(float x) { return ((int64_t)nearbyintf(x + x - 0.5f)) >> 1; }

static double vox_or_zero(const float* vol, const int64_t dims[3], int64_t x, int64_t y, int64_t z)
{
  if (x < 0 || y < 0 || z < 0 || x >= dims[0] || y >= dims[1] || z >= dims[2]) return 0.0;
  return (double)vol[x + dims[0] * (y + dims[1] * z)];
}

/* NearestNeighborInterpolateImageFunction: the voxel at the rounded index (ITK does not test
 * the bounds; every sample of a ray lies inside the index box, and out-of-range reads 0 here). */
double orc_itk_nn_interp(const float* vol, const int64_t dims[3], const float idx[3])
{
  return vox_or_zero(vol, dims, itk_round_half_up(idx[0]), itk_round_half_up(idx[1]), itk_round_half_up(idx[2]));
}

/* WindowedSincInterpolateImageFunction<Image<float,3>, 4, LanczosWindowFunction<4>,
 * ConstantBoundaryCondition, float>: base = Floor(index), distance in double; per dimension
 * x runs distance + 3 down to distance - 4 and the weight is Lanczos(x) * Sinc(x) with
 * Sinc(x) = sin(pi x) / (pi x) and Lanczos(x) = sin(pi x / 4) / (pi x / 4) (1 at x = 0); a zero
 * distance gives the unit impulse on offset 0. The neighbourhood (offsets -3..4, x fastest)
 * is summed in iteration order, each term the pixel times the x, y and z weights in that order;
 * pixels outside the image are 0. */
double orc_itk_sinc_interp(const float* vol, const int64_t dims[3], const float idx[3])
{
  const double pi = 3.14159265358979323846;
  int64_t base[3];
  double w[3][8];
  int k, i, j, l;
  for (k = 0; k < 3; ++k) {
    base[k] = itk_floor(idx[k]);
    const double dist = (double)idx[k] - (double)base[k];
    if (dist == 0.0) {
      for (i = 0; i < 8; ++i) w[k][i] = (i == 3) ? 1.0 : 0.0;
    } else {
      double x = dist + 4.0;
      for (i = 0; i < 8; ++i) {
        x -= 1.0;
        const double zl = (pi / 4.0) * x;
        const double lanczos = (x == 0.0) ? 1.0 : sin(zl) / zl;
        const double px = pi * x;
        const double sinc = (x == 0.0) ? 1.0 : sin(px) / px;
        w[k][i] = lanczos * sinc;
      }
    }
  }
  double sum = 0.0;
  for (l = 0; l < 8; ++l)
    for (j = 0; j < 8; ++j)
      for (i = 0; i < 8; ++i) {
        double v = vox_or_zero(vol, dims, base[0] - 3 + i, base[1] - 3 + j, base[2] - 3 + l);
        v *= w[0][i];
        v *= w[1][j];
        v *= w[2][l];
        sum += v;
      }
  return sum;
}

static void orc_bspline_line(double* c, int64_t n, int64_t stride);

/* BSplineDecompositionImageFilter (order 3) of a whole volume: the voxels in double, filtered
 * along x, then y, then z with the mirror-boundary recursion of orc_bspline_line. */
void orc_bspline_coefficients(const float* vol, const int64_t dims[3], double* coef)
{
  const int64_t X = dims[0], Y = dims[1], Z = dims[2];
  int64_t i, a, b;
  for (i = 0; i < X * Y * Z; ++i) coef[i] = (double)vol[i];
  for (b = 0; b < Z; ++b)
    for (a = 0; a < Y; ++a) orc_bspline_line(coef + X * (a + Y * b), X, 1);
  for (b = 0; b < Z; ++b)
    for (a = 0; a < X; ++a) orc_bspline_line(coef + a + X * Y * b, Y, X);
  for (b = 0; b < Y; ++b)
    for (a = 0; a < X; ++a) orc_bspline_line(coef + a + X * b, Z, X * Y);
}

/* ITK's single-reflection mirror (BSplineInterpolateImageFunction::
 * ApplyMirrorBoundaryConditions): i < 0 -> -i, i >= n - 1 -> 2 (n - 1) - i; length 1 -> 0. */
static int64_t itk_mirror(int64_t i, int64_t n)
{
  if (n == 1) return 0;
  if (i < 0) i = -i;
  if (i >= n - 1) i = 2 * (n - 1) - i;
  return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
}

/* BSplineInterpolateImageFunction (order 3, double coefficients)::EvaluateAtContinuousIndex:
 * support floor(x) - 1 .. floor(x) + 2 per dimension, weights of w = x - (floor(x)) as ITK's
 * SetInterpolationWeights writes them, mirrored indices, and the 64 points in ITK's order
 * (p = i + 4 j + 16 l, x fastest), each weight the product over x, y, z. */
double orc_itk_bspline_interp(const double* coef, const int64_t dims[3], const float idx[3])
{
  int64_t ind[3][4];
  double w[3][4];
  int k, p;
  for (k = 0; k < 3; ++k) {
    const int64_t b = (int64_t)floor((double)idx[k]) - 1;
    int q;
    for (q = 0; q < 4; ++q) ind[k][q] = b + q;
    const double t = (double)idx[k] - (double)ind[k][1];
    w[k][3] = (1.0 / 6.0) * t * t * t;
    w[k][0] = (1.0 / 6.0) + 0.5 * t * (t - 1.0) - w[k][3];
    w[k][2] = t + w[k][0] - 2.0 * w[k][3];
    w[k][1] = 1.0 - w[k][0] - w[k][2] - w[k][3];
    for (q = 0; q < 4; ++q) ind[k][q] = itk_mirror(ind[k][q], dims[k]);
  }
  double v = 0.0;
  for (p = 0; p < 64; ++p) {
    const int i = p % 4, j = (p / 4) % 4, l = p / 16;
    double wt = 1.0;
    wt *= w[0][i];
    wt *= w[1][j];
    wt *= w[2][l];
    v += wt * coef[ind[0][i] + dims[0] * (ind[1][j] + dims[1] * ind[2][l])];
  }
  return v;
}

/* ------------------------------------------------------------------------------------ */
/* DRR, FP32 restatement of ComputeLineInts<Kernel> (xregRayCastLineIntCPU.cpp:105-290). */

typedef struct {
  const float* vol;
  const int64_t* dims;
  float aabb_max[3];
  const float* p2i;
  const orc_cam_f* cams;
  const float* xforms;
  const int64_t* cam_for_proj;
  float step;
  int32_t kernel;
  int64_t npx;
  int32_t cols0;
  float* out;
  int64_t samples;
  int32_t* steps;     /* optional: per ray, the last sample index n (-1: the ray misses) */
  int32_t steps_only; /* set-up only: write steps, take no samples, leave out untouched */
  int32_t interp;     /* RayCaster::InterpMethod: 0 linear, 1 NN, 2 sinc, 3 B-spline */
  const double* coef; /* B-spline coefficients (interp 3) */
  int32_t aa_rays;    /* RayCasterCPU::aa_fact_: 0 off, k >= 1 jittered rays per pixel */
  uint64_t aa_seed;   /* seed of orc_aa_jitter */
} drr_f32_ctx;

/* SplitMix64's output function (Steele, Lea and Flood, OOPSLA 2014). */
static uint64_t splitmix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Anti-aliasing jitter of ray k of range index `pixel`, component 0 (column) or 1 (row): a
 * uniform draw on [-0.5, 0.5) in steps of 2^-24, standing in for the reference's
 * uniform_real_distribution<float>(-0.5, 0.5) over a random_device-seeded mt19937
 * (xregRayCastLineIntCPU.cpp:162-174), which no two runs share; this one is a pure function of
 * its arguments (counter-based SplitMix64), the same as the device's. */
float orc_aa_jitter(uint64_t seed, uint64_t pixel, int32_t k, int32_t comp)
{
  const uint64_t counter = (pixel << 17) | ((uint64_t)k << 1) | (uint64_t)comp;
  const uint64_t h = splitmix64(seed + 0x9E3779B97F4A7C15ull * (counter + 1));
  return (float)(uint32_t)(h >> 40) * 0x1.0p-24f - 0.5f;
}

/* One ray through detector index (colf, rowf) of `cam` under pose-composed M (:198-279):
 * returns the line value (the kernel's initial value when the ray misses the volume) and, in
 * *n_out, the last sample index (-1: miss). With take_samples = 0 only the set-up runs. */
static float ray_f32(const drr_f32_ctx* c, const orc_cam_f* cam, const float M[16], float colf, float rowf,
                     int take_samples, int64_t* n_out, int64_t* samples)
{
  const float aabb_min[3] = {0.0f, 0.0f, 0.0f};
  const float kTol = 1.0e-3f; /* RayCasterCPU::kVOL_BB_STEP_INC_TOL (xregRayCastBaseCPU.h:37) */
  float det[3];
  orc_cam_ind_to_det_f32(cam, colf, rowf, det);                               /* :204 */
  float ph[3], dp[3], d[3];
  orc_affine_pt_f32(M, cam->pinhole, ph);                                    /* :211 */
  orc_affine_pt_f32(M, det, dp);                                             /* :215-216 */
  d[0] = dp[0] - ph[0];
  d[1] = dp[1] - ph[1];
  d[2] = dp[2] - ph[2];

  float acc = (c->kernel == 0) ? 0.0f : -FLT_MAX;                           /* :221 */
  float t0 = 0.0f, t1 = 0.0f;
  const int32_t hit = orc_ray_rect_f32(aabb_min, c->aabb_max, ph, d, 1, &t0, &t1); /* :227 */
  *n_out = -1;
  if (hit && ((t1 - t0) > (float)(2.0f * kTol))) {                            /* :233 */
    t0 += kTol;
    t1 -= kTol;
    const float start[3] = {ph[0] + t0 * d[0], ph[1] + t0 * d[1], ph[2] + t0 * d[2]};
    /* Eigen norm(): sqrt of the linear 3-element redux of squares */
    const float dlen = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);   /* :243-244 */
    const float L = (t1 - t0) * dlen;                                        /* :245-246 */
    /* step length in index space: ||M3x3 * (normalized(det - pinhole) * step)|| (:249-251) */
    float v[3] = {det[0] - cam->pinhole[0], det[1] - cam->pinhole[1], det[2] - cam->pinhole[2]};
    const float vz = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    if (vz > 0.0f) {
      const float vn = sqrtf(vz);
      v[0] = v[0] / vn;
      v[1] = v[1] / vn;
      v[2] = v[2] / vn;
    }
    v[0] = v[0] * c->step;
    v[1] = v[1] * c->step;
    v[2] = v[2] * c->step;
    float sv[3];
    linear_vec_f32(M, v, sv);
    const float s = sqrtf((sv[0] * sv[0] + sv[1] * sv[1]) + sv[2] * sv[2]);
    const int64_t n = (int64_t)(L / s);                                      /* :253-254 */
    *n_out = n;
    if (!take_samples) return acc;
    const float scale = s / dlen;                                            /* :263-264 */
    const float dx = d[0] * scale, dy = d[1] * scale, dz = d[2] * scale;
    float p[3] = {start[0], start[1], start[2]};
    for (int64_t i = 0; i <= n; ++i) {                                       /* :270-277 */
      double dv;
      switch (c->interp) {
        case 1: dv = orc_itk_nn_interp(c->vol, c->dims, p); break;
        case 2: dv = orc_itk_sinc_interp(c->vol, c->dims, p); break;
        case 3: dv = orc_itk_bspline_interp(c->coef, c->dims, p); break;
        default: dv = orc_itk_linear_interp(c->vol, c->dims, p); break;
      }
      const float val = (float)dv;
      if (c->kernel == 0) {
        acc = acc + val;
      } else {
        acc = (acc < val) ? val : acc; /* std::max(a, b) = (a < b) ? b : a */
      }
      p[0] = p[0] + dx;
      p[1] = p[1] + dy;
      p[2] = p[2] + dz;
    }
    *samples += n + 1;
    acc = acc * c->step;                                                     /* :279 */
  }
  return acc;
}

static void drr_f32_range(void* vctx, int64_t begin, int64_t end)
{
  drr_f32_ctx* c = (drr_f32_ctx*)vctx;
  int64_t samples = 0;
  /* rays per pixel and their weight (:151-157) */
  const int32_t rays = (c->aa_rays > 0) ? c->aa_rays : 1;
  const float inv = 1.0f / (float)rays;
  for (int64_t r = begin; r < end; ++r) {
    /* index decomposition uses camera 0's size (:151-152, 179-183) */
    const int64_t proj = r / c->npx;
    const int64_t off = r - c->npx * proj;
    const int64_t row = off / c->cols0;
    const int64_t col = off - (int64_t)c->cols0 * row;
    const orc_cam_f* cam = &c->cams[c->cam_for_proj[proj]];
    float M[16];
    orc_affine_mul_f32(c->p2i, &c->xforms[16 * proj], M);                      /* :207-208 */
    float aa_sum = 0.0f;
    int32_t k;
    for (k = 0; k < rays; ++k) {                                               /* :186-283 */
      float colf = (float)col, rowf = (float)row;
      if (c->aa_rays > 0) {
        colf += orc_aa_jitter(c->aa_seed, (uint64_t)r, k, 0);
        rowf += orc_aa_jitter(c->aa_seed, (uint64_t)r, k, 1);
      }
      int64_t n;
      const float acc = ray_f32(c, cam, M, colf, rowf, !c->steps_only, &n, &samples);
      if (c->steps && k == 0) c->steps[r] = (int32_t)n;
      aa_sum += acc * inv;                                                     /* :282 */
    }
    if (c->steps_only) continue;
    /* store with the same kernel (:285) */
    if (c->kernel == 0) {
      c->out[r] = c->out[r] + aa_sum;
    } else {
      c->out[r] = (c->out[r] < aa_sum) ? aa_sum : c->out[r];
    }
  }
  __atomic_fetch_add(&c->samples, samples, __ATOMIC_RELAXED);
}

static void pre_compute_f32(float* out, const orc_cam_f* cams, const int64_t* cam_for_proj,
                            int64_t num_projs, int64_t npx, int32_t store, const float* bg,
                            float default_bg)
{
  /* RayCasterCPU::pre_compute (lib/ray_cast/xregRayCastBaseCPU.cpp:128-158) */
  (void)cams;
  if (bg) {
    for (int64_t p = 0; p < num_projs; ++p) {
      memcpy(out + p * npx, bg + cam_for_proj[p] * npx, sizeof(float) * (size_t)npx);
    }
  }
  if (store == 0 && !bg) {
    for (int64_t i = 0; i < num_projs * npx; ++i) out[i] = default_bg;
  }
}

void orc_drr_f32(const float* vol, const int64_t dims[3], const float p2i[16],
                 const orc_cam_f* cams, int32_t num_cams,
                 const float* xforms, const int64_t* cam_for_proj, int64_t num_projs,
                 float step, int32_t kernel, int32_t store, const float* bg, float default_bg,
                 float* out, int32_t threads, int64_t* samples_out)
{
  orc_drr_interp_f32(vol, dims, p2i, cams, num_cams, xforms, cam_for_proj, num_projs, step, kernel, store, bg,
                     default_bg, 0, 0, 0, out, threads, samples_out);
}

/* orc_drr_f32 with the interpolator RayCasterLineIntCPU picks from interp_method
 * (xregRayCastLineIntCPU.cpp:126-147): 0 linear, 1 nearest neighbour, 2 windowed sinc,
 * 3 cubic B-spline (coefficients computed once here; the reference recomputes them in
 * every range's interpolator from the same volume, with the same result), and aa_rays
 * jittered rays per pixel (RayCasterCPU::aa_fact_, :150-183; orc_aa_jitter from aa_seed). */
void orc_drr_interp_f32(const float* vol, const int64_t dims[3], const float p2i[16],
                        const orc_cam_f* cams, int32_t num_cams,
                        const float* xforms, const int64_t* cam_for_proj, int64_t num_projs,
                        float step, int32_t kernel, int32_t store, const float* bg, float default_bg,
                        int32_t interp, int32_t aa_rays, uint64_t aa_seed, float* out, int32_t threads,
                        int64_t* samples_out)
{
  (void)num_cams;
  drr_f32_ctx c;
  double* coef = NULL;
  memset(&c, 0, sizeof(c));
  if (interp == 3) {
    coef = (double*)malloc(sizeof(double) * (size_t)(dims[0] * dims[1] * dims[2]));
    orc_bspline_coefficients(vol, dims, coef);
  }
  c.interp = interp;
  c.coef = coef;
  c.aa_rays = aa_rays;
  c.aa_seed = aa_seed;
  c.vol = vol;
  c.dims = dims;
  /* ITKImageIndexBoundsAsEigen: max index = size - 1 (lib/itk/xregITKBasicImageUtils.h:56-75) */
  c.aabb_max[0] = (float)(dims[0] - 1);
  c.aabb_max[1] = (float)(dims[1] - 1);
  c.aabb_max[2] = (float)(dims[2] - 1);
  c.p2i = p2i;
  c.cams = cams;
  c.xforms = xforms;
  c.cam_for_proj = cam_for_proj;
  c.step = step;
  c.kernel = kernel;
  c.cols0 = cams[0].cols;
  c.npx = (int64_t)cams[0].rows * cams[0].cols;
  c.out = out;
  pre_compute_f32(out, cams, cam_for_proj, num_projs, c.npx, store, bg, default_bg);
  orc_parallel_for(drr_f32_range, &c, num_projs * c.npx, 2048, threads);
  if (samples_out) *samples_out = c.samples;
  free(coef);
}

/* DRR, FP64 oracle: identical algorithm, every quantity in double. */

typedef struct {
  const double* vol;
  const int64_t* dims;
  double aabb_max[3];
  const double* p2i;
  const orc_cam_d* cams;
  const double* xforms;
  const int64_t* cam_for_proj;
  double step;
  int32_t kernel;
  int64_t npx;
  int32_t cols0;
  double* out;
  int64_t samples;
  int32_t* steps;     /* as in drr_f32_ctx */
  int32_t steps_only;
} drr_f64_ctx;

static void drr_f64_range(void* vctx, int64_t begin, int64_t end)
{
  drr_f64_ctx* c = (drr_f64_ctx*)vctx;
  const double aabb_min[3] = {0.0, 0.0, 0.0};
  const double kTol = 1.0e-3;
  int64_t samples = 0;
  for (int64_t r = begin; r < end; ++r) {
    const int64_t proj = r / c->npx;
    const int64_t off = r - c->npx * proj;
    const int64_t row = off / c->cols0;
    const int64_t col = off - (int64_t)c->cols0 * row;
    const orc_cam_d* cam = &c->cams[c->cam_for_proj[proj]];
    double det[3];
    orc_cam_ind_to_det_f64(cam, (double)col, (double)row, det);
    double M[16];
    affine_mul_f64(c->p2i, &c->xforms[16 * proj], M);
    double ph[3], dp[3], d[3];
    affine_pt_f64(M, cam->pinhole, ph);
    affine_pt_f64(M, det, dp);
    d[0] = dp[0] - ph[0];
    d[1] = dp[1] - ph[1];
    d[2] = dp[2] - ph[2];
    double acc = (c->kernel == 0) ? 0.0 : -DBL_MAX;
    double t0 = 0.0, t1 = 0.0;
    const int32_t hit = orc_ray_rect_f64(aabb_min, c->aabb_max, ph, d, 1, &t0, &t1);
    if (c->steps) c->steps[r] = -1;
    if (hit && ((t1 - t0) > 2.0 * kTol)) {
      t0 += kTol;
      t1 -= kTol;
      const double start[3] = {ph[0] + t0 * d[0], ph[1] + t0 * d[1], ph[2] + t0 * d[2]};
      const double dlen = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      const double L = (t1 - t0) * dlen;
      double v[3] = {det[0] - cam->pinhole[0], det[1] - cam->pinhole[1], det[2] - cam->pinhole[2]};
      const double vn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      if (vn > 0.0) {
        v[0] /= vn;
        v[1] /= vn;
        v[2] /= vn;
      }
      v[0] *= c->step;
      v[1] *= c->step;
      v[2] *= c->step;
      double sv[3];
      for (int i = 0; i < 3; ++i) sv[i] = M[4 * i] * v[0] + M[4 * i + 1] * v[1] + M[4 * i + 2] * v[2];
      const double s = sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]);
      const int64_t n = (int64_t)(L / s);
      if (c->steps) c->steps[r] = (int32_t)n;
      if (c->steps_only) continue;
      const double scale = s / dlen;
      const double dx = d[0] * scale, dy = d[1] * scale, dz = d[2] * scale;
      for (int64_t i = 0; i <= n; ++i) {
        /* positions evaluated directly (no accumulated drift) */
        const double p[3] = {start[0] + (double)i * dx, start[1] + (double)i * dy,
                             start[2] + (double)i * dz};
        const double val = orc_linear_interp_f64(c->vol, c->dims, p);
        if (c->kernel == 0) {
          acc += val;
        } else if (acc < val) {
          acc = val;
        }
      }
      samples += n + 1;
      acc *= c->step;
    }
    if (c->steps_only) continue;
    if (c->kernel == 0) {
      c->out[r] += acc;
    } else if (c->out[r] < acc) {
      c->out[r] = acc;
    }
  }
  __atomic_fetch_add(&c->samples, samples, __ATOMIC_RELAXED);
}

void orc_drr_f64(const double* vol, const int64_t dims[3], const double p2i[16],
                 const orc_cam_d* cams, int32_t num_cams,
                 const double* xforms, const int64_t* cam_for_proj, int64_t num_projs,
                 double step, int32_t kernel, int32_t store, const double* bg, double default_bg,
                 double* out, int32_t threads, int64_t* samples_out)
{
  (void)num_cams;
  drr_f64_ctx c;
  memset(&c, 0, sizeof(c));
  c.vol = vol;
  c.dims = dims;
  c.aabb_max[0] = (double)(dims[0] - 1);
  c.aabb_max[1] = (double)(dims[1] - 1);
  c.aabb_max[2] = (double)(dims[2] - 1);
  c.p2i = p2i;
  c.cams = cams;
  c.xforms = xforms;
  c.cam_for_proj = cam_for_proj;
  c.step = step;
  c.kernel = kernel;
  c.cols0 = cams[0].cols;
  c.npx = (int64_t)cams[0].rows * cams[0].cols;
  c.out = out;
  if (bg) {
    for (int64_t p = 0; p < num_projs; ++p) {
      memcpy(out + p * c.npx, bg + cam_for_proj[p] * c.npx, sizeof(double) * (size_t)c.npx);
    }
  }
  if (store == 0 && !bg) {
    for (int64_t i = 0; i < num_projs * c.npx; ++i) out[i] = default_bg;
  }
  orc_parallel_for(drr_f64_range, &c, num_projs * c.npx, 2048, threads);
  if (samples_out) *samples_out = c.samples;
}

/* Per-ray step counts: the ray set-up of orc_drr_f32 / orc_drr_f64 alone (no samples), writing
 * every ray's last sample index n = floor(chord / step_len) (xregRayCastLineIntCPU.cpp:253-254),
 * -1 where the ray misses, to steps_out[proj][row][col]. */
void orc_ray_steps_f32(const int64_t dims[3], const float p2i[16], const orc_cam_f* cams,
                       const float* xforms, const int64_t* cam_for_proj, int64_t num_projs,
                       float step, int32_t* steps_out, int32_t threads)
{
  drr_f32_ctx c;
  memset(&c, 0, sizeof(c));
  c.dims = dims;
  c.aabb_max[0] = (float)(dims[0] - 1);
  c.aabb_max[1] = (float)(dims[1] - 1);
  c.aabb_max[2] = (float)(dims[2] - 1);
  c.p2i = p2i;
  c.cams = cams;
  c.xforms = xforms;
  c.cam_for_proj = cam_for_proj;
  c.step = step;
  c.cols0 = cams[0].cols;
  c.npx = (int64_t)cams[0].rows * cams[0].cols;
  c.steps = steps_out;
  c.steps_only = 1;
  orc_parallel_for(drr_f32_range, &c, num_projs * c.npx, 2048, threads);
}

void orc_ray_steps_f64(const int64_t dims[3], const double p2i[16], const orc_cam_d* cams,
                       const double* xforms, const int64_t* cam_for_proj, int64_t num_projs,
                       double step, int32_t* steps_out, int32_t threads)
{
  drr_f64_ctx c;
  memset(&c, 0, sizeof(c));
  c.dims = dims;
  c.aabb_max[0] = (double)(dims[0] - 1);
  c.aabb_max[1] = (double)(dims[1] - 1);
  c.aabb_max[2] = (double)(dims[2] - 1);
  c.p2i = p2i;
  c.cams = cams;
  c.xforms = xforms;
  c.cam_for_proj = cam_for_proj;
  c.step = step;
  c.cols0 = cams[0].cols;
  c.npx = (int64_t)cams[0].rows * cams[0].cols;
  c.steps = steps_out;
  c.steps_only = 1;
  orc_parallel_for(drr_f64_range, &c, num_projs * c.npx, 2048, threads);
}

/* ------------------------------------------------------------------------------------ */
/* Eigen 3.3.4 redux_impl<sum, LinearVectorizedTraversal, NoUnrolling> with SSE Packet4f
 * (Eigen/src/Core/Redux.h): two 4-lane accumulators over 8-element blocks, one trailing
 * packet, predux = (a0 + a2) + (a1 + a3) (Eigen/src/Core/arch/SSE/PacketMath.h), then the
 * scalar tail. Assumes a 16-byte aligned start. parity unpinned (Eigen absent). */
float orc_eigen_sum_f32(const float* x, int64_t n)
{
  const int64_t ps = 4;
  const int64_t aligned2 = (n / (2 * ps)) * (2 * ps);
  const int64_t aligned = (n / ps) * ps;
  float res;
  if (aligned) {
    float p0[4] = {x[0], x[1], x[2], x[3]};
    if (aligned > ps) {
      float p1[4] = {x[4], x[5], x[6], x[7]};
      for (int64_t i = 2 * ps; i < aligned2; i += 2 * ps) {
        for (int l = 0; l < 4; ++l) {
          p0[l] = p0[l] + x[i + l];
          p1[l] = p1[l] + x[i + ps + l];
        }
      }
      for (int l = 0; l < 4; ++l) p0[l] = p0[l] + p1[l];
      if (aligned > aligned2) {
        for (int l = 0; l < 4; ++l) p0[l] = p0[l] + x[aligned2 + l];
      }
    }
    res = (p0[0] + p0[2]) + (p0[1] + p0[3]);
    for (int64_t i = aligned; i < n; ++i) res = res + x[i];
  } else {
    res = x[0];
    for (int64_t i = 1; i < n; ++i) res = res + x[i];
  }
  return res;
}

static double sum_f64(const double* x, int64_t n)
{
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += x[i];
  return s;
}

/* ComputeImage2DMeanStdDev (xregImgSimMetric2DNCCCPU.cpp:52-64): mean = sum/n,
 * std = max(1e-6, sqrt(sum((x-mean)^2)/(n-1))). */
static void mean_std_f32(const float* x, int64_t n, float* tmp, float* mean_out, float* std_out)
{
  const float mean = orc_eigen_sum_f32(x, n) / (float)n;
  for (int64_t i = 0; i < n; ++i) {
    const float t = x[i] - mean;
    tmp[i] = t * t;
  }
  const float var_sum = orc_eigen_sum_f32(tmp, n);
  const float s = sqrtf(var_sum / (float)(n - 1));
  *mean_out = mean;
  *std_out = (s > 1.0e-6f) ? s : 1.0e-6f;
}

static void mean_std_masked_f32(const float* x, const uint8_t* m, int64_t n, int64_t len,
                                float* mean_out, float* std_out)
{
  /* ComputeImage2DMeanStdDevWithMask (xregImgSimMetric2DNCCCPU.cpp:77-115) */
  float mean = 0.0f;
  for (int64_t i = 0; i < n; ++i) {
    if (m[i]) mean += x[i];
  }
  mean /= (float)len;
  float sd = 0.0f;
  for (int64_t i = 0; i < n; ++i) {
    if (m[i]) {
      const float t = x[i] - mean;
      sd += t * t;
    }
  }
  sd = sqrtf(sd / (float)(len - 1));
  *mean_out = mean;
  *std_out = (sd > 1.0e-6f) ? sd : 1.0e-6f;
}

void orc_zero_mean_inplace_f32(float* img, int64_t npix)
{
  float* tmp = (float*)malloc(sizeof(float) * (size_t)npix);
  float mean, sd;
  mean_std_f32(img, npix, tmp, &mean, &sd);
  for (int64_t i = 0; i < npix; ++i) img[i] = img[i] - mean;
  free(tmp);
}

void orc_ncc_f32(const float* fixed, const float* mov, int64_t npix, int64_t num_mov,
                 const uint8_t* mask, float* sim_out)
{
  /* ImgSimMetric2DNCCCPU::process_mask + compute (xregImgSimMetric2DNCCCPU.cpp:164-236) */
  float* zf = (float*)malloc(sizeof(float) * (size_t)npix);
  float* zm = (float*)malloc(sizeof(float) * (size_t)npix);
  float* tmp = (float*)malloc(sizeof(float) * (size_t)npix);
  float fmean, fsd;
  int64_t len = 0;
  if (!mask) {
    mean_std_f32(fixed, npix, tmp, &fmean, &fsd);
  } else {
    for (int64_t i = 0; i < npix; ++i) len += mask[i] ? 1 : 0;
    mean_std_masked_f32(fixed, mask, npix, len, &fmean, &fsd);
  }
  for (int64_t i = 0; i < npix; ++i) zf[i] = fixed[i] - fmean;
  for (int64_t m = 0; m < num_mov; ++m) {
    const float* x = mov + m * npix;
    float mmean, msd, sim;
    if (!mask) {
      mean_std_f32(x, npix, tmp, &mmean, &msd);
      for (int64_t i = 0; i < npix; ++i) zm[i] = x[i] - mmean;
      for (int64_t i = 0; i < npix; ++i) tmp[i] = zf[i] * zm[i];
      const float dot = orc_eigen_sum_f32(tmp, npix);
      sim = dot / (((float)npix * fsd) * msd);
    } else {
      mean_std_masked_f32(x, mask, npix, len, &mmean, &msd);
      sim = 0.0f;
      for (int64_t i = 0; i < npix; ++i) {
        if (mask[i]) sim += zf[i] * (x[i] - mmean);
      }
      sim /= ((float)len * fsd) * msd;
    }
    sim_out[m] = (1.0f - sim) * 0.5f;                                          /* :204 */
  }
  free(zf);
  free(zm);
  free(tmp);
}

void orc_ncc_f64(const double* fixed, const double* mov, int64_t npix, int64_t num_mov,
                 const uint8_t* mask, double* sim_out)
{
  int64_t len = npix;
  if (mask) {
    len = 0;
    for (int64_t i = 0; i < npix; ++i) len += mask[i] ? 1 : 0;
  }
  double fm = 0.0;
  for (int64_t i = 0; i < npix; ++i) {
    if (!mask || mask[i]) fm += fixed[i];
  }
  fm /= (double)len;
  double fv = 0.0;
  for (int64_t i = 0; i < npix; ++i) {
    if (!mask || mask[i]) fv += (fixed[i] - fm) * (fixed[i] - fm);
  }
  double fsd = sqrt(fv / (double)(len - 1));
  if (fsd < 1.0e-6) fsd = 1.0e-6;
  for (int64_t m = 0; m < num_mov; ++m) {
    const double* x = mov + m * npix;
    double mm = 0.0;
    for (int64_t i = 0; i < npix; ++i) {
      if (!mask || mask[i]) mm += x[i];
    }
    mm /= (double)len;
    double mv = 0.0, dot = 0.0;
    for (int64_t i = 0; i < npix; ++i) {
      if (!mask || mask[i]) {
        mv += (x[i] - mm) * (x[i] - mm);
        dot += (fixed[i] - fm) * (x[i] - mm);
      }
    }
    double msd = sqrt(mv / (double)(len - 1));
    if (msd < 1.0e-6) msd = 1.0e-6;
    sim_out[m] = (1.0 - dot / ((double)len * fsd * msd)) * 0.5;
  }
}

/* ------------------------------------------------------------------------------------ */
/* Gradient images: ImgSimMetric2DGradImgCPU (xregImgSimMetric2DGradImgCPU.cpp:32-102)
 * calls cv::GaussianBlur(img, Size(k,k), 0, 0) then cv::Sobel(img, -1, 1,0) / (0,1).
 * OpenCV 3.4: Sobel ksize 3 = separable [-1,0,1] x [1,2,1]; GaussianBlur with sigma 0 and
 * odd ksize <= 7 uses the fixed table (ksize 5: [1,4,6,4,1]/16, ksize 3: [1,2,1]/4);
 * the default border is BORDER_REFLECT_101; the row filter runs before the column filter.
 * parity unpinned: OpenCV is absent (its IPP path for float blur is not restated). */

static inline int32_t reflect101(int32_t p, int32_t n)
{
  if (n == 1) return 0;
  while (p < 0 || p >= n) {
    if (p < 0) p = -p;
    if (p >= n) p = 2 * (n - 1) - p;
  }
  return p;
}

static const float* gauss_table_f32(int32_t ksize)
{
  static const float k1[1] = {1.0f};
  static const float k3[3] = {0.25f, 0.5f, 0.25f};
  static const float k5[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
  static const float k7[7] = {0.03125f, 0.109375f, 0.21875f, 0.28125f, 0.21875f, 0.109375f, 0.03125f};
  switch (ksize) {
    case 1: return k1;
    case 3: return k3;
    case 5: return k5;
    case 7: return k7;
    default: return NULL;
  }
}

static void blur_f32(const float* src, int32_t rows, int32_t cols, int32_t ksize, float* dst)
{
  const float* k = gauss_table_f32(ksize);
  const int32_t r = ksize / 2;
  float* tmp = (float*)malloc(sizeof(float) * (size_t)rows * cols);
  /* row pass: symmetric kernel, k[r]*s0 + sum_j k[r+j]*(s[-j] + s[+j]) */
  for (int32_t y = 0; y < rows; ++y) {
    const float* s = src + (int64_t)y * cols;
    for (int32_t x = 0; x < cols; ++x) {
      float acc = k[r] * s[x];
      for (int32_t j = 1; j <= r; ++j) {
        acc = acc + k[r + j] * (s[reflect101(x - j, cols)] + s[reflect101(x + j, cols)]);
      }
      tmp[(int64_t)y * cols + x] = acc;
    }
  }
  for (int32_t y = 0; y < rows; ++y) {
    for (int32_t x = 0; x < cols; ++x) {
      float acc = k[r] * tmp[(int64_t)y * cols + x];
      for (int32_t j = 1; j <= r; ++j) {
        acc = acc + k[r + j] * (tmp[(int64_t)reflect101(y - j, rows) * cols + x] +
                                tmp[(int64_t)reflect101(y + j, rows) * cols + x]);
      }
      dst[(int64_t)y * cols + x] = acc;
    }
  }
  free(tmp);
}

void orc_sobel_f32(const float* img, int32_t rows, int32_t cols, int32_t smooth_ksize,
                   float* gx, float* gy)
{
  const float* src = img;
  float* sm = NULL;
  if (smooth_ksize > 1) {
    sm = (float*)malloc(sizeof(float) * (size_t)rows * cols);
    blur_f32(img, rows, cols, smooth_ksize, sm);
    src = sm;
  }
  float* tmp = (float*)malloc(sizeof(float) * (size_t)rows * cols);
  /* gx: row derivative s[x+1] - s[x-1], then column smoothing (t[-1] + t[+1]) + 2 t[0] */
  for (int32_t y = 0; y < rows; ++y) {
    const float* s = src + (int64_t)y * cols;
    for (int32_t x = 0; x < cols; ++x) {
      tmp[(int64_t)y * cols + x] = s[reflect101(x + 1, cols)] - s[reflect101(x - 1, cols)];
    }
  }
  for (int32_t y = 0; y < rows; ++y) {
    const float* a = tmp + (int64_t)reflect101(y - 1, rows) * cols;
    const float* b = tmp + (int64_t)y * cols;
    const float* c = tmp + (int64_t)reflect101(y + 1, rows) * cols;
    for (int32_t x = 0; x < cols; ++x) gx[(int64_t)y * cols + x] = (a[x] + c[x]) + b[x] * 2.0f;
  }
  /* gy: row smoothing (s[-1] + s[+1]) + 2 s[0], then column derivative t[+1] - t[-1] */
  for (int32_t y = 0; y < rows; ++y) {
    const float* s = src + (int64_t)y * cols;
    for (int32_t x = 0; x < cols; ++x) {
      tmp[(int64_t)y * cols + x] = (s[reflect101(x - 1, cols)] + s[reflect101(x + 1, cols)]) + s[x] * 2.0f;
    }
  }
  for (int32_t y = 0; y < rows; ++y) {
    const float* a = tmp + (int64_t)reflect101(y - 1, rows) * cols;
    const float* c = tmp + (int64_t)reflect101(y + 1, rows) * cols;
    for (int32_t x = 0; x < cols; ++x) gy[(int64_t)y * cols + x] = c[x] - a[x];
  }
  free(tmp);
  free(sm);
}

static void blur_f64(const double* src, int32_t rows, int32_t cols, int32_t ksize, double* dst)
{
  const float* kf = gauss_table_f32(ksize);
  const int32_t r = ksize / 2;
  double k[7];
  for (int i = 0; i < ksize; ++i) k[i] = (double)kf[i];
  double* tmp = (double*)malloc(sizeof(double) * (size_t)rows * cols);
  for (int32_t y = 0; y < rows; ++y) {
    for (int32_t x = 0; x < cols; ++x) {
      double acc = 0.0;
      for (int32_t j = -r; j <= r; ++j) acc += k[r + j] * src[(int64_t)y * cols + reflect101(x + j, cols)];
      tmp[(int64_t)y * cols + x] = acc;
    }
  }
  for (int32_t y = 0; y < rows; ++y) {
    for (int32_t x = 0; x < cols; ++x) {
      double acc = 0.0;
      for (int32_t j = -r; j <= r; ++j) acc += k[r + j] * tmp[(int64_t)reflect101(y + j, rows) * cols + x];
      dst[(int64_t)y * cols + x] = acc;
    }
  }
  free(tmp);
}

void orc_sobel_f64(const double* img, int32_t rows, int32_t cols, int32_t smooth_ksize,
                   double* gx, double* gy)
{
  const double* src = img;
  double* sm = NULL;
  if (smooth_ksize > 1) {
    sm = (double*)malloc(sizeof(double) * (size_t)rows * cols);
    blur_f64(img, rows, cols, smooth_ksize, sm);
    src = sm;
  }
  for (int32_t y = 0; y < rows; ++y) {
    const int32_t ym = reflect101(y - 1, rows), yp = reflect101(y + 1, rows);
    for (int32_t x = 0; x < cols; ++x) {
      const int32_t xm = reflect101(x - 1, cols), xp = reflect101(x + 1, cols);
#define S(yy, xx) src[(int64_t)(yy) * cols + (xx)]
      gx[(int64_t)y * cols + x] = (S(ym, xp) - S(ym, xm)) + 2.0 * (S(y, xp) - S(y, xm)) + (S(yp, xp) - S(yp, xm));
      gy[(int64_t)y * cols + x] = (S(yp, xm) - S(ym, xm)) + 2.0 * (S(yp, x) - S(ym, x)) + (S(yp, xp) - S(ym, xp));
#undef S
    }
  }
  free(sm);
}

void orc_grad_ncc_f32(const float* fixed, const float* mov, int32_t rows, int32_t cols,
                      int64_t num_mov, int32_t smooth_ksize, float* sim_out)
{
  /* ImgSimMetric2DGradNCCCPU::compute (xregImgSimMetric2DGradNCCCPU.cpp:51-65):
   * NCC on gx and on gy, sim = 0.5 * (ncc_x + ncc_y). */
  const int64_t npix = (int64_t)rows * cols;
  float* fgx = (float*)malloc(sizeof(float) * (size_t)npix);
  float* fgy = (float*)malloc(sizeof(float) * (size_t)npix);
  float* mgx = (float*)malloc(sizeof(float) * (size_t)(npix * num_mov));
  float* mgy = (float*)malloc(sizeof(float) * (size_t)(npix * num_mov));
  float* sx = (float*)malloc(sizeof(float) * (size_t)num_mov);
  float* sy = (float*)malloc(sizeof(float) * (size_t)num_mov);
  orc_sobel_f32(fixed, rows, cols, smooth_ksize, fgx, fgy);
  for (int64_t m = 0; m < num_mov; ++m) {
    orc_sobel_f32(mov + m * npix, rows, cols, smooth_ksize, mgx + m * npix, mgy + m * npix);
  }
  orc_ncc_f32(fgx, mgx, npix, num_mov, NULL, sx);
  orc_ncc_f32(fgy, mgy, npix, num_mov, NULL, sy);
  for (int64_t m = 0; m < num_mov; ++m) sim_out[m] = (float)(0.5 * (double)(sx[m] + sy[m]));
  free(fgx);
  free(fgy);
  free(mgx);
  free(mgy);
  free(sx);
  free(sy);
}

void orc_grad_ncc_f64(const double* fixed, const double* mov, int32_t rows, int32_t cols,
                      int64_t num_mov, int32_t smooth_ksize, double* sim_out)
{
  const int64_t npix = (int64_t)rows * cols;
  double* fgx = (double*)malloc(sizeof(double) * (size_t)npix);
  double* fgy = (double*)malloc(sizeof(double) * (size_t)npix);
  double* mgx = (double*)malloc(sizeof(double) * (size_t)(npix * num_mov));
  double* mgy = (double*)malloc(sizeof(double) * (size_t)(npix * num_mov));
  double* sx = (double*)malloc(sizeof(double) * (size_t)num_mov);
  double* sy = (double*)malloc(sizeof(double) * (size_t)num_mov);
  orc_sobel_f64(fixed, rows, cols, smooth_ksize, fgx, fgy);
  for (int64_t m = 0; m < num_mov; ++m) {
    orc_sobel_f64(mov + m * npix, rows, cols, smooth_ksize, mgx + m * npix, mgy + m * npix);
  }
  orc_ncc_f64(fgx, mgx, npix, num_mov, NULL, sx);
  orc_ncc_f64(fgy, mgy, npix, num_mov, NULL, sy);
  for (int64_t m = 0; m < num_mov; ++m) sim_out[m] = 0.5 * (sx[m] + sy[m]);
  free(fgx);
  free(fgy);
  free(mgx);
  free(mgy);
  free(sx);
  free(sy);
}

/* ------------------------------------------------------------------------------------ */
/* Patch NCC: ImgSimMetric2DPatchCommon::setup_patches (xregImgSimMetric2DPatchCommon.cpp:
 * 256-297: centres r .. size-1-r, row-major, stride s), detail::ComputePatchMeanStdDev
 * (xregImgSimMetric2DPatchNCCCPU.cpp:558-619), fixed preprocessing (:398-409) and the
 * per-patch score + combine (:190-287) with unit weights. */

typedef struct {
  const float* fixed;
  const float* mov;
  int32_t cols;
  int32_t r;
  int32_t diam;
  int32_t ppr; /* patches per row */
  int32_t stride;
  float* fhat; /* num_patches * diam * diam */
  float* vals;         /* per used index */
  const uint8_t* mask; /* optional fixed-image mask (non-zero = use the pixel) */
  const float* wgt;    /* per-patch weights, or NULL for unit weights */
  const int64_t* inds; /* patch indices used, or NULL for 0 .. num_patches - 1 */
  int32_t wic;         /* weight_patch_sims_in_combine_ */
  int32_t mstats;      /* use_mask_for_patch_stats_ (with a mask) */
} patch_f32_ctx;

/* ComputePatchMeanStdDev (:558-619): over the whole patch, or over the pixels `mask` keeps
 * (mask non-NULL: use_mask_for_patch_stats_); the sum is divided only when n > 1 */
static void patch_stats_f32(const float* img, int32_t cols, int32_t r0, int32_t c0, int32_t diam,
                            const uint8_t* mask, float* mean_out, float* std_out, int64_t* n_out)
{
  float mean = 0.0f;
  int64_t n = 0;
  for (int32_t y = 0; y < diam; ++y) {
    for (int32_t x = 0; x < diam; ++x) {
      const int64_t o = (int64_t)(r0 + y) * cols + c0 + x;
      if (!mask || mask[o]) {
        mean += img[o];
        ++n;
      }
    }
  }
  float var = 0.0f;
  if (n > 1) {
    mean /= (float)n;
    for (int32_t y = 0; y < diam; ++y) {
      for (int32_t x = 0; x < diam; ++x) {
        const int64_t o = (int64_t)(r0 + y) * cols + c0 + x;
        if (!mask || mask[o]) {
          const float t = img[o] - mean;
          var += t * t;
        }
      }
    }
    var /= (float)n - 1.0f;
  }
  const float s = sqrtf(var);
  *mean_out = mean;
  *std_out = (s > 1.0e-6f) ? s : 1.0e-6f;
  *n_out = n;
}

static void patch_fixed_range(void* vctx, int64_t begin, int64_t end)
{
  patch_f32_ctx* c = (patch_f32_ctx*)vctx;
  const int32_t d = c->diam;
  for (int64_t p = begin; p < end; ++p) {
    const int32_t r0 = (int32_t)(p / c->ppr) * c->stride;
    const int32_t c0 = (int32_t)(p % c->ppr) * c->stride;
    float mean, sd;
    int64_t n;
    patch_stats_f32(c->fixed, c->cols, r0, c0, d, c->mstats ? c->mask : NULL, &mean, &sd, &n);
    const float denom = sd * (float)n;
    float* dst = c->fhat + p * d * d;
    for (int32_t y = 0; y < d; ++y) {
      for (int32_t x = 0; x < d; ++x) {
        dst[y * d + x] = (c->fixed[(int64_t)(r0 + y) * c->cols + c0 + x] - mean) / denom;
      }
    }
  }
}

static void patch_mov_range(void* vctx, int64_t begin, int64_t end)
{
  patch_f32_ctx* c = (patch_f32_ctx*)vctx;
  const int32_t d = c->diam;
  for (int64_t k = begin; k < end; ++k) {
    const int64_t p = c->inds ? c->inds[k] : k; /* local -> global patch index (:204-206) */
    const int32_t r0 = (int32_t)(p / c->ppr) * c->stride;
    const int32_t c0 = (int32_t)(p % c->ppr) * c->stride;
    const float w = c->wgt ? c->wgt[p] : 1.0f;
    if (c->wic && !(fabsf(w) > 1.0e-6f)) {
      c->vals[k] = 0.0f; /* low-weight patch skipped (:212) */
      continue;
    }
    float mean, sd;
    int64_t n;
    /* patch statistics over the whole patch, or its unmasked pixels (use_mask_for_patch_stats_) */
    patch_stats_f32(c->mov, c->cols, r0, c0, d, c->mstats ? c->mask : NULL, &mean, &sd, &n);
    const float* fh = c->fhat + p * d * d;
    float accum = 0.0f;
    for (int32_t y = 0; y < d; ++y) {
      for (int32_t x = 0; x < d; ++x) {
        const int64_t o = (int64_t)(r0 + y) * c->cols + c0 + x;
        if (!c->mask || c->mask[o]) accum += ((c->mov[o] - mean) / sd) * fh[y * d + x];
      }
    }
    /* weighted by the patch weight when weight_patch_sims_in_combine_ (:252-253) */
    c->vals[k] = (c->wic ? w : 1.0f) * (1.0f - accum);
  }
}

/* ImgSimMetric2DPatchCommon::compute_weights with a mask and the default options
 * (xregImgSimMetric2DPatchCommon.cpp:306-400): fraction of unmasked pixels in the patch, then
 * normalised to sum to one (normalize_weights_as_prob_), all in float. */
static void patch_mask_weights_f32(const uint8_t* mask, int32_t cols, int32_t ppr, int64_t np, int32_t stride,
                                   int32_t d, float* w)
{
  float wsum = 0.0f;
  for (int64_t p = 0; p < np; ++p) {
    const int32_t r0 = (int32_t)(p / ppr) * stride, c0 = (int32_t)(p % ppr) * stride;
    int64_t cnt = 0;
    for (int32_t y = 0; y < d; ++y)
      for (int32_t x = 0; x < d; ++x) cnt += mask[(int64_t)(r0 + y) * cols + c0 + x] ? 1 : 0;
    w[p] = (float)cnt / (float)((int64_t)d * d);
  }
  for (int64_t p = 0; p < np; ++p) wsum += w[p];
  for (int64_t p = 0; p < np; ++p) w[p] /= wsum;
}

void orc_patch_ncc_f32(const float* fixed, const float* mov, int32_t rows, int32_t cols,
                       int64_t num_mov, int32_t patch_radius, int32_t patch_stride,
                       float* sim_out, int32_t threads)
{
  orc_patch_ncc_masked_f32(fixed, mov, NULL, rows, cols, num_mov, patch_radius, patch_stride, sim_out, threads);
}

void orc_patch_ncc_masked_f32(const float* fixed, const float* mov, const uint8_t* mask, int32_t rows,
                              int32_t cols, int64_t num_mov, int32_t patch_radius, int32_t patch_stride,
                              float* sim_out, int32_t threads)
{
  float* wgt = NULL;
  if (mask) {
    const int32_t diam = 2 * patch_radius + 1;
    const int32_t ppr = (cols - 1 - 2 * patch_radius) / patch_stride + 1;
    const int32_t ppc = (rows - 1 - 2 * patch_radius) / patch_stride + 1;
    const int64_t np = (int64_t)ppr * ppc;
    wgt = (float*)malloc(sizeof(float) * (size_t)np);
    patch_mask_weights_f32(mask, cols, ppr, np, patch_stride, diam, wgt);
  }
  orc_patch_ncc_weighted_f32(fixed, mov, mask, rows, cols, num_mov, patch_radius, patch_stride, wgt, NULL, 0, 1, 0,
                             0, sim_out, threads);
  free(wgt);
}

void orc_patch_weights_f32(const float* wgt_img, const uint8_t* mask, int32_t rows, int32_t cols,
                           int32_t patch_radius, int32_t patch_stride, int32_t use_mask_for_weighting,
                           int32_t normalize, float* w_out)
{
  const int32_t diam = 2 * patch_radius + 1;
  const int32_t ppr = (cols - 1 - 2 * patch_radius) / patch_stride + 1;
  const int32_t ppc = (rows - 1 - 2 * patch_radius) / patch_stride + 1;
  const int64_t np = (int64_t)ppr * ppc;
  const int use_mask_wgts = use_mask_for_weighting && mask;
  (void)rows;
  if (!wgt_img && !use_mask_wgts) {
    for (int64_t p = 0; p < np; ++p) w_out[p] = 1.0f; /* otherwise keep weight at 1 (:384) */
    return;
  }
  if (wgt_img) {
    /* the weight image's value at the patch centre, zero where the mask is off there when
     * the mask is used for weighting (:318-342) */
    for (int64_t p = 0; p < np; ++p) {
      const int64_t cr = patch_radius + (p / ppr) * patch_stride, cc = patch_radius + (p % ppr) * patch_stride;
      w_out[p] = wgt_img[cr * cols + cc];
      if (use_mask_wgts && !mask[cr * cols + cc]) w_out[p] = 0.0f;
    }
  } else {
    /* fraction of unmasked pixels (:344-368) */
    for (int64_t p = 0; p < np; ++p) {
      const int32_t r0 = (int32_t)(p / ppr) * patch_stride, c0 = (int32_t)(p % ppr) * patch_stride;
      int64_t cnt = 0;
      for (int32_t y = 0; y < diam; ++y)
        for (int32_t x = 0; x < diam; ++x) cnt += mask[(int64_t)(r0 + y) * cols + c0 + x] ? 1 : 0;
      w_out[p] = (float)cnt / (float)((int64_t)diam * diam);
    }
  }
  if (normalize) { /* normalize_weights_as_prob_ (:370-381), float sum in patch order */
    float wsum = 0.0f;
    for (int64_t p = 0; p < np; ++p) wsum += w_out[p];
    for (int64_t p = 0; p < np; ++p) w_out[p] /= wsum;
  }
}

void orc_patch_variance_image_f32(const float* img, const uint8_t* mask, int32_t rows, int32_t cols,
                                  int32_t patch_radius, int32_t patch_stride, float* out)
{
  /* ImgSimMetric2DPatchNCCCPU::process_mask with use_fixed_img_patch_variances_as_wgts_
   * (xregImgSimMetric2DPatchNCCCPU.cpp:331-440): ComputePatchMeanStdDev (:558-619) over the
   * unmasked pixels of the patch when there is a mask, std clamped at 1e-6, squared, at the
   * patch centre; zero elsewhere */
  const int32_t diam = 2 * patch_radius + 1;
  const int32_t ppr = (cols - 1 - 2 * patch_radius) / patch_stride + 1;
  const int32_t ppc = (rows - 1 - 2 * patch_radius) / patch_stride + 1;
  for (int64_t i = 0; i < (int64_t)rows * cols; ++i) out[i] = 0.0f;
  for (int64_t p = 0; p < (int64_t)ppr * ppc; ++p) {
    const int32_t r0 = (int32_t)(p / ppr) * patch_stride, c0 = (int32_t)(p % ppr) * patch_stride;
    float mean = 0.0f;
    int64_t n = 0;
    for (int32_t y = 0; y < diam; ++y)
      for (int32_t x = 0; x < diam; ++x) {
        const int64_t o = (int64_t)(r0 + y) * cols + c0 + x;
        if (!mask || mask[o]) {
          mean += img[o];
          ++n;
        }
      }
    float var = 0.0f;
    if (n > 1) {
      mean /= (float)n;
      for (int32_t y = 0; y < diam; ++y)
        for (int32_t x = 0; x < diam; ++x) {
          const int64_t o = (int64_t)(r0 + y) * cols + c0 + x;
          if (!mask || mask[o]) {
            const float t = img[o] - mean;
            var += t * t;
          }
        }
      var /= (float)n - 1.0f;
    }
    const float sd = sqrtf(var) > 1.0e-6f ? sqrtf(var) : 1.0e-6f;
    out[(int64_t)(r0 + patch_radius) * cols + c0 + patch_radius] = sd * sd;
  }
}

static void patch_ncc_core_f32(const float* fixed, const float* mov, const uint8_t* mask, int32_t rows,
                               int32_t cols, int64_t num_mov, int32_t patch_radius, int32_t patch_stride,
                               const float* wgt, const int64_t* inds, int64_t num_inds, int32_t weight_in_combine,
                               int32_t compute_mean, int32_t mask_stats, float* sim_out, float* vals_out,
                               int32_t threads);

void orc_patch_ncc_weighted_f32(const float* fixed, const float* mov, const uint8_t* mask, int32_t rows,
                                int32_t cols, int64_t num_mov, int32_t patch_radius, int32_t patch_stride,
                                const float* wgt, const int64_t* inds, int64_t num_inds, int32_t weight_in_combine,
                                int32_t compute_mean, int32_t mask_stats, float* sim_out, int32_t threads)
{
  patch_ncc_core_f32(fixed, mov, mask, rows, cols, num_mov, patch_radius, patch_stride, wgt, inds, num_inds,
                     weight_in_combine, compute_mean, mask_stats, sim_out, NULL, threads);
}

void orc_patch_scores_f32(int32_t metric, const float* fixed, const float* mov, const uint8_t* mask, int32_t rows,
                          int32_t cols, int64_t num_mov, int32_t smooth_ksize, int32_t patch_radius,
                          int32_t patch_stride, int32_t mask_stats, float* scores_out, int32_t threads)
{
  /* sim_vals_for_each_patch_: every patch's unweighted score 1 - accum
   * (xregImgSimMetric2DPatchNCCCPU.cpp:228-250); Patch-Grad-NCC 0.5 (x + y) in float
   * (xregImgSimMetric2DPatchGradNCCCPU.cpp:226-240) */
  const int32_t ppr = (cols - 1 - 2 * patch_radius) / patch_stride + 1;
  const int32_t ppc = (rows - 1 - 2 * patch_radius) / patch_stride + 1;
  const int64_t np = (int64_t)ppr * ppc;
  const int64_t npix = (int64_t)rows * cols;
  float sim;
  if (metric == 2) {
    for (int64_t m = 0; m < num_mov; ++m) {
      patch_ncc_core_f32(fixed, mov + m * npix, mask, rows, cols, 1, patch_radius, patch_stride, NULL, NULL, 0, 0, 0,
                         mask_stats, &sim, scores_out + m * np, threads);
    }
    return;
  }
  float* fgx = (float*)malloc(sizeof(float) * (size_t)npix);
  float* fgy = (float*)malloc(sizeof(float) * (size_t)npix);
  float* mgx = (float*)malloc(sizeof(float) * (size_t)npix);
  float* mgy = (float*)malloc(sizeof(float) * (size_t)npix);
  float* vy = (float*)malloc(sizeof(float) * (size_t)np);
  orc_sobel_f32(fixed, rows, cols, smooth_ksize, fgx, fgy);
  for (int64_t m = 0; m < num_mov; ++m) {
    float* vx = scores_out + m * np;
    orc_sobel_f32(mov + m * npix, rows, cols, smooth_ksize, mgx, mgy);
    patch_ncc_core_f32(fgx, mgx, mask, rows, cols, 1, patch_radius, patch_stride, NULL, NULL, 0, 0, 0, mask_stats,
                       &sim, vx, threads);
    patch_ncc_core_f32(fgy, mgy, mask, rows, cols, 1, patch_radius, patch_stride, NULL, NULL, 0, 0, 0, mask_stats,
                       &sim, vy, threads);
    for (int64_t p = 0; p < np; ++p) vx[p] = (float)(0.5 * (double)(vx[p] + vy[p]));
  }
  free(fgx);
  free(fgy);
  free(mgx);
  free(mgy);
  free(vy);
}

static void patch_ncc_core_f32(const float* fixed, const float* mov, const uint8_t* mask, int32_t rows,
                               int32_t cols, int64_t num_mov, int32_t patch_radius, int32_t patch_stride,
                               const float* wgt, const int64_t* inds, int64_t num_inds, int32_t weight_in_combine,
                               int32_t compute_mean, int32_t mask_stats, float* sim_out, float* vals_out,
                               int32_t threads)
{
  const int32_t diam = 2 * patch_radius + 1;
  const int32_t ppr = (cols - 1 - 2 * patch_radius) / patch_stride + 1;
  const int32_t ppc = (rows - 1 - 2 * patch_radius) / patch_stride + 1;
  const int64_t np = (int64_t)ppr * ppc;
  const int64_t nu = inds ? num_inds : np;
  const int64_t npix = (int64_t)rows * cols;
  patch_f32_ctx c;
  c.fixed = fixed;
  c.cols = cols;
  c.r = patch_radius;
  c.diam = diam;
  c.ppr = ppr;
  c.stride = patch_stride;
  c.fhat = (float*)malloc(sizeof(float) * (size_t)(np * diam * diam));
  c.vals = (float*)malloc(sizeof(float) * (size_t)(nu > 0 ? nu : 1));
  c.mask = mask;
  c.wgt = wgt;
  c.inds = inds;
  c.wic = weight_in_combine;
  c.mstats = mask_stats && mask;
  orc_parallel_for(patch_fixed_range, &c, np, 256, threads);
  for (int64_t m = 0; m < num_mov; ++m) {
    c.mov = mov + m * npix;
    orc_parallel_for(patch_mov_range, &c, nu, 256, threads);
    if (vals_out) memcpy(vals_out + m * nu, c.vals, sizeof(float) * (size_t)nu);
    float sum = 0.0f;
    for (int64_t k = 0; k < nu; ++k) sum += c.vals[k];
    /* combine (:262-287): the mean over the used patches, else the weighted sum divided by
     * the float sum of the used patches' weights, else the plain sum */
    if (compute_mean) {
      sum /= (float)nu;
    } else if (weight_in_combine) {
      float tot_wgt = 0.0f;
      for (int64_t k = 0; k < nu; ++k) tot_wgt += wgt ? wgt[inds ? inds[k] : k] : 1.0f;
      sum /= tot_wgt;
    }
    sim_out[m] = sum;
  }
  free(c.fhat);
  free(c.vals);
}

typedef struct {
  const double* fixed;
  const double* mov;
  int32_t cols;
  int32_t diam;
  int32_t ppr;
  int32_t stride;
  double* fmean;
  double* fsd;
  double* vals;
  const uint8_t* mask;
  const double* wgt;
  const int64_t* inds;
  int32_t wic;
  int32_t mstats;
  double* fn; /* per patch: pixels the fixed statistics used */
} patch_f64_ctx;

static void patch_stats_f64(const double* img, int32_t cols, int32_t r0, int32_t c0, int32_t d,
                            const uint8_t* mask, double* mean_out, double* sd_out, double* n_out)
{
  double mean = 0.0, n = 0.0;
  for (int32_t y = 0; y < d; ++y) {
    for (int32_t x = 0; x < d; ++x) {
      const int64_t o = (int64_t)(r0 + y) * cols + c0 + x;
      if (!mask || mask[o]) {
        mean += img[o];
        n += 1.0;
      }
    }
  }
  double var = 0.0;
  if (n > 1.0) {
    mean /= n;
    for (int32_t y = 0; y < d; ++y) {
      for (int32_t x = 0; x < d; ++x) {
        const int64_t o = (int64_t)(r0 + y) * cols + c0 + x;
        if (!mask || mask[o]) {
          const double t = img[o] - mean;
          var += t * t;
        }
      }
    }
    var /= n - 1.0;
  }
  const double sd = sqrt(var);
  *mean_out = mean;
  *sd_out = (sd > 1.0e-6) ? sd : 1.0e-6;
  *n_out = n;
}

static void patch_fixed_range_f64(void* vctx, int64_t begin, int64_t end)
{
  patch_f64_ctx* c = (patch_f64_ctx*)vctx;
  for (int64_t p = begin; p < end; ++p) {
    const int32_t r0 = (int32_t)(p / c->ppr) * c->stride;
    const int32_t c0 = (int32_t)(p % c->ppr) * c->stride;
    patch_stats_f64(c->fixed, c->cols, r0, c0, c->diam, c->mstats ? c->mask : NULL, &c->fmean[p], &c->fsd[p],
                    &c->fn[p]);
  }
}

static void patch_mov_range_f64(void* vctx, int64_t begin, int64_t end)
{
  patch_f64_ctx* c = (patch_f64_ctx*)vctx;
  const int32_t d = c->diam;
  for (int64_t k = begin; k < end; ++k) {
    const int64_t p = c->inds ? c->inds[k] : k;
    const int32_t r0 = (int32_t)(p / c->ppr) * c->stride;
    const int32_t c0 = (int32_t)(p % c->ppr) * c->stride;
    const double w = c->wgt ? c->wgt[p] : 1.0;
    if (c->wic && !(fabs(w) > 1.0e-6)) {
      c->vals[k] = 0.0;
      continue;
    }
    double mean, sd, nm;
    patch_stats_f64(c->mov, c->cols, r0, c0, d, c->mstats ? c->mask : NULL, &mean, &sd, &nm);
    double acc = 0.0;
    for (int32_t y = 0; y < d; ++y) {
      for (int32_t x = 0; x < d; ++x) {
        const int64_t o = (int64_t)(r0 + y) * c->cols + c0 + x;
        if (!c->mask || c->mask[o]) acc += ((c->mov[o] - mean) / sd) * ((c->fixed[o] - c->fmean[p]) / (c->fsd[p] * c->fn[p]));
      }
    }
    c->vals[k] = (c->wic ? w : 1.0) * (1.0 - acc);
  }
}

void orc_patch_ncc_f64(const double* fixed, const double* mov, int32_t rows, int32_t cols,
                       int64_t num_mov, int32_t patch_radius, int32_t patch_stride,
                       double* sim_out, int32_t threads)
{
  orc_patch_ncc_masked_f64(fixed, mov, NULL, rows, cols, num_mov, patch_radius, patch_stride, sim_out, threads);
}

void orc_patch_ncc_masked_f64(const double* fixed, const double* mov, const uint8_t* mask, int32_t rows,
                              int32_t cols, int64_t num_mov, int32_t patch_radius, int32_t patch_stride,
                              double* sim_out, int32_t threads)
{
  double* wgt = NULL;
  if (mask) {
    const int32_t diam = 2 * patch_radius + 1;
    const int32_t ppr = (cols - 1 - 2 * patch_radius) / patch_stride + 1;
    const int32_t ppc = (rows - 1 - 2 * patch_radius) / patch_stride + 1;
    const int64_t np = (int64_t)ppr * ppc;
    wgt = (double*)malloc(sizeof(double) * (size_t)np);
    double wsum = 0.0;
    for (int64_t p = 0; p < np; ++p) {
      const int32_t r0 = (int32_t)(p / ppr) * patch_stride, c0 = (int32_t)(p % ppr) * patch_stride;
      int64_t cnt = 0;
      for (int32_t y = 0; y < diam; ++y)
        for (int32_t x = 0; x < diam; ++x) cnt += mask[(int64_t)(r0 + y) * cols + c0 + x] ? 1 : 0;
      wgt[p] = (double)cnt / ((double)diam * diam);
      wsum += wgt[p];
    }
    for (int64_t p = 0; p < np; ++p) wgt[p] /= wsum;
  }
  orc_patch_ncc_weighted_f64(fixed, mov, mask, rows, cols, num_mov, patch_radius, patch_stride, wgt, NULL, 0, 1, 0,
                             0, sim_out, threads);
  free(wgt);
}

void orc_patch_ncc_weighted_f64(const double* fixed, const double* mov, const uint8_t* mask, int32_t rows,
                                int32_t cols, int64_t num_mov, int32_t patch_radius, int32_t patch_stride,
                                const double* wgt, const int64_t* inds, int64_t num_inds, int32_t weight_in_combine,
                                int32_t compute_mean, int32_t mask_stats, double* sim_out, int32_t threads)
{
  const int32_t diam = 2 * patch_radius + 1;
  const int32_t ppr = (cols - 1 - 2 * patch_radius) / patch_stride + 1;
  const int32_t ppc = (rows - 1 - 2 * patch_radius) / patch_stride + 1;
  const int64_t np = (int64_t)ppr * ppc;
  const int64_t nu = inds ? num_inds : np;
  const int64_t npix = (int64_t)rows * cols;
  patch_f64_ctx c;
  c.fixed = fixed;
  c.cols = cols;
  c.diam = diam;
  c.ppr = ppr;
  c.stride = patch_stride;
  c.fmean = (double*)malloc(sizeof(double) * (size_t)np);
  c.fsd = (double*)malloc(sizeof(double) * (size_t)np);
  c.fn = (double*)malloc(sizeof(double) * (size_t)np);
  c.vals = (double*)malloc(sizeof(double) * (size_t)(nu > 0 ? nu : 1));
  c.mask = mask;
  c.wgt = wgt;
  c.inds = inds;
  c.wic = weight_in_combine;
  c.mstats = mask_stats && mask;
  double norm = 1.0;
  if (compute_mean) {
    norm = (double)nu;
  } else if (weight_in_combine) {
    norm = 0.0;
    for (int64_t k = 0; k < nu; ++k) norm += wgt ? wgt[inds ? inds[k] : k] : 1.0;
  }
  orc_parallel_for(patch_fixed_range_f64, &c, np, 256, threads);
  for (int64_t m = 0; m < num_mov; ++m) {
    c.mov = mov + m * npix;
    orc_parallel_for(patch_mov_range_f64, &c, nu, 256, threads);
    sim_out[m] = sum_f64(c.vals, nu) / norm;
  }
  free(c.fmean);
  free(c.fsd);
  free(c.fn);
  free(c.vals);
}

void orc_patch_grad_ncc_f32(const float* fixed, const float* mov, int32_t rows, int32_t cols,
                            int64_t num_mov, int32_t smooth_ksize, int32_t patch_radius,
                            int32_t patch_stride, float* sim_out, int32_t threads)
{
  orc_patch_grad_ncc_masked_f32(fixed, mov, NULL, rows, cols, num_mov, smooth_ksize, patch_radius, patch_stride,
                                sim_out, threads);
}

void orc_patch_grad_ncc_masked_f32(const float* fixed, const float* mov, const uint8_t* mask, int32_t rows,
                                   int32_t cols, int64_t num_mov, int32_t smooth_ksize, int32_t patch_radius,
                                   int32_t patch_stride, float* sim_out, int32_t threads)
{
  float* wgt = NULL;
  if (mask) {
    const int32_t diam = 2 * patch_radius + 1;
    const int32_t ppr = (cols - 1 - 2 * patch_radius) / patch_stride + 1;
    const int32_t ppc = (rows - 1 - 2 * patch_radius) / patch_stride + 1;
    const int64_t np = (int64_t)ppr * ppc;
    wgt = (float*)malloc(sizeof(float) * (size_t)np);
    patch_mask_weights_f32(mask, cols, ppr, np, patch_stride, diam, wgt);
  }
  orc_patch_grad_ncc_weighted_f32(fixed, mov, mask, rows, cols, num_mov, smooth_ksize, patch_radius, patch_stride,
                                  wgt, NULL, 0, 1, 0, 0, sim_out, threads);
  free(wgt);
}

void orc_patch_grad_ncc_weighted_f32(const float* fixed, const float* mov, const uint8_t* mask, int32_t rows,
                                     int32_t cols, int64_t num_mov, int32_t smooth_ksize, int32_t patch_radius,
                                     int32_t patch_stride, const float* wgt, const int64_t* inds, int64_t num_inds,
                                     int32_t weight_in_combine, int32_t compute_mean, int32_t mask_stats,
                                     float* sim_out, int32_t threads)
{
  /* ImgSimMetric2DPatchGradNCCCPU::compute (xregImgSimMetric2DPatchGradNCCCPU.cpp:104-224):
   * both directions use the parent's patches and weights (enforce_same_patches_in_both_x_and_y_,
   * :126-135) */
  const int64_t npix = (int64_t)rows * cols;
  float* fgx = (float*)malloc(sizeof(float) * (size_t)npix);
  float* fgy = (float*)malloc(sizeof(float) * (size_t)npix);
  float* mgx = (float*)malloc(sizeof(float) * (size_t)npix);
  float* mgy = (float*)malloc(sizeof(float) * (size_t)npix);
  orc_sobel_f32(fixed, rows, cols, smooth_ksize, fgx, fgy);
  for (int64_t m = 0; m < num_mov; ++m) {
    float sx, sy;
    orc_sobel_f32(mov + m * npix, rows, cols, smooth_ksize, mgx, mgy);
    orc_patch_ncc_weighted_f32(fgx, mgx, mask, rows, cols, 1, patch_radius, patch_stride, wgt, inds, num_inds,
                               weight_in_combine, compute_mean, mask_stats, &sx, threads);
    orc_patch_ncc_weighted_f32(fgy, mgy, mask, rows, cols, 1, patch_radius, patch_stride, wgt, inds, num_inds,
                               weight_in_combine, compute_mean, mask_stats, &sy, threads);
    sim_out[m] = (float)(0.5 * (double)(sx + sy));
  }
  free(fgx);
  free(fgy);
  free(mgx);
  free(mgy);
}

void orc_patch_grad_ncc_f64(const double* fixed, const double* mov, int32_t rows, int32_t cols,
                            int64_t num_mov, int32_t smooth_ksize, int32_t patch_radius,
                            int32_t patch_stride, double* sim_out, int32_t threads)
{
  orc_patch_grad_ncc_masked_f64(fixed, mov, NULL, rows, cols, num_mov, smooth_ksize, patch_radius, patch_stride,
                                sim_out, threads);
}

void orc_patch_grad_ncc_masked_f64(const double* fixed, const double* mov, const uint8_t* mask, int32_t rows,
                                   int32_t cols, int64_t num_mov, int32_t smooth_ksize, int32_t patch_radius,
                                   int32_t patch_stride, double* sim_out, int32_t threads)
{
  const int64_t npix = (int64_t)rows * cols;
  double* fgx = (double*)malloc(sizeof(double) * (size_t)npix);
  double* fgy = (double*)malloc(sizeof(double) * (size_t)npix);
  double* mgx = (double*)malloc(sizeof(double) * (size_t)npix);
  double* mgy = (double*)malloc(sizeof(double) * (size_t)npix);
  orc_sobel_f64(fixed, rows, cols, smooth_ksize, fgx, fgy);
  for (int64_t m = 0; m < num_mov; ++m) {
    double sx, sy;
    orc_sobel_f64(mov + m * npix, rows, cols, smooth_ksize, mgx, mgy);
    orc_patch_ncc_masked_f64(fgx, mgx, mask, rows, cols, 1, patch_radius, patch_stride, &sx, threads);
    orc_patch_ncc_masked_f64(fgy, mgy, mask, rows, cols, 1, patch_radius, patch_stride, &sy, threads);
    sim_out[m] = 0.5 * (sx + sy);
  }
  free(fgx);
  free(fgy);
  free(mgx);
  free(mgy);
}

void orc_patch_grad_ncc_weighted_f64(const double* fixed, const double* mov, const uint8_t* mask, int32_t rows,
                                     int32_t cols, int64_t num_mov, int32_t smooth_ksize, int32_t patch_radius,
                                     int32_t patch_stride, const double* wgt, const int64_t* inds, int64_t num_inds,
                                     int32_t weight_in_combine, int32_t compute_mean, int32_t mask_stats,
                                     double* sim_out, int32_t threads)
{
  const int64_t npix = (int64_t)rows * cols;
  double* fgx = (double*)malloc(sizeof(double) * (size_t)npix);
  double* fgy = (double*)malloc(sizeof(double) * (size_t)npix);
  double* mgx = (double*)malloc(sizeof(double) * (size_t)npix);
  double* mgy = (double*)malloc(sizeof(double) * (size_t)npix);
  orc_sobel_f64(fixed, rows, cols, smooth_ksize, fgx, fgy);
  for (int64_t m = 0; m < num_mov; ++m) {
    double sx, sy;
    orc_sobel_f64(mov + m * npix, rows, cols, smooth_ksize, mgx, mgy);
    orc_patch_ncc_weighted_f64(fgx, mgx, mask, rows, cols, 1, patch_radius, patch_stride, wgt, inds, num_inds,
                               weight_in_combine, compute_mean, mask_stats, &sx, threads);
    orc_patch_ncc_weighted_f64(fgy, mgy, mask, rows, cols, 1, patch_radius, patch_stride, wgt, inds, num_inds,
                               weight_in_combine, compute_mean, mask_stats, &sy, threads);
    sim_out[m] = 0.5 * (sx + sy);
  }
  free(fgx);
  free(fgy);
  free(mgx);
  free(mgy);
}

/* ---------------------------------------------------------------------------------------- */
/* HU -> linear attenuation (lib/image/xregHUToLinAtt.cpp:45-68) */

void orc_hu_to_lin_att(const float* hu, int64_t n, double hu_lower, double mu_water, double mu_air, float* out)
{
  const double hu_scale = (mu_water - mu_air) * 1.0e-3;
  const double mu_lower = (hu_lower * hu_scale) + mu_water;
  for (int64_t i = 0; i < n; ++i) {
    const double a = ((double)hu[i] * hu_scale) + mu_water - mu_lower;
    /* std::max(a, 0.0) returns a unless a < 0 (a NaN passes through) */
    out[i] = (float)((a < 0.0) ? 0.0 : a);
  }
}

/* ---------------------------------------------------------------------------------------- */
/* multi-resolution fixed images: ITK 5.1 DiscreteGaussianImageFilter + BSpline resampling as
 * used by DownsampleImage (lib/itk/xregITKResampleUtils.h:51-120, 181-186). ITK is not part
 * of the reference tree; its published algorithms are restated here. */

/* Abramowitz & Stegun 9.8.1-9.8.4 (the approximations itk::GaussianOperator uses) */
static double orc_bessel_i0(double x)
{
  double ax = fabs(x), t;
  if (ax < 3.75) {
    t = (x / 3.75) * (x / 3.75);
    return 1.0 + t * (3.5156229 + t * (3.0899424 + t * (1.2067492 + t * (0.2659732 + t * (0.0360768 + t * 0.0045813)))));
  }
  t = 3.75 / ax;
  return (exp(ax) / sqrt(ax)) *
         (0.39894228 + t * (0.01328592 + t * (0.00225319 + t * (-0.00157565 + t * (0.00916281 + t * (-0.02057706 +
          t * (0.02635537 + t * (-0.01647633 + t * 0.00392377))))))));
}

static double orc_bessel_i1(double x)
{
  double ax = fabs(x), t, r;
  if (ax < 3.75) {
    t = (x / 3.75) * (x / 3.75);
    r = ax * (0.5 + t * (0.87890594 + t * (0.51498869 + t * (0.15084934 + t * (0.02658733 + t * (0.00301532 +
         t * 0.00032411))))));
  } else {
    t = 3.75 / ax;
    r = (exp(ax) / sqrt(ax)) *
        (0.39894228 + t * (-0.03988024 + t * (-0.00362018 + t * (0.00163801 + t * (-0.01031555 + t * (0.02282967 +
         t * (-0.02895312 + t * (0.01787654 - t * 0.00420059))))))));
  }
  return (x < 0.0) ? -r : r;
}

int32_t orc_gaussian_kernel(double variance, double max_error, int32_t max_width, double* k, int32_t cap)
{
  double half[1024];
  int32_t nh = 0, i;
  double sum = 0.0;
  const double et = exp(-variance);
  half[nh++] = et * orc_bessel_i0(variance);
  sum += half[0];
  half[nh++] = et * orc_bessel_i1(variance);
  sum += 2.0 * half[1];
  for (i = 2; sum < 1.0 - max_error && i < 1023; ++i) {
    half[nh++] = half[i - 2] - 2 * (i - 1) / variance * half[i - 1];
    sum += 2.0 * half[i];
    if (half[i] <= 0.0) break;
    if (nh > max_width) break;
  }
  for (i = 0; i < nh; ++i) half[i] /= sum;
  const int32_t r = nh - 1;
  if (2 * r + 1 > cap) return -1;
  for (i = 0; i <= r; ++i) k[r + i] = k[r - i] = half[i];
  return 2 * r + 1;
}

static void orc_bspline_line(double* c, int64_t n, int64_t stride)
{
  const double z = sqrt(3.0) - 2.0;
  int64_t i;
  if (n == 1) return;
  for (i = 0; i < n; ++i) c[i * stride] *= (1.0 - z) * (1.0 - 1.0 / z);
  {
    const int64_t horizon = (int64_t)ceil(log(1.0e-10) / log(fabs(z)));
    double zn = z;
    if (horizon < n) {
      double sum = c[0];
      for (i = 1; i < horizon; ++i) {
        sum += zn * c[i * stride];
        zn *= z;
      }
      c[0] = sum;
    } else {
      const double iz = 1.0 / z;
      double z2n = pow(z, (double)(n - 1));
      double sum = c[0] + z2n * c[(n - 1) * stride];
      z2n *= z2n * iz;
      for (i = 1; i <= n - 2; ++i) {
        sum += (zn + z2n) * c[i * stride];
        zn *= z;
        z2n *= iz;
      }
      c[0] = sum / (1.0 - zn * zn);
    }
  }
  for (i = 1; i < n; ++i) c[i * stride] += z * c[(i - 1) * stride];
  c[(n - 1) * stride] = (z / (z * z - 1.0)) * (z * c[(n - 2) * stride] + c[(n - 1) * stride]);
  for (i = n - 2; i >= 0; --i) c[i * stride] = z * (c[(i + 1) * stride] - c[i * stride]);
}

static int64_t orc_mirror(int64_t i, int64_t n)
{
  int64_t n2;
  if (n == 1) return 0;
  n2 = 2 * n - 2;
  if (i < 0) i = -i;
  i %= n2;
  return (i >= n) ? n2 - i : i;
}

void orc_downsample_image(const float* img, int32_t rows, int32_t cols, double row_sp, double col_sp,
                          double factor, double sigma, float* out, int32_t out_rows, int32_t out_cols)
{
  const int64_t n = (int64_t)rows * cols;
  float* sm = (float*)malloc((size_t)n * sizeof(float));
  double* tmp = (double*)malloc((size_t)n * sizeof(double));
  int64_t r, c, i, j;
  if (factor < 1.0 && fabs(sigma) > 1.0e-6) {
    double k[129];
    const double s = (sigma < 0.0) ? 0.5 / factor : sigma;
    const int32_t nk = orc_gaussian_kernel(s * s, 0.01, 32, k, 129);
    const int32_t R = nk / 2;
    for (r = 0; r < rows; ++r)
      for (c = 0; c < cols; ++c) {
        double acc = 0.0;
        int32_t q;
        for (q = -R; q <= R; ++q) {
          int64_t cc = c + q;
          cc = cc < 0 ? 0 : (cc >= cols ? cols - 1 : cc);
          acc += k[q + R] * (double)img[r * cols + cc];
        }
        tmp[r * cols + c] = acc;
      }
    for (r = 0; r < rows; ++r)
      for (c = 0; c < cols; ++c) {
        double acc = 0.0;
        int32_t q;
        for (q = -R; q <= R; ++q) {
          int64_t rr = r + q;
          rr = rr < 0 ? 0 : (rr >= rows ? rows - 1 : rr);
          acc += k[q + R] * tmp[rr * cols + c];
        }
        sm[r * cols + c] = (float)acc;
      }
  } else {
    memcpy(sm, img, (size_t)n * sizeof(float));
  }
  for (i = 0; i < n; ++i) tmp[i] = (double)sm[i];
  for (r = 0; r < rows; ++r) orc_bspline_line(tmp + r * cols, cols, 1);
  for (c = 0; c < cols; ++c) orc_bspline_line(tmp + c, rows, cols);
  for (i = 0; i < out_rows; ++i)
    for (j = 0; j < out_cols; ++j) {
      const double x = ((col_sp / factor) * (double)j) * (1.0 / col_sp);
      const double y = ((row_sp / factor) * (double)i) * (1.0 / row_sp);
      double v = 0.0;
      if (x >= -0.5 && x < cols - 0.5 && y >= -0.5 && y < rows - 0.5) {
        double wx[4], wy[4], t;
        int64_t bx = (int64_t)floor(x) - 1, by = (int64_t)floor(y) - 1, p;
        t = x - (double)(bx + 1);
        wx[3] = (1.0 / 6.0) * t * t * t;
        wx[0] = 1.0 / 6.0 + 0.5 * t * (t - 1.0) - wx[3];
        wx[2] = t + wx[0] - 2.0 * wx[3];
        wx[1] = 1.0 - wx[0] - wx[2] - wx[3];
        t = y - (double)(by + 1);
        wy[3] = (1.0 / 6.0) * t * t * t;
        wy[0] = 1.0 / 6.0 + 0.5 * t * (t - 1.0) - wy[3];
        wy[2] = t + wy[0] - 2.0 * wy[3];
        wy[1] = 1.0 - wy[0] - wy[2] - wy[3];
        for (p = 0; p < 16; ++p) {
          const int64_t a = p % 4, b = p / 4;
          v += (wx[a] * wy[b]) * tmp[orc_mirror(by + b, rows) * cols + orc_mirror(bx + a, cols)];
        }
      }
      out[i * out_cols + j] = (float)v;
    }
  free(sm);
  free(tmp);
}
