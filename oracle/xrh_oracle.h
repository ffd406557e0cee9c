/*
 * xrh_oracle.h -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * CPU restatement of xReg's intensity-based 2D/3D registration hot path:
 *   - the line-integral DRR of RayCasterLineIntCPU
 *       (lib/ray_cast/xregRayCastLineIntCPU.cpp:105-349,
 *        lib/ray_cast/xregRayCastBaseCPU.cpp:128-158),
 *   - the NCC / Grad-NCC / Patch-NCC / Patch-Grad-NCC CPU similarity metrics
 *       (lib/regi/sim_metrics_2d/xregImgSimMetric2D{NCC,GradImg,GradNCC,PatchNCC,
 *        PatchGradNCC}CPU.cpp) and the mean-over-views combiner
 *       (xregImgSimMetric2DCombine.cpp:79-98).
 *
 * Two flavours of every function:
 *   *_f32  follows the reference's FP32 arithmetic and operation order (Eigen 3.3
 *          SSE reductions, ITK 5.1 trilinear in double rounded to float per sample,
 *          OpenCV 3.4 separable Sobel / fixed-table Gaussian with BORDER_REFLECT_101),
 *   *_f64  the same algorithm carried out entirely in double (the FP64 oracle).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library. The product path (libxrh.so) never links or calls it.
 */
#ifndef XRH_ORACLE_H
#define XRH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Camera as xReg's CameraModel holds it (lib/transforms/xregPerspectiveXform.h:108-273).
 * frame: 0 = origin at focal point, detector on +z; 1 = detector on -z (xReg default);
 *        2 = origin on the detector. Matrices are row-major. */
typedef struct orc_cam_f {
  int32_t rows;
  int32_t cols;
  int32_t frame;
  float focal;
  float kinv[9];
  float einv[16];
  float pinhole[3];
} orc_cam_f;

typedef struct orc_cam_d {
  int32_t rows;
  int32_t cols;
  int32_t frame;
  int32_t pad_;
  double focal;
  double kinv[9];
  double einv[16];
  double pinhole[3];
} orc_cam_d;

/* ---- geometry helpers (exported so tests can pin them one by one) ---- */
void orc_mat3_inverse_f32(const float m[9], float out[9]);
void orc_affine_inverse_f32(const float m[16], float out[16]);
void orc_affine_inverse_f64(const double m[16], double out[16]);
void orc_affine_mul_f32(const float a[16], const float b[16], float out[16]);
void orc_affine_pt_f32(const float m[16], const float p[3], float out[3]);
void orc_cam_naive_f32(float focal, int32_t rows, int32_t cols, float row_sp, float col_sp,
                       orc_cam_f* out);
void orc_cam_ind_to_det_f32(const orc_cam_f* cam, float col, float row, float out[3]);
void orc_cam_ind_to_det_f64(const orc_cam_d* cam, double col, double row, double out[3]);
int32_t orc_ray_rect_f32(const float mn[3], const float mx[3], const float p0[3],
                         const float d[3], int32_t limit_to_segment, float* t0, float* t1);
int32_t orc_ray_rect_f64(const double mn[3], const double mx[3], const double p0[3],
                         const double d[3], int32_t limit_to_segment, double* t0, double* t1);
double orc_itk_linear_interp(const float* vol, const int64_t dims[3], const float idx[3]);
/* nearest neighbour, Lanczos-4 windowed sinc and cubic B-spline as the reference's ITK
 * interpolators evaluate them (B-spline from orc_bspline_coefficients' double coefficients) */
double orc_itk_nn_interp(const float* vol, const int64_t dims[3], const float idx[3]);
double orc_itk_sinc_interp(const float* vol, const int64_t dims[3], const float idx[3]);
void orc_bspline_coefficients(const float* vol, const int64_t dims[3], double* coef);
double orc_itk_bspline_interp(const double* coef, const int64_t dims[3], const float idx[3]);
double orc_linear_interp_f64(const double* vol, const int64_t dims[3], const double idx[3]);

/* ---- DRR (RayCasterLineIntCPU::compute for one volume) ----
 * vol: x-fastest (i + X*(j + Y*k)); p2i: ITK physical -> continuous index (row-major);
 * xforms: num_projs row-major 4x4 cam->phys poses; cam_for_proj: camera of each proj.
 * store: 0 REPLACE (pre-fill with bg or default_bg), 1 ACCUM (keep out's contents).
 * kernel: 0 SUM, 1 MAX. bg: num_cams images (rows*cols each) or NULL.
 * samples_out (optional): total trilinear samples taken (sum over rays of n+1).
 * threads: 0 = all hardware threads. Results do not depend on the thread count. */
void orc_drr_f32(const float* vol, const int64_t dims[3], const float p2i[16],
                 const orc_cam_f* cams, int32_t num_cams,
                 const float* xforms, const int64_t* cam_for_proj, int64_t num_projs,
                 float step, int32_t kernel, int32_t store, const float* bg, float default_bg,
                 float* out, int32_t threads, int64_t* samples_out);
/* the same with RayCaster::InterpMethod interp (0 linear, 1 NN, 2 sinc, 3 B-spline) and
 * RayCasterCPU's anti-aliasing: aa_rays jittered rays per pixel (0: off), jitter from aa_seed */
void orc_drr_interp_f32(const float* vol, const int64_t dims[3], const float p2i[16],
                        const orc_cam_f* cams, int32_t num_cams,
                        const float* xforms, const int64_t* cam_for_proj, int64_t num_projs,
                        float step, int32_t kernel, int32_t store, const float* bg, float default_bg,
                        int32_t interp, int32_t aa_rays, uint64_t aa_seed, float* out, int32_t threads,
                        int64_t* samples_out);
/* the jitter of ray k of range index `pixel`, component 0 (column) / 1 (row): [-0.5, 0.5) */
float orc_aa_jitter(uint64_t seed, uint64_t pixel, int32_t k, int32_t comp);

void orc_drr_f64(const double* vol, const int64_t dims[3], const double p2i[16],
                 const orc_cam_d* cams, int32_t num_cams,
                 const double* xforms, const int64_t* cam_for_proj, int64_t num_projs,
                 double step, int32_t kernel, int32_t store, const double* bg, double default_bg,
                 double* out, int32_t threads, int64_t* samples_out);

/* Per-ray last sample index n (-1: miss) of orc_drr_f32 / orc_drr_f64, set-up only. */
void orc_ray_steps_f32(const int64_t dims[3], const float p2i[16], const orc_cam_f* cams,
                       const float* xforms, const int64_t* cam_for_proj, int64_t num_projs,
                       float step, int32_t* steps_out, int32_t threads);
void orc_ray_steps_f64(const int64_t dims[3], const double p2i[16], const orc_cam_d* cams,
                       const double* xforms, const int64_t* cam_for_proj, int64_t num_projs,
                       double step, int32_t* steps_out, int32_t threads);

/* ---- similarity metrics (values are xReg's: smaller is better) ---- */
/* NCC: returns (1 - ncc)/2 per moving image. mov is NOT modified (the reference zero-means
 * it in place; callers that need that side effect use orc_zero_mean_inplace_f32). */
void orc_ncc_f32(const float* fixed, const float* mov, int64_t npix, int64_t num_mov,
                 const uint8_t* mask, float* sim_out);
void orc_ncc_f64(const double* fixed, const double* mov, int64_t npix, int64_t num_mov,
                 const uint8_t* mask, double* sim_out);
void orc_zero_mean_inplace_f32(float* img, int64_t npix);

/* Gradient images (OpenCV Sobel ksize 3, optional GaussianBlur ksize x ksize, sigma 0,
 * BORDER_REFLECT_101). smooth_ksize 0 disables smoothing. */
void orc_sobel_f32(const float* img, int32_t rows, int32_t cols, int32_t smooth_ksize,
                   float* gx, float* gy);
void orc_sobel_f64(const double* img, int32_t rows, int32_t cols, int32_t smooth_ksize,
                   double* gx, double* gy);

void orc_grad_ncc_f32(const float* fixed, const float* mov, int32_t rows, int32_t cols,
                      int64_t num_mov, int32_t smooth_ksize, float* sim_out);
void orc_grad_ncc_f64(const double* fixed, const double* mov, int32_t rows, int32_t cols,
                      int64_t num_mov, int32_t smooth_ksize, double* sim_out);

/* Patch NCC with the reference defaults: stride patch_stride, unit weights, no mask,
 * weighted-sum-then-normalise combine (xregImgSimMetric2DPatchNCCCPU.cpp:190-287). */
void orc_patch_ncc_f32(const float* fixed, const float* mov, int32_t rows, int32_t cols,
                       int64_t num_mov, int32_t patch_radius, int32_t patch_stride,
                       float* sim_out, int32_t threads);
void orc_patch_ncc_f64(const double* fixed, const double* mov, int32_t rows, int32_t cols,
                       int64_t num_mov, int32_t patch_radius, int32_t patch_stride,
                       double* sim_out, int32_t threads);

/* The same with a fixed-image mask (non-zero = use) and the reference's default mask options
 * (ImgSimMetric2DPatchCommon.h:150-157): patch weight = fraction of unmasked pixels, weights
 * normalised to sum to one, patches of weight <= 1e-6 skipped, patch statistics over the
 * whole patch, the correlation summed over unmasked pixels only, weighted combine. */
void orc_patch_ncc_masked_f32(const float* fixed, const float* mov, const uint8_t* mask, int32_t rows,
                              int32_t cols, int64_t num_mov, int32_t patch_radius, int32_t patch_stride,
                              float* sim_out, int32_t threads);
void orc_patch_ncc_masked_f64(const double* fixed, const double* mov, const uint8_t* mask, int32_t rows,
                              int32_t cols, int64_t num_mov, int32_t patch_radius, int32_t patch_stride,
                              double* sim_out, int32_t threads);
void orc_patch_grad_ncc_masked_f32(const float* fixed, const float* mov, const uint8_t* mask, int32_t rows,
                                   int32_t cols, int64_t num_mov, int32_t smooth_ksize, int32_t patch_radius,
                                   int32_t patch_stride, float* sim_out, int32_t threads);
void orc_patch_grad_ncc_masked_f64(const double* fixed, const double* mov, const uint8_t* mask, int32_t rows,
                                   int32_t cols, int64_t num_mov, int32_t smooth_ksize, int32_t patch_radius,
                                   int32_t patch_stride, double* sim_out, int32_t threads);
void orc_patch_grad_ncc_f32(const float* fixed, const float* mov, int32_t rows, int32_t cols,
                            int64_t num_mov, int32_t smooth_ksize, int32_t patch_radius,
                            int32_t patch_stride, float* sim_out, int32_t threads);
void orc_patch_grad_ncc_f64(const double* fixed, const double* mov, int32_t rows, int32_t cols,
                            int64_t num_mov, int32_t smooth_ksize, int32_t patch_radius,
                            int32_t patch_stride, double* sim_out, int32_t threads);

/* ImgSimMetric2DPatchCommon::compute_weights (xregImgSimMetric2DPatchCommon.cpp:306-400):
 * per-patch weights from a weight image (its value at each patch centre; NULL for none) and
 * a mask (NULL for none): with a weight image, the centre values, zeroed where the mask is off
 * at the centre when use_mask_for_weighting; with only a mask used for weighting, the
 * unmasked fraction of each patch; with neither, 1. normalize: divided by their float sum. */
void orc_patch_weights_f32(const float* wgt_img, const uint8_t* mask, int32_t rows, int32_t cols,
                           int32_t patch_radius, int32_t patch_stride, int32_t use_mask_for_weighting,
                           int32_t normalize, float* w_out);

/* The fixed-image patch variances the patch metrics can use as a weight image
 * (use_fixed_img_patch_variances_as_wgts, xregImgSimMetric2DPatchNCCCPU.cpp:331-440): per patch,
 * the variance over its pixels (its unmasked pixels with a mask; n - 1 denominator, std
 * clamped at 1e-6), written at the patch centre; zero elsewhere. */
void orc_patch_variance_image_f32(const float* img, const uint8_t* mask, int32_t rows, int32_t cols,
                                  int32_t patch_radius, int32_t patch_stride, float* out);

/* Patch NCC / Patch-Grad-NCC with the reference's general combine
 * (xregImgSimMetric2DPatchNCCCPU.cpp:186-287): per-patch weights wgt (NULL = 1), the patch
 * indices used in order (inds, num_inds; NULL = all), weight_patch_sims_in_combine_ and
 * compute_mean_of_patch_sims_. A mask restricts the correlation to unmasked pixels; with mask_stats
 * (use_mask_for_patch_stats_) the patch means and deviations, fixed and moving, are taken over
 * the unmasked pixels too and the fixed patch is scaled by their count. */
void orc_patch_ncc_weighted_f32(const float* fixed, const float* mov, const uint8_t* mask, int32_t rows,
                                int32_t cols, int64_t num_mov, int32_t patch_radius, int32_t patch_stride,
                                const float* wgt, const int64_t* inds, int64_t num_inds, int32_t weight_in_combine,
                                int32_t compute_mean, int32_t mask_stats, float* sim_out, int32_t threads);
void orc_patch_ncc_weighted_f64(const double* fixed, const double* mov, const uint8_t* mask, int32_t rows,
                                int32_t cols, int64_t num_mov, int32_t patch_radius, int32_t patch_stride,
                                const double* wgt, const int64_t* inds, int64_t num_inds, int32_t weight_in_combine,
                                int32_t compute_mean, int32_t mask_stats, double* sim_out, int32_t threads);
void orc_patch_grad_ncc_weighted_f32(const float* fixed, const float* mov, const uint8_t* mask, int32_t rows,
                                     int32_t cols, int64_t num_mov, int32_t smooth_ksize, int32_t patch_radius,
                                     int32_t patch_stride, const float* wgt, const int64_t* inds, int64_t num_inds,
                                     int32_t weight_in_combine, int32_t compute_mean, int32_t mask_stats,
                                     float* sim_out, int32_t threads);
void orc_patch_grad_ncc_weighted_f64(const double* fixed, const double* mov, const uint8_t* mask, int32_t rows,
                                     int32_t cols, int64_t num_mov, int32_t smooth_ksize, int32_t patch_radius,
                                     int32_t patch_stride, const double* wgt, const int64_t* inds, int64_t num_inds,
                                     int32_t weight_in_combine, int32_t compute_mean, int32_t mask_stats,
                                     double* sim_out, int32_t threads);

/* Every grid patch's unweighted score per moving image (save_all_per_patch_scores: Patch-NCC
 * 1 - accum, metric 3 Patch-Grad-NCC 0.5 (x + y) in float), num_mov x num_patches. */
void orc_patch_scores_f32(int32_t metric, const float* fixed, const float* mov, const uint8_t* mask, int32_t rows,
                          int32_t cols, int64_t num_mov, int32_t smooth_ksize, int32_t patch_radius,
                          int32_t patch_stride, int32_t mask_stats, float* scores_out, int32_t threads);

/* Eigen 3.3 SSE sum restatement, exported for its own test. */
float orc_eigen_sum_f32(const float* x, int64_t n);

/* HUToLinAttFilter::GenerateData (lib/image/xregHUToLinAtt.cpp:45-68): per voxel in double,
 * std::max((hu * s + mu_water) - mu_lower, 0.0) rounded to float, s = (mu_water - mu_air) * 1e-3,
 * mu_lower = hu_lower * s + mu_water. */
/* itk::GaussianOperator::GenerateCoefficients (ITK 5.1): symmetric discrete-Gaussian kernel
 * (e^-t I_n(t), Bessel recurrence, truncated at 1 - max_error of the mass, normalised) for
 * variance t in pixels. Writes up to cap taps into k; returns the number of taps (2r+1). */
int32_t orc_gaussian_kernel(double variance, double max_error, int32_t max_width, double* k, int32_t cap);
/* DownsampleImage with the default B-spline interpolator (lib/itk/xregITKResampleUtils.h:
 * 51-120): DiscreteGaussian (x then y, edge replication, double between passes, float result),
 * cubic B-spline decomposition (x then y, mirror, tolerance 1e-10) and resampling at
 * (spacing / factor) * index / spacing. out is out_rows x out_cols = floor(n * factor + 0.5). */
void orc_downsample_image(const float* img, int32_t rows, int32_t cols, double row_sp, double col_sp,
                          double factor, double sigma, float* out, int32_t out_rows, int32_t out_cols);

void orc_hu_to_lin_att(const float* hu, int64_t n, double hu_lower, double mu_water, double mu_air, float* out);

#ifdef __cplusplus
}
#endif

#endif
