/*
 * xrh.h -- C-ABI of the MI355X (gfx950) DRR + image-similarity hot path.
 *
 * This is the device layer that replaces xReg's lib/opencl backend (Boost.Compute queues,
 * OpenCL-C kernel strings, ViennaCL GEMV reductions) underneath the RayCaster and
 * ImgSimMetric2D interfaces. Conventions:
 *   - extern "C", plain pointers and sizes; no C++ or torch types cross this boundary;
 *   - every entry point returns an xrh_status; on failure xrh_last_error() (thread-local)
 *     holds a message; no exception ever escapes;
 *   - host arrays are caller-owned and only read during the call; device buffers passed
 *     in ("_dev") are caller-owned device memory (xrh_device_alloc or any HIP allocation);
 *     a range that starts inside an xrh_device_alloc allocation and runs past its end is
 *     refused with XRH_ERR_INVALID before any work is queued (other pointers are trusted);
 *   - all device work is ordered on the context's stream; entry points that return host
 *     results synchronise that stream before returning (the reference's compute() calls
 *     are synchronous: xregRayCastLineIntOCL.cpp:192-197).
 * Matrices are row-major 4x4 (affine, last row 0 0 0 1) or row-major 3x3.
 */
#ifndef XRH_H
#define XRH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XRH_ABI_VERSION 1

typedef int32_t xrh_status;
enum {
  XRH_OK = 0,
  XRH_ERR_INVALID = 1,     /* bad argument (xregASSERT / xregThrow in the reference)   */
  XRH_ERR_HIP = 2,         /* a HIP runtime call failed                                 */
  XRH_ERR_OOM = 3,         /* device allocation failed                                  */
  XRH_ERR_BUDGET = 4,      /* request exceeds the context's HBM budget                  */
  XRH_ERR_UNSUPPORTED = 5, /* UnsupportedOperationException (xregRayCastInterface.h:59) */
  XRH_ERR_STATE = 6        /* call made in the wrong order (e.g. before allocation)     */
};

/* Storage type of a resident volume. FP16/BF16 halve the gather bytes; interpolation and
 * accumulation stay FP32. */
enum { XRH_F32 = 0, XRH_F16 = 1, XRH_BF16 = 2 };

/* RayCastLineIntKernel (lib/ray_cast/xregRayCastInterface.h:575-579). */
enum { XRH_LINE_INT_SUM = 0, XRH_LINE_INT_MAX = 1 };

/* RayCaster::ProjPixelStoreMethod (lib/ray_cast/xregRayCastInterface.h:71-75). */
enum { XRH_STORE_REPLACE = 0, XRH_STORE_ACCUM = 1 };

/* Ray set-up arithmetic.
 * XRH_GEOM_FP32: RayCasterLineIntCPU's float arithmetic in its operation order with no
 *   fused multiply-add (xregRayCastLineIntCPU.cpp:204-277), so step counts and sample
 *   positions match the CPU reference bit for bit; samples advance by repeated addition.
 * XRH_GEOM_FP64: pose composition and per-ray set-up (entry/exit, chord length, step
 *   count) in double; samples are placed at start + i*step in float. */
enum { XRH_GEOM_FP32 = 0, XRH_GEOM_FP64 = 1 };

/* Volume interpolation: RayCaster::InterpMethod (lib/ray_cast/xregRayCastInterface.h:61-67),
 * the interpolators RayCasterLineIntCPU creates per range (xregRayCastLineIntCPU.cpp:115-147). */
enum {
  XRH_INTERP_LINEAR = 0,  /* itk::LinearInterpolateImageFunction                           */
  XRH_INTERP_NN = 1,      /* itk::NearestNeighborInterpolateImageFunction                  */
  XRH_INTERP_SINC = 2,    /* itk::WindowedSincInterpolateImageFunction, radius 4, Lanczos
                             window, constant-zero boundary                               */
  XRH_INTERP_BSPLINE = 3  /* itk::BSplineInterpolateImageFunction, order 3                 */
};

/* ImgSimMetric2D subclasses on the hot path (lib/regi/sim_metrics_2d/). */
enum {
  XRH_SIM_NCC = 0,            /* ImgSimMetric2DNCCCPU            */
  XRH_SIM_GRAD_NCC = 1,       /* ImgSimMetric2DGradNCCCPU        */
  XRH_SIM_PATCH_NCC = 2,      /* ImgSimMetric2DPatchNCCCPU       */
  XRH_SIM_PATCH_GRAD_NCC = 3  /* ImgSimMetric2DPatchGradNCCCPU   */
};

typedef struct xrh_ctx xrh_ctx;
typedef struct xrh_volume xrh_volume;
typedef struct xrh_fixed xrh_fixed;
typedef struct xrh_batch xrh_batch;

/* Pinhole camera exactly as xReg's CameraModel stores it
 * (lib/transforms/xregPerspectiveXform.h:108-273): coord_frame is CameraCoordFrame
 * (0 = origin at focal point, detector on +z; 1 = detector on -z, the default;
 * 2 = origin on the detector). intrins_inv / extrins_inv / pinhole are the float members
 * CameraModel::setup computes (xregPerspectiveXform.cpp:230-254). */
typedef struct xrh_camera {
  int32_t num_det_rows;
  int32_t num_det_cols;
  int32_t coord_frame;
  float focal_len;
  float intrins_inv[9];
  float extrins_inv[16];
  float pinhole[3];
  float det_row_spacing;
  float det_col_spacing;
} xrh_camera;

/* Similarity-metric parameters. smooth_ksize is
 * ImgSimMetric2DGradImgParamInterface::smooth_img_before_sobel_kernel_radius (an odd
 * GaussianBlur aperture 1,3,5,7 or 0 = no smoothing; xregImgSimMetric2DGradImgCPU.cpp:48-55);
 * patch_radius / patch_stride are ImgSimMetric2DPatchCommon's (defaults 5 and 1,
 * xregImgSimMetric2DPatchCommon.h:141-143); patch_mask_stats is its use_mask_for_patch_stats_
 * (default 0, .h:149): with a mask, every patch's mean and standard deviation, fixed and
 * moving, are taken over its unmasked pixels only (ComputePatchMeanStdDev,
 * xregImgSimMetric2DPatchNCCCPU.cpp:558-619) and the fixed patch is scaled by their count. */
typedef struct xrh_sim_params {
  int32_t metric;
  int32_t smooth_ksize;
  int32_t patch_radius;
  int32_t patch_stride;
  int32_t patch_mask_stats;
} xrh_sim_params;

/* ---------------------------------------------------------------------------------- */
/* errors, version                                                                      */

/* Message of the last failed call on this thread (empty string if none). */
const char* xrh_last_error(void);
int32_t xrh_abi_version(void);

/* ---------------------------------------------------------------------------------- */
/* context: one HIP device, one stream, an HBM budget for everything it allocates.
 * Replaces the OpenCL context/queue pair the OCL classes take in their constructors
 * (xregRayCastBaseOCL.h:38-60) and the device selection of xregOpenCLSys.cpp:31-83. */

xrh_status xrh_ctx_create(int32_t device, uint64_t hbm_budget_bytes, xrh_ctx** out);
xrh_status xrh_ctx_destroy(xrh_ctx* ctx);
/* Waits for all work queued on the context's stream. */
xrh_status xrh_ctx_synchronize(xrh_ctx* ctx);
/* The context's hipStream_t, for callers that order their own work against it. */
void* xrh_ctx_stream(xrh_ctx* ctx);
/* Bytes currently charged to the budget. */
uint64_t xrh_ctx_bytes_in_use(const xrh_ctx* ctx);

/* Device memory charged to the context budget (DRR buffers, background images...). */
xrh_status xrh_device_alloc(xrh_ctx* ctx, uint64_t bytes, void** dptr);
xrh_status xrh_device_free(xrh_ctx* ctx, void* dptr);
/* Stream-ordered copies; _sync variants wait for completion. */
xrh_status xrh_copy_h2d(xrh_ctx* ctx, void* dst_dev, const void* src_host, uint64_t bytes);
xrh_status xrh_copy_d2h(xrh_ctx* ctx, void* dst_host, const void* src_dev, uint64_t bytes);
/* Device-to-device copy by a plain streaming kernel (16-byte loads and stores when both
 * pointers and the size are 16-byte aligned), synchronous. Replaces the OCL buffer-to-buffer
 * copies of the projection sync objects (lib/ray_cast/xregRayCastSyncBuf.cpp:27-194). Its
 * dispatch "xrh_copy_kernel" reads and writes exactly `bytes`, which makes it the reference
 * point for calibrating memory-traffic counters. */
xrh_status xrh_copy_d2d(xrh_ctx* ctx, void* dst_dev, const void* src_dev, uint64_t bytes);

/* ---------------------------------------------------------------------------------- */
/* volumes: RayCaster::set_volume / vols_changed (xregRayCastInterface.h:87-90, 425;
 * the OCL image3d upload of xregRayCastBaseOCL.cpp:446-465).
 *   vox          x-fastest float voxels (ITK buffer order), size[0]*size[1]*size[2]
 *   idx_to_phys  ITKImagePhysicalPointTransformsAsEigen (lib/itk/xregITKBasicImageUtils.h:
 *                131-168) as row-major 4x4 float
 *   storage      XRH_F32 / XRH_F16 / XRH_BF16
 * The volume is copied into a device-resident, edge-padded layout; the host array is not
 * retained. */
xrh_status xrh_volume_create(xrh_ctx* ctx, const float* vox, const int64_t size[3],
                             const float idx_to_phys[16], int32_t storage, xrh_volume** out);
xrh_status xrh_volume_destroy(xrh_volume* vol);
/* Same, with an explicit HBM layout (0 = padded linear, 1 = quad-packed FP32 corners,
 * 2 = oct-packed 16-bit corners, 3 = FP32 trilinear-polynomial coefficients per cell,
 * 4 = FP32 bilinear coefficients of the face normal to x per voxel, 5 = the same faces stored
 * in 4x4x4-cell bricks); xrh_volume_create picks 4 for FP32 and 2 otherwise
 * (XRH_VOLUME_LAYOUT=linear, =quad, =coef or =brick overrides). */
xrh_status xrh_volume_create_layout(xrh_ctx* ctx, const float* vox, const int64_t size[3],
                                    const float idx_to_phys[16], int32_t storage,
                                    int32_t layout, xrh_volume** out);
/* HUToLinAtt (lib/image/xregHUToLinAtt.cpp:45-80) fused into volume ingest: the voxels are
 * Hounsfield units and are converted on the device, per voxel in double and rounded to
 * float, to max((hu * s + mu_water) - mu_lower, 0) with s = (mu_water - mu_air) * 1e-3 and
 * mu_lower = hu_lower * s + mu_water, before the layout is built. The reference's defaults
 * are hu_lower -1000, mu_water 0.02683, mu_air 0.02485e-4 (xregHUToLinAtt.h:73-76).
 * layout -1 picks the default for the storage type. */
xrh_status xrh_volume_create_hu(xrh_ctx* ctx, const float* vox_hu, const int64_t size[3],
                                const float idx_to_phys[16], int32_t storage, int32_t layout,
                                double hu_lower, double mu_water, double mu_air, xrh_volume** out);
/* Device bytes held by the volume (its layout, empty-space map and any interpolation grids). */
uint64_t xrh_volume_bytes(const xrh_volume* vol);
/* Storage type and HBM layout a volume was built with (the automatic choice included: a
 * volume whose packed layout would pass 4 GiB is kept linear). Volumes batched together must
 * share both. */
xrh_status xrh_volume_info(const xrh_volume* vol, int32_t* storage, int32_t* layout);

/* ---------------------------------------------------------------------------------- */
/* line-integral ray casting: RayCasterLineIntCPU::compute(vol_idx)
 * (lib/ray_cast/xregRayCastLineIntCPU.cpp:294-349) including RayCasterCPU::pre_compute
 * (xregRayCastBaseCPU.cpp:128-158).
 *   cams, num_cams          camera models (host); all must share num_det_rows/cols
 *                           (xregRayCastBaseCPU.cpp:54-70)
 *   xforms_cam_to_phys      num_projs row-major 4x4 poses (host), xforms_cam_to_itk_phys_
 *   cam_for_proj            camera index of each projection (host), cam_model_for_proj_
 *   step                    ray_step_size_ (physical units)
 *   kernel                  XRH_LINE_INT_SUM / _MAX
 *   store                   XRH_STORE_REPLACE (pre-fill) / XRH_STORE_ACCUM (add into drr_dev)
 *   bg_dev                  NULL, or num_cams background images on the device, copied into
 *                           each projection before casting under either store method, so
 *                           ACCUM then adds onto the background, not onto drr_dev's previous
 *                           contents (use_bg_projs_)
 *   default_bg              default_bg_pixel_val_ used by REPLACE without bg images
 *   geom                    XRH_GEOM_FP32 / XRH_GEOM_FP64
 *   drr_dev                 [num_projs][rows][cols] float on the device
 * Returns once the work is queued on the context stream (not synchronised). */
xrh_status xrh_line_int(xrh_ctx* ctx, const xrh_volume* vol, const xrh_camera* cams,
                        int32_t num_cams, const float* xforms_cam_to_phys,
                        const int64_t* cam_for_proj, int64_t num_projs, float step,
                        int32_t kernel, int32_t store, const float* bg_dev, float default_bg,
                        int32_t geom, float* drr_dev);

/* xrh_line_int without any allocation. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, semantics and DRRs as xrh_line_int; the
 * pose, camera and tile tables go through a fixed staging area compiled into the library
 * (about 0.8 MiB of device memory per GPU, held until the process ends) instead of the
 * context's growable scratch, and the default DRR launch variant is used (tuned variants are
 * not consulted). Limits: num_projs <= XRH_NOALLOC_MAX_PROJS, num_cams <=
 * XRH_NOALLOC_MAX_CAMS; larger calls fail with XRH_ERR_INVALID before anything is queued.
 * Calls are serialised within the process and each returns after its work has finished
 * (synchronous), so overlapping calls never share the staging; its contents are those of the
 * last call. */
#define XRH_NOALLOC_MAX_PROJS 1024
#define XRH_NOALLOC_MAX_CAMS 64
xrh_status xrh_line_int_noalloc(xrh_ctx* ctx, const xrh_volume* vol, const xrh_camera* cams,
                                int32_t num_cams, const float* xforms_cam_to_phys,
                                const int64_t* cam_for_proj, int64_t num_projs, float step,
                                int32_t kernel, int32_t store, const float* bg_dev, float default_bg,
                                int32_t geom, float* drr_dev);

/* xrh_line_int with RayCasterLineIntCPU's other sampling options: the volume interpolator
 * chosen by RayCaster::set_interp_method (xregRayCastInterface.h:165-191; the interpolators of
 * xregRayCastLineIntCPU.cpp:115-147, evaluated at every sample, :270-277) and the anti-aliasing
 * factor of RayCasterCPU::set_anti_alias_factor (xregRayCastBaseCPU.h:113-130; jittered rays
 * averaged per pixel, xregRayCastLineIntCPU.cpp:150-183). Same arguments as xrh_line_int except:
 *   interp   XRH_INTERP_*
 *   aa_rays  0: one ray through each pixel's index (no anti-aliasing); k in 1..XRH_MAX_AA_RAYS:
 *            k rays through the index jittered by uniform [-0.5, 0.5) offsets in column and row,
 *            their values averaged (aa_sum += value * (1/k), float) and the average stored with
 *            the line kernel; k = 1 still jitters, as in the reference
 *   aa_seed  the jitter is a pure function of (aa_seed, projection * rows * cols + pixel, ray,
 *            component) -- counter-based SplitMix64 -- so a seed reproduces a cast exactly; the
 *            reference seeds from std::random_device and never repeats. Ignored when aa_rays = 0.
 *   (no geom) rays are always set up with XRH_GEOM_FP32, the reference's own arithmetic, so
 *            each sample lands where the reference's does (nearest-neighbour lookups depend
 *            on it)
 * XRH_INTERP_LINEAR with aa_rays = 0 is xrh_line_int with XRH_GEOM_FP32. Otherwise: nearest
 * neighbour rounds each continuous index half up as ITK 5.1 does on x86-64; windowed sinc takes
 * the 8x8x8 voxels around floor(index) with Lanczos-4 weights and zeros outside the volume; the
 * B-spline samples 4x4x4 cubic B-spline coefficients (computed on the device in double with
 * ITK's mirror-boundary recursive filter) with mirrored indices; linear (with anti-aliasing)
 * follows ITK's LinearInterpolateImageFunction. Their weights and sums are FP64, as ITK
 * evaluates them, each sample rounded to float once; the line sums are FP32 as in the reference.
 * The first such call per volume for linear, nearest neighbour or sinc (one grid, 4 B per voxel)
 * and for B-spline (another, 8 B per voxel) builds that grid: synchronous, allocated from the
 * volume's context, counted in xrh_volume_bytes and freed with the volume; the volume is
 * therefore not const here. A volume's grids are built from its stored voxels, so a 16-bit
 * volume interpolates its 16-bit values. Empty-space skipping is not used. */
#define XRH_MAX_AA_RAYS 4096
xrh_status xrh_line_int_interp(xrh_ctx* ctx, xrh_volume* vol, const xrh_camera* cams,
                               int32_t num_cams, const float* xforms_cam_to_phys,
                               const int64_t* cam_for_proj, int64_t num_projs, float step,
                               int32_t kernel, int32_t store, const float* bg_dev, float default_bg,
                               int32_t interp, int32_t aa_rays, uint64_t aa_seed, float* drr_dev);

/* Exact number of trilinear samples the reference algorithm takes for the same call (sum
 * over rays of n+1, the S_pose of the roofline byte model) and, optionally, how many of them
 * the kernel evaluates after empty-space skipping (the others lie where every voxel of their
 * trilinear support is exactly zero). Synchronous. */
xrh_status xrh_line_int_sample_count(xrh_ctx* ctx, const xrh_volume* vol,
                                     const xrh_camera* cams, int32_t num_cams,
                                     const float* xforms_cam_to_phys,
                                     const int64_t* cam_for_proj, int64_t num_projs,
                                     float step, int32_t geom, uint64_t* samples_out,
                                     uint64_t* evaluated_out);

/* ---------------------------------------------------------------------------------- */
/* fixed images + similarity: ImgSimMetric2D::set_fixed_image + allocate_resources
 * (xregImgSimMetric2D.h:69-81) process the fixed image once (zero-mean and std for NCC;
 * Sobel gradients for Grad-NCC; per-patch statistics for the patch metrics); compute()
 * (xregImgSimMetric2D.h:88) evaluates num_mov moving images stored contiguously on the
 * device. mask: optional host uint8 image (non-zero = use), or NULL. NCC and Grad-NCC use
 * the masked pixel count and sums (xregImgSimMetric2DNCCCPU.cpp:186-198); the patch metrics
 * follow ImgSimMetric2DPatchCommon's default mask options (xregImgSimMetric2DPatchCommon.h:
 * 150-157; .cpp:306-400): patch weight = unmasked fraction, normalised to sum to one,
 * patches of weight <= 1e-6 skipped, patch statistics over the whole patch, correlation over
 * unmasked pixels (xregImgSimMetric2DPatchNCCCPU.cpp:212-285). */
xrh_status xrh_fixed_create(xrh_ctx* ctx, const float* fixed_host, int32_t rows, int32_t cols,
                            const xrh_sim_params* params, const uint8_t* mask_host,
                            xrh_fixed** out);
xrh_status xrh_fixed_destroy(xrh_fixed* fixed);

/* The patch list of a patch metric, as ImgSimMetric2DPatchCommon keeps it and its SimAux
 * debug record stores it per compute call (xregImgSimMetric2DPatchCommon.h:49-61,
 * xregImgSimMetric2DPatchNCCCPU.cpp:290-298, 443-485): inclusive start / stop rows and columns
 * of each patch, rows of centres outer, and its weight (the weights last set with
 * xrh_fixed_set_patch_weights; before that 1, or the normalised unmasked fraction with a
 * mask). *num_patches receives the count; with capacity 0 nothing else is written, otherwise
 * capacity must hold it and any of the arrays may be NULL. Synchronous. */
xrh_status xrh_fixed_patches(const xrh_fixed* fixed, int64_t capacity, int32_t* start_row, int32_t* start_col,
                             int32_t* stop_row, int32_t* stop_col, float* weight, int64_t* num_patches);

/* How a patch metric combines its per-patch scores s_p into an image score, after
 * ImgSimMetric2DPatchNCCCPU::compute (xregImgSimMetric2DPatchNCCCPU.cpp:186-287) with the
 * state of ImgSimMetric2DPatchCommon (xregImgSimMetric2DPatchCommon.h:141-172):
 *   weights      num_patches floats, the PatchInfo weights (a weight image's values at the
 *                patch centres, mask fractions, normalised or not: the caller's
 *                compute_weights, xregImgSimMetric2DPatchCommon.cpp:306-400); NULL = the
 *                weights the image was created with (unit, or the mask's);
 *   weights_y    Patch-Grad-NCC only: the y direction's weights when they differ from the x
 *                direction's (its sub-metrics weighted by their own gradient image's patch
 *                variances, xregImgSimMetric2DPatchGradNCCCPU.cpp:56-79); NULL = the same.
 *                Needs non-NULL weights;
 *   patch_inds   the patches used, in order, repeats allowed (patch_inds_to_use_: a random
 *                draw, set_patches_to_use); NULL = every patch (num_inds ignored);
 *   weight_in_combine, compute_mean   weight_patch_sims_in_combine_ and
 *                compute_mean_of_patch_sims_.
 * Score = sum over the used indices of (weight_in_combine ? w_p : 1) s_p, patches with
 * |w_p| <= 1e-6 skipped when weight_in_combine, divided by num_inds if compute_mean, else by
 * the float sum of the used weights if weight_in_combine (which must not be zero), else by 1;
 * Patch-Grad-NCC averages its two directions' scores, each with its own weights and divisor.
 * Both directions use the same patches (the reference's default
 * enforce_same_patches_in_both_x_and_y_). Index checks and the divisor check fail with
 * nothing changed. Synchronises the context's stream first: work queued before the call,
 * batched and graph-replayed launches included, uses the previous combine; every later launch
 * that reads this fixed image uses the new one. */
xrh_status xrh_fixed_set_patch_weights(xrh_fixed* fixed, const float* weights, const float* weights_y,
                                       const int64_t* patch_inds, int64_t num_inds, int32_t weight_in_combine,
                                       int32_t compute_mean);

/* A patch metric's fixed image as the device holds it (rows x cols floats into host_out):
 * which = 0 the image the first direction correlates with (Patch-NCC: the fixed image;
 * Patch-Grad-NCC: its x gradient after the optional blur, ImgSimMetric2DGradImgCPU's
 * fixed_grad_img_x_, xregImgSimMetric2DGradImgCPU.cpp:54-65), which = 1 the y gradient
 * (Patch-Grad-NCC only). For weights derived from it (patch variances). Synchronous. */
xrh_status xrh_fixed_image(const xrh_fixed* fixed, int32_t which, float* host_out);

/* sim_host_out: num_mov floats (synchronises); sim_dev_out: optional device destination
 * (stream-ordered). At least one must be non-NULL. The moving images are not modified
 * (unlike the reference NCC, which zero-means them in place: xregImgSimMetric2DNCCCPU.h). */
xrh_status xrh_sim_compute(xrh_ctx* ctx, const xrh_fixed* fixed, const float* mov_dev,
                           int64_t num_mov, float* sim_dev_out, float* sim_host_out);

/* The same for a patch metric, and every grid patch's unweighted score per moving image into
 * scores_host_out (num_mov x num_patches floats, image-major): ImgSimMetric2DPatchCommon's
 * save_all_per_patch_scores record (xregImgSimMetric2DPatchCommon.h:88-95; the values of
 * xregImgSimMetric2DPatchNCCCPU.cpp:247-250, Patch-Grad-NCC 0.5 (x + y),
 * xregImgSimMetric2DPatchGradNCCCPU.cpp:226-240), for patches the combine skips too. Runs the
 * general patch kernel (a debugging aid, not the fast path); synchronous. */
xrh_status xrh_sim_patch_scores(xrh_ctx* ctx, const xrh_fixed* fixed, const float* mov_dev, int64_t num_mov,
                                float* sim_host_out, float* scores_host_out);

/* ---------------------------------------------------------------------------------- */
/* batched evaluation: one optimiser generation (all poses x all views, any mix of
 * volumes, detector sizes and fixed images) as one fused DRR + similarity pass, replayed
 * from a cached hipGraph when the shape repeats. This is what Intensity2D3DRegi::obj_fn
 * (xregIntensity2D3DRegi.cpp:571-696) does per call through ray_caster_->compute() and
 * sim_metrics_[v]->compute(), collapsed into one submission.
 *
 * A view binds a resident volume, a camera, a processed fixed image, the ray step and the
 * line-integral kernel. An item is (view index, FP64 row-major cam->phys pose). */
typedef struct xrh_view {
  const xrh_volume* vol;
  const xrh_fixed* fixed;
  xrh_camera cam;
  float step;
  int32_t kernel;
} xrh_view;

xrh_status xrh_batch_create(xrh_ctx* ctx, const xrh_view* views, int32_t num_views,
                            int64_t max_items, int32_t geom, xrh_batch** out);
xrh_status xrh_batch_destroy(xrh_batch* batch);

/* Serving form: a batch sized for up to max_views views, max_items items and detectors up to
 * max_rows x max_cols, whose view table is (re)set with xrh_batch_set_views between
 * evaluations (e.g. the registrations that are ready at a scheduling step). Replacing the
 * views drops cached graphs; device memory is allocated once, at creation. */
xrh_status xrh_batch_create_empty(xrh_ctx* ctx, int32_t max_views, int64_t max_items,
                                  int32_t max_rows, int32_t max_cols, int32_t geom,
                                  xrh_batch** out);
xrh_status xrh_batch_set_views(xrh_batch* batch, const xrh_view* views, int32_t num_views);

/* use_graph: capture / replay hipGraphs per item->view signature (default 1).
 * time_kernels: submit eagerly with HIP events around the DRR kernel and the whole
 * submission, readable through xrh_batch_last_timing (default 0). */
xrh_status xrh_batch_set_options(xrh_batch* batch, int32_t use_graph, int32_t time_kernels);

/* Evaluates n items. sim_host_out receives n similarity values (synchronises).
 * drr_dev_out: optional device buffer receiving every item's DRR back to back (in item
 * order, each rows*cols of its view); NULL skips the copy-out when the metric allows. */
xrh_status xrh_batch_eval(xrh_batch* batch, const int32_t* item_view, const double* item_pose,
                          int64_t n, float* sim_host_out, float* drr_dev_out);

/* Several objects per item, as Intensity2D3DRegi::obj_fn casts a multi-object registration
 * (xregIntensity2D3DRegi.cpp:571-629: the first volume with the REPLACE store -- onto the
 * static volumes' background images when the registration has them, :593-625 -- the others
 * with ACCUM, into the same projection, then the similarity). Entry i * num_objs + k of
 * entry_view / entry_pose (FP64 row-major cam->phys) is object k of item i; the entries of
 * one item share the detector size, and the first entry's view supplies the fixed image the
 * item is scored against. Per pixel the DRR is the first object's line integral added onto
 * the background (or 0) and the others' added in object order -- the float operations of a
 * REPLACE cast followed by ACCUM casts, in one launch. bg_dev: NULL, or background images
 * back to back in view order (rows x cols of each view, up to the last view an item's first
 * object uses); the first object of an item is cast onto its view's image.
 * n_items * num_objs <= max_items. Returns
 * n_items similarity values (synchronises); drr_dev_out as in xrh_batch_eval. */
xrh_status xrh_batch_eval_objects(xrh_batch* batch, const int32_t* entry_view, const double* entry_pose,
                                  int64_t n_items, int32_t num_objs, const float* bg_dev,
                                  float* sim_host_out, float* drr_dev_out);

/* Asynchronous form for the serving loop: queues the same work and returns; the values
 * land in sim_host_out (pinned by the batch) once xrh_batch_wait returns. */
xrh_status xrh_batch_launch(xrh_batch* batch, const int32_t* item_view,
                            const double* item_pose, int64_t n);
xrh_status xrh_batch_wait(xrh_batch* batch, float* sim_host_out);

/* Per-launch timing of the last xrh_batch_eval / launch on the context stream, measured
 * with HIP events around the DRR kernel and around the whole submission (milliseconds). */
xrh_status xrh_batch_last_timing(const xrh_batch* batch, float* drr_kernel_ms, float* total_ms);

/* DRR launch variants. Every variant casts the same rays with the same arithmetic, so DRRs
 * are identical under all of them (fused NCC sums may differ in the last FP64 bits: the
 * per-tile summation order follows the wave footprint).
 *   cell_reuse  1: one sample per loop iteration, the current cell's 32 bytes kept in
 *               registers and loaded only when a sample enters a new cell; 0: two samples per
 *               iteration, each loading its cell; -1: by launch size (1 from 16 workgroups
 *               per CU up)
 *   wave_shape  pixels per wave64 inside the 16x16 workgroup tile: 0 = 8x8, 1 = 16x4, 2 = 4x16
 *   bands       detector-row bands of the XCD-aware workgroup order, 1..16 (default 8)
 * The environment variable XRH_CELL_REUSE, when set, still overrides cell_reuse. */
typedef struct xrh_drr_variant {
  int32_t cell_reuse;
  int32_t wave_shape;
  int32_t bands;
} xrh_drr_variant;

typedef struct xrh_tune_result {
  xrh_drr_variant chosen;  /* cell_reuse resolved to 0 / 1 */
  float chosen_ms;         /* DRR kernel time of the chosen variant (minimum over reps) */
  float default_ms;        /* the same for the default variant {-1, 0, 8} */
  int32_t variants_timed;
} xrh_tune_result;

/* Autotuner of the ray-tile layout and sample loop (no counterpart in the reference, whose
 * OpenCL caster uses one fixed 1-D NDRange, xregRayCastLineIntOCL.cpp:160-197): evaluates
 * the given submission once per variant (the default, then cell_reuse {0,1} x wave_shape
 * {0,1,2} x bands {4,8,16}, then the default again), each reps + 1 times with HIP events
 * around the DRR kernel (the minimum of the last reps), and keeps the fastest if it beats the
 * default's better timing by 1%, for this batch and, in the context, for later batch and
 * xrh_line_int launches with the same storage, layout, geometry mode, line kernel, largest
 * detector and workgroups per CU (power-of-two bucket). Synchronous; runs 20 * (reps + 1)
 * submissions. */
xrh_status xrh_batch_autotune(xrh_batch* batch, const int32_t* item_view, const double* item_pose,
                              int64_t n_items, int32_t reps, xrh_tune_result* out);

/* Forces a DRR variant for this batch's later submissions; NULL returns to the context's
 * tuned table / the default. */
xrh_status xrh_batch_set_drr_variant(xrh_batch* batch, const xrh_drr_variant* variant);

/* Exact trilinear sample count of a batch (as xrh_line_int_sample_count; evaluated_out may
 * be NULL). */
xrh_status xrh_batch_sample_count(xrh_batch* batch, const int32_t* item_view,
                                  const double* item_pose, int64_t n, uint64_t* samples_out,
                                  uint64_t* evaluated_out);

/* Per-ray step counts of a batch: for every item (in item order, rows * cols int32 per item,
 * laid out like drr_dev_out of xrh_batch_eval) the last sample index n that
 * RayCasterLineIntCPU's floor(chord / step) gives (xregRayCastLineIntCPU.cpp:253-254) under
 * the batch's geometry mode, -1 where the ray misses the volume. Lets a parity check tell a
 * pixel that differs from the CPU reference by a whole sample (a step-count flip) from one
 * whose samples differ. Synchronous. */
xrh_status xrh_batch_ray_steps(xrh_batch* batch, const int32_t* item_view, const double* item_pose,
                               int64_t n, int32_t* steps_dev_out);

/* ---------------------------------------------------------------------------------- */
/* multi-resolution: DownsampleImage (lib/itk/xregITKResampleUtils.h:51-120, 181-186) as
 * DownsampleProjData applies it to fixed images (lib/image/xregProjData.cpp:40-50): ITK 5.1's
 * DiscreteGaussianImageFilter (variance sigma^2 in pixels, sigma < 0 selects 0.5 / factor,
 * MaximumError 0.01, MaximumKernelWidth 32; skipped when factor >= 1 or |sigma| <= 1e-6),
 * then ResampleImageFilter with a cubic B-spline interpolator (mirror boundary) onto the
 * grid of spacing / factor. Images are row-major [rows][cols] float on the host. */

/* Output size of DownsampleImage: floor(n * factor + 0.5) per axis. */
xrh_status xrh_downsample_size(int32_t rows, int32_t cols, double factor, int32_t* out_rows,
                               int32_t* out_cols);
/* Downsamples img (pixel spacing row_spacing x col_spacing) into out
 * (out_rows x out_cols, which must be xrh_downsample_size's). Synchronous. */
xrh_status xrh_downsample_image(xrh_ctx* ctx, const float* img, int32_t rows, int32_t cols,
                                double row_spacing, double col_spacing, double factor, double sigma,
                                float* out, int32_t out_rows, int32_t out_cols);

#ifdef __cplusplus
}
#endif

#endif /* XRH_H */
