// xreg_hip.h -- drop-in HIP subclasses of xReg's RayCaster / ImgSimMetric2D interfaces.
//
// The classes keep the method names, argument meaning, call order and error behaviour of
//   RayCaster                 lib/ray_cast/xregRayCastInterface.h:43-434
//   RayCastLineIntParamInterface  xregRayCastInterface.h:581-591
//   ImgSimMetric2D            lib/regi/sim_metrics_2d/xregImgSimMetric2D.h:42-156
//   ImgSimMetric2DGradImgParamInterface  xregImgSimMetric2DGradImgParamInterface.h:31-39
//   ImgSimMetric2DPatchCommon xregImgSimMetric2DPatchCommon.h:38-181
//   ImgSimMetric2DCombineMean xregImgSimMetric2DCombine.h:108
// so that Intensity2D3DRegi and its drivers can run unchanged. xReg's ITK / Eigen / OpenCV
// value types are replaced by small POD equivalents here (this build has none of those
// libraries); in the xReg tree the same classes would take itk::Image / Eigen arguments and
// forward their buffers to the same C-ABI calls (include/xrh.h). Device memory, kernels and
// reductions live behind that C-ABI; this layer only does bookkeeping.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/xrh.h"

namespace xreg {

using size_type = std::size_t;
using CoordScalar = float;

// xregThrow -> StringMessageException (lib/common/xregExceptionUtils.h:96)
struct StringMessageException : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// 4x4 affine, row-major, float (Eigen::Transform<float,3,Affine> in xReg, xregCommon.h:65)
struct FrameTransform {
  std::array<float, 16> m{};
  static FrameTransform Identity();
  FrameTransform operator*(const FrameTransform& o) const;  // Eigen Affine*Affine order
  float& operator()(int r, int c) { return m[4 * r + c]; }
  float operator()(int r, int c) const { return m[4 * r + c]; }
};
using FrameTransformList = std::vector<FrameTransform>;

// CameraModel (lib/transforms/xregPerspectiveXform.h:108-273), float members
struct CameraModel {
  enum CameraCoordFrame { kORIGIN_AT_FOCAL_PT_DET_POS_Z = 0, kORIGIN_AT_FOCAL_PT_DET_NEG_Z, kORIGIN_ON_DETECTOR };
  CameraCoordFrame coord_frame_type = kORIGIN_AT_FOCAL_PT_DET_NEG_Z;
  float focal_len = 0;
  size_type num_det_rows = 0, num_det_cols = 0;
  float det_row_spacing = 0, det_col_spacing = 0;
  std::array<float, 9> intrins{}, intrins_inv{};
  FrameTransform extrins = FrameTransform::Identity(), extrins_inv = FrameTransform::Identity();
  std::array<float, 3> pinhole_pt{};

  // CameraModel::setup(f, nr, nc, rs, cs) (xregPerspectiveXform.cpp:230-254)
  void setup(float focal_len, size_type nr, size_type nc, float rs, float cs);
  // CameraModel::setup(intrins, extrins, nr, nc, rs, cs) (xregPerspectiveXform.cpp:300-330)
  void setup(const std::array<float, 9>& intrins, const FrameTransform& extrins, size_type nr, size_type nc,
             float rs, float cs);
  xrh_camera to_xrh() const;
};

// itk::Image<float,3> stand-in: x-fastest buffer, spacing, origin, direction (row-major)
struct Volume {
  std::vector<float> buf;
  std::array<size_type, 3> size{};
  std::array<double, 3> spacing{1, 1, 1};
  std::array<double, 3> origin{0, 0, 0};
  std::array<double, 9> direction{1, 0, 0, 0, 1, 0, 0, 0, 1};
  // ITKImagePhysicalPointTransformsAsEigen (lib/itk/xregITKBasicImageUtils.h:131-168)
  FrameTransform idx_to_phys() const;
};
using VolPtr = std::shared_ptr<Volume>;
using VolList = std::vector<VolPtr>;

// itk::Image<float,2> stand-in for projections / fixed images: row-major rows x cols
struct Image2D {
  std::vector<float> buf;
  size_type rows = 0, cols = 0;
  float row_spacing = 1, col_spacing = 1;
};
using ImagePtr = std::shared_ptr<Image2D>;
using ImageMaskPtr = std::shared_ptr<std::vector<unsigned char>>;

// One device context per process and device, shared by all HIP classes (the OpenCL
// context/queue pair of the OCL classes).
class HIPContext {
 public:
  static HIPContext& get(int device = 0);
  xrh_ctx* ctx() const { return ctx_; }
  ~HIPContext();

 private:
  explicit HIPContext(int device);
  xrh_ctx* ctx_ = nullptr;
};

void check_xrh(xrh_status s);  // non-zero status -> exception (Unsupported / StringMessage)

// ------------------------------------------------------------------------------------------

// Sync-buffer and HDF5 types of xReg, used here only as return types of interface methods the
// HIP classes do not implement (the device buffer hand-off replaces the sync objects).
class RayCastSyncOCLBuf;
class RayCastSyncHostBuf;
// xReg's H5ReadWriteInterface (lib/hdf5/xregHDF5ReadWriteInterface.h): the records below
// derive from it; in the xReg tree their write() / read() are the reference's writers over the
// same fields (this repo has no HDF5 C++ API; xreg_amd.h5io writes the same layout).
class H5ReadWriteInterface {
 public:
  virtual ~H5ReadWriteInterface() = default;
};

// ImgSimMetric2DPatchCommon::PatchInfo (xregImgSimMetric2DPatchCommon.h:49-61): inclusive patch
// bounds and weight.
struct PatchInfo {
  size_type start_row, start_col, stop_row, stop_col;
  float weight;
};
using PatchInfoList = std::vector<PatchInfo>;
using PatchIndexList = std::vector<size_type>;

// ImgSimMetric2DPatchNCCCPU::SimAux (xregImgSimMetric2DPatchNCCCPU.h:57-65): per compute call,
// the patch list and the indices of the patches used.
struct PatchNCCSimAux : public H5ReadWriteInterface {
  std::vector<PatchInfoList> patch_infos_per_compute_call;
  std::vector<PatchIndexList> patch_indices_per_compute_call;
};

// ImgSimMetric2DPatchGradNCCCPU::SimAux: the x- and y-gradient records.
struct PatchGradNCCSimAux : public H5ReadWriteInterface {
  std::shared_ptr<PatchNCCSimAux> sim_aux_x, sim_aux_y;
};

class RayCaster {
 public:
  using PixelScalar2D = float;
  using CameraModelList = std::vector<CameraModel>;
  using CamModelAssocList = std::vector<size_type>;
  class UnsupportedOperationException {};
  enum InterpMethod { kRAY_CAST_INTERP_LINEAR = 0, kRAY_CAST_INTERP_NN, kRAY_CAST_INTERP_SINC, kRAY_CAST_INTERP_BSPLINE };
  enum ProjPixelStoreMethod { kRAY_CAST_PIXEL_REPLACE = 0, kRAY_CAST_PIXEL_ACCUM };

  RayCaster() = default;
  virtual ~RayCaster() = default;
  RayCaster(const RayCaster&) = delete;
  RayCaster& operator=(const RayCaster&) = delete;

  void set_volume(VolPtr v);
  void set_volumes(const VolList& vols);
  size_type num_vols() const { return vols_.size(); }
  size_type num_camera_models() const { return camera_models_.size(); }
  void set_camera_model(const CameraModel& c);
  void set_camera_models(const CameraModelList& c);
  const CameraModelList& camera_models() const { return camera_models_; }
  const CameraModel& camera_model(size_type i = 0) const { return camera_models_[i]; }
  void set_proj_cam_model(size_type proj_idx, size_type cam_idx) { cam_model_for_proj_[proj_idx] = cam_idx; }
  const CamModelAssocList& camera_model_proj_associations() const { return cam_model_for_proj_; }
  void set_camera_model_proj_associations(const CamModelAssocList& a);
  void distribute_xforms_among_cam_models(const FrameTransformList& xforms);
  void distribute_xform_among_cam_models(const FrameTransform& xform);
  void set_ray_step_size(const CoordScalar& s) { ray_step_size_ = s; }
  CoordScalar ray_step_size() const { return ray_step_size_; }
  virtual void set_num_projs(size_type n);
  size_type num_projs() const { return num_projs_; }
  void set_interp_method(const InterpMethod& m) { interp_method_ = m; }
  InterpMethod interp_method() const { return interp_method_; }
  void use_linear_interp() { interp_method_ = kRAY_CAST_INTERP_LINEAR; }
  void use_nn_interp() { interp_method_ = kRAY_CAST_INTERP_NN; }
  void use_sinc_interp() { interp_method_ = kRAY_CAST_INTERP_SINC; }
  void use_bspline_interp() { interp_method_ = kRAY_CAST_INTERP_BSPLINE; }
  void set_xforms_cam_to_itk_phys(const FrameTransformList& xforms);
  const FrameTransformList& xforms_cam_to_itk_phys() const { return xforms_cam_to_itk_phys_; }
  FrameTransform& xform_cam_to_itk_phys(size_type i) { return xforms_cam_to_itk_phys_[i]; }
  const FrameTransform& xform_cam_to_itk_phys(size_type i) const { return xforms_cam_to_itk_phys_[i]; }
  FrameTransform xform_img_center_to_itk_phys(size_type vol_idx = 0) const;
  FrameTransform xform_cam_wrt_carm_center_of_rot(size_type cam_idx = 0) const;
  void post_multiply_all_xforms(const FrameTransform& x);
  void pre_multiply_all_xforms(const FrameTransform& x);
  virtual void allocate_resources();
  virtual void compute(size_type vol_idx = 0) = 0;
  virtual ImagePtr proj(size_type proj_idx) = 0;
  virtual PixelScalar2D* raw_host_pixel_buf() = 0;
  virtual void use_external_host_pixel_buf(void* buf) = 0;
  virtual size_type max_num_projs_possible() const = 0;
  void set_proj_store_method(ProjPixelStoreMethod m) { proj_store_meth_ = m; }
  ProjPixelStoreMethod proj_store_method() const { return proj_store_meth_; }
  void use_proj_store_replace_method() { proj_store_meth_ = kRAY_CAST_PIXEL_REPLACE; }
  void use_proj_store_accum_method() { proj_store_meth_ = kRAY_CAST_PIXEL_ACCUM; }
  virtual void use_other_proj_buf(RayCaster* other) = 0;
  void set_use_bg_projs(bool b) { use_bg_projs_ = b; }
  bool use_bg_projs() const { return use_bg_projs_; }
  void set_bg_proj(ImagePtr& p, bool use = true);
  void set_bg_projs(std::vector<ImagePtr>& ps, bool use = true);
  size_type max_num_projs() const { return max_num_projs_; }
  PixelScalar2D default_bg_pixel_val() const { return default_bg_pixel_val_; }
  void set_default_bg_pixel_val(PixelScalar2D v) { default_bg_pixel_val_ = v; }

  // Device buffer hand-off for HIP similarity metrics (the RayCastSyncOCLBuf role,
  // lib/ray_cast/xregRayCastSyncBuf.h:35-192): the DRRs never leave HBM.
  virtual const float* device_pixel_buf() const { throw UnsupportedOperationException(); }
  // xregRayCastInterface.h:324-328 (the base throws, xregRayCastInterface.cpp:299-307); the HIP
  // caster keeps that behaviour: metrics bind to device_pixel_buf() instead
  virtual RayCastSyncOCLBuf* to_ocl_buf() { throw UnsupportedOperationException(); }
  virtual RayCastSyncHostBuf* to_host_buf() { throw UnsupportedOperationException(); }

 protected:
  virtual void vols_changed() {}
  virtual void camera_models_changed() {}

  VolList vols_;
  CameraModelList camera_models_;
  size_type num_projs_ = 0;
  size_type max_num_projs_ = 0;
  FrameTransformList xforms_cam_to_itk_phys_;
  CoordScalar ray_step_size_ = 1;
  InterpMethod interp_method_ = kRAY_CAST_INTERP_LINEAR;
  CamModelAssocList cam_model_for_proj_;
  PixelScalar2D default_bg_pixel_val_ = 0;
  ProjPixelStoreMethod proj_store_meth_ = kRAY_CAST_PIXEL_REPLACE;
  bool use_bg_projs_ = false;
  bool bg_projs_updated_ = false;
  std::vector<ImagePtr> bg_projs_for_each_cam_;
  bool resources_allocated_ = false;
};

enum RayCastLineIntKernel { kRAY_CAST_LINE_INT_SUM_KERNEL = 0, kRAY_CAST_LINE_INT_MAX_KERNEL };

class RayCastLineIntParamInterface {
 public:
  RayCastLineIntKernel kernel_id() const { return kernel_id_; }
  void set_kernel_id(RayCastLineIntKernel k) { kernel_id_ = k; }

 protected:
  RayCastLineIntKernel kernel_id_ = kRAY_CAST_LINE_INT_SUM_KERNEL;
};

// RayCasterLineIntCPU / RayCasterLineIntOCL counterpart. geom selects the ray set-up
// arithmetic (XRH_GEOM_FP32 matches RayCasterLineIntCPU; XRH_GEOM_FP64 is more exact).
class RayCasterLineIntHIP : public RayCaster, public RayCastLineIntParamInterface {
 public:
  explicit RayCasterLineIntHIP(int device = 0, int32_t geom = XRH_GEOM_FP32, int32_t storage = XRH_F32);
  ~RayCasterLineIntHIP() override;
  void set_num_projs(size_type n) override;
  void allocate_resources() override;
  void compute(size_type vol_idx = 0) override;
  ImagePtr proj(size_type proj_idx) override;
  PixelScalar2D* raw_host_pixel_buf() override;
  void use_external_host_pixel_buf(void* buf) override;
  size_type max_num_projs_possible() const override;
  void use_other_proj_buf(RayCaster* other) override;
  const float* device_pixel_buf() const override { return out_->buf ? out_->buf->dev : nullptr; }
  // floats in the buffer device_pixel_buf() points to (0 before allocate_resources)
  size_type device_pixel_capacity() const { return out_->buf ? out_->buf->capacity : 0; }
  // exact trilinear sample count of the next compute() (roofline accounting)
  uint64_t sample_count(size_type vol_idx = 0);
  // RayCasterCPU::set_anti_alias_factor / anti_alias_factor (xregRayCastBaseCPU.h:113-130):
  // 0 (default) casts one ray per pixel; k >= 1 casts k jittered rays per pixel and averages
  // them (k = 1 still jitters).
  void set_anti_alias_factor(size_type aa_fact) { aa_fact_ = aa_fact; }
  size_type anti_alias_factor() const { return aa_fact_; }
  // The jitter seed (no counterpart in the reference, which seeds from std::random_device on
  // every compute()): fixed, the jitter and so the DRRs repeat exactly; by default every
  // compute() draws a new seed from std::random_device, as the reference does.
  void set_anti_alias_seed(uint64_t seed)
  {
    aa_seed_ = seed;
    aa_seed_fixed_ = true;
  }
  void use_random_anti_alias_seed() { aa_seed_fixed_ = false; }

 protected:
  void vols_changed() override;

 private:
  void sync_to_host();
  std::vector<xrh_camera> xrh_cams() const;
  // A DRR device allocation, freed with its last holder. `version` is a process-wide ticket
  // taken at allocation and at every compute() that writes it (never reused), so every caster
  // reading the buffer knows whether its host copy is current.
  struct DeviceDrr {
    xrh_ctx* ctx = nullptr;
    float* dev = nullptr;
    size_type capacity = 0;  // floats
    uint64_t version = next_version();
    ~DeviceDrr();
    static uint64_t next_version();
  };
  // The slot a caster's allocate_resources() fills. use_other_proj_buf shares the other
  // caster's slot (as RayCasterOCL shares a pointer to the other caster's buffer member, so a
  // later re-allocation by the owner is seen), and the shared slot keeps the allocation alive
  // even if the owner is destroyed first.
  struct DrrSlot {
    std::shared_ptr<DeviceDrr> buf;
  };
  DeviceDrr& out_drr(size_type need) const;  // the buffer compute() writes; throws if too small
  xrh_ctx* ctx_;
  int32_t geom_, storage_;
  std::vector<xrh_volume*> dev_vols_;
  std::shared_ptr<DrrSlot> own_ = std::make_shared<DrrSlot>();
  std::shared_ptr<DrrSlot> out_ = own_;
  float* bg_dev_ = nullptr;
  std::vector<float> host_buf_;
  float* ext_host_buf_ = nullptr;
  bool host_stale_ = true;
  const DeviceDrr* synced_buf_ = nullptr;  // the allocation and version the host copy reflects
  uint64_t synced_version_ = 0;
  size_type aa_fact_ = 0;
  uint64_t aa_seed_ = 0;
  bool aa_seed_fixed_ = false;
};

// ------------------------------------------------------------------------------------------

class ImgSimMetric2D {
 public:
  using Scalar = float;
  using ScalarList = std::vector<Scalar>;
  class UnsupportedOperationException {};
  ImgSimMetric2D() = default;
  virtual ~ImgSimMetric2D() = default;
  ImgSimMetric2D(const ImgSimMetric2D&) = delete;
  ImgSimMetric2D& operator=(const ImgSimMetric2D&) = delete;

  void set_fixed_image(ImagePtr f) { fixed_img_ = f; }
  ImagePtr fixed_image() { return fixed_img_; }
  virtual void allocate_resources() { sim_vals_.resize(num_mov_imgs_); }
  virtual void compute() = 0;
  void set_num_moving_images(size_type n) { num_mov_imgs_ = n; }
  size_type num_moving_images() const { return num_mov_imgs_; }
  Scalar& sim_val(size_type i) { return sim_vals_[i]; }
  const Scalar& sim_val(size_type i) const { return sim_vals_[i]; }
  ScalarList& sim_vals() { return sim_vals_; }
  const ScalarList& sim_vals() const { return sim_vals_; }
  virtual void set_mov_imgs_buf_from_ray_caster(RayCaster* rc, size_type proj_offset = 0);
  virtual void set_mov_imgs_host_buf(Scalar* buf, size_type proj_offset = 0) = 0;
  void set_mask(ImageMaskPtr m);
  ImageMaskPtr mask() { return mask_; }
  void set_save_aux_info(bool s) { save_aux_info_ = s; }
  // ImgSimMetric2D::aux_info (xregImgSimMetric2D.cpp:99-102): none by default; the patch
  // metrics return their per-compute-call patch record when save_aux_info is set
  virtual std::shared_ptr<H5ReadWriteInterface> aux_info() { return nullptr; }

 protected:
  ImagePtr fixed_img_;
  size_type num_mov_imgs_ = 0;
  ScalarList sim_vals_;
  ImageMaskPtr mask_;
  bool mask_updated_ = true;
  bool save_aux_info_ = false;
};

// HIP base: binds the moving images either to a HIP ray caster's device buffer (no copy)
// or to a host buffer (uploaded per compute, like RayCastSyncOCLBuf's host->device path).
class ImgSimMetric2DHIP : public ImgSimMetric2D {
 public:
  explicit ImgSimMetric2DHIP(int32_t metric, int device = 0);
  ~ImgSimMetric2DHIP() override;
  void allocate_resources() override;
  void compute() override;
  void set_mov_imgs_buf_from_ray_caster(RayCaster* rc, size_type proj_offset = 0) override;
  void set_mov_imgs_host_buf(Scalar* buf, size_type proj_offset = 0) override;
  const xrh_fixed* fixed_handle() const { return fixed_; }
  // the patch metrics' record (PatchNCCSimAux / PatchGradNCCSimAux), else null
  std::shared_ptr<H5ReadWriteInterface> aux_info() override { return sim_aux_; }

 protected:
  // hooks of the patch metrics around each xrh_sim_compute call; fixed_gen counts the fixed
  // images allocate_resources has created
  virtual void before_sim(xrh_fixed*, uint64_t /*fixed_gen*/) {}
  virtual void after_sim() {}
  // where a patch metric wants every grid patch's score of this call (num_mov x num_patches,
  // through xrh_sim_patch_scores), or null
  virtual std::vector<float>* patch_score_sink() { return nullptr; }
  xrh_sim_params params_{};
  std::shared_ptr<H5ReadWriteInterface> sim_aux_;

 private:
  xrh_ctx* ctx_;
  xrh_fixed* fixed_ = nullptr;
  xrh_sim_params created_params_{};  // the parameters fixed_ was processed with
  uint64_t fixed_gen_ = 0;
  RayCaster* rc_ = nullptr;
  size_type proj_off_ = 0;
  Scalar* host_mov_ = nullptr;
  float* mov_dev_ = nullptr;
  size_type mov_dev_capacity_ = 0;
};

class ImgSimMetric2DNCCHIP : public ImgSimMetric2DHIP {
 public:
  explicit ImgSimMetric2DNCCHIP(int device = 0) : ImgSimMetric2DHIP(XRH_SIM_NCC, device) {}
};

class ImgSimMetric2DGradImgParamInterface {
 public:
  virtual ~ImgSimMetric2DGradImgParamInterface() = default;
  virtual size_type smooth_img_before_sobel_kernel_radius() const = 0;
  virtual void set_smooth_img_before_sobel_kernel_radius(size_type r) = 0;
};

class ImgSimMetric2DGradNCCHIP : public ImgSimMetric2DHIP, public ImgSimMetric2DGradImgParamInterface {
 public:
  explicit ImgSimMetric2DGradNCCHIP(int device = 0) : ImgSimMetric2DHIP(XRH_SIM_GRAD_NCC, device)
  {
    params_.smooth_ksize = 5;  // xregImgSimMetric2DGradImgCPU.h:81
  }
  size_type smooth_img_before_sobel_kernel_radius() const override { return (size_type)params_.smooth_ksize; }
  void set_smooth_img_before_sobel_kernel_radius(size_type r) override { params_.smooth_ksize = (int32_t)r; }
};

// ImgSimMetric2DPatchCommon (xregImgSimMetric2DPatchCommon.h:38-181, .cpp:27-493): the patch
// grid, the patch weights (a weight image, the mask, normalisation: compute_weights,
// .cpp:306-400), the patches used per compute call (all; a weight-proportional random draw
// with a minimum centre separation, .cpp:402-493; or a caller's list) and how their scores are
// combined (PatchNCCCPU::compute, xregImgSimMetric2DPatchNCCCPU.cpp:186-287). This is host
// bookkeeping: the device scores every patch of the grid and applies the weights and the
// patches used through xrh_fixed_set_patch_weights. The weights are recomputed whenever an
// input to them (weight image, mask, options, grid) changes. Refused with
// UnsupportedOperationException: moving-image patch variances as weights
// (use_mov_img_patch_variances_as_wgts), whose reference weights each patch with the variance
// found at its offset within the parallel chunk that scores it
// (xregImgSimMetric2DPatchNCCCPU.cpp:197, 208-210), so the result depends on the partition.
class ImgSimMetric2DPatchCommonHIP {
 public:
  using WgtImgPtr = ImagePtr;
  virtual ~ImgSimMetric2DPatchCommonHIP() = default;
  size_type patch_radius() const { return (size_type)pp_->patch_radius; }
  void set_patch_radius(size_type r) { pp_->patch_radius = (int32_t)r; }
  size_type patch_stride() const { return (size_type)pp_->patch_stride; }
  void set_patch_stride(size_type s) { pp_->patch_stride = (int32_t)s; }
  bool compute_mean_of_patch_sims() const { return compute_mean_; }
  void set_compute_mean_of_patch_sims(bool b) { compute_mean_ = b; }
  bool weight_patch_sims_in_combine() const { return wic_; }
  void set_weight_patch_sims_in_combine(bool b) { wic_ = b; }
  bool use_mask_for_patch_weighting() const { return use_mask_wgt_; }
  void set_use_mask_for_patch_weighting(bool b);
  // with a mask, patch means and deviations over the unmasked pixels (the fixed image is
  // processed again at the next compute when this changes)
  bool use_mask_for_patch_stats() const { return pp_->patch_mask_stats != 0; }
  void set_use_mask_for_patch_stats(bool u) { pp_->patch_mask_stats = u ? 1 : 0; }
  // ImgSimMetric2DPatchCommon::sim_vals_for_each_patch (xregImgSimMetric2DPatchCommon.h:88-95):
  // per used patch (local index), the unweighted score of every moving image of the last call
  bool save_all_per_patch_scores() const { return save_scores_; }
  void set_save_all_per_patch_scores(bool s) { save_scores_ = s; }
  using ListOfSimScalarLists = std::vector<std::vector<float>>;
  const ListOfSimScalarLists& sim_vals_for_each_patch() const { return per_patch_; }
  bool normalize_weights_as_prob() const { return normalize_; }
  void set_normalize_weights_as_prob(bool b);
  bool choose_rand_patches() const { return choose_rand_; }
  void set_choose_rand_patches(bool b) { choose_rand_ = b; }
  size_type num_rand_patches() const { return num_rand_; }
  void set_num_rand_patches(size_type n) { num_rand_ = n; }
  double rand_patch_min_pixels_sep() const { return rand_sep_; }
  void set_rand_patch_min_pixels_sep(double d) { rand_sep_ = d; }
  // the grid as of the last compute call, with the weights in force
  const PatchInfoList& patch_infos() const { return patch_infos_; }
  // patches used per compute call: the caller's list, the random draw's size, or the grid's
  size_type num_patches() const;
  void set_patches_to_use(const PatchIndexList& inds);
  void reset_patches_to_use();
  void set_wgt_img(WgtImgPtr w);
  WgtImgPtr wgt_img() const { return wgt_img_; }

 protected:
  explicit ImgSimMetric2DPatchCommonHIP(xrh_sim_params* p);
  // before each compute: the grid of a new fixed image, the weights, this call's patches, and
  // the combine handed to the device when it differs from what the fixed image holds
  void prepare_patches(xrh_fixed* fixed, uint64_t fixed_gen, const Image2D& fixed_img,
                       const std::vector<unsigned char>* mask);
  // the SimAux record of this call: the patch list and the indices used (y: with the y
  // direction's weights when they differ)
  void record_patches(std::shared_ptr<PatchNCCSimAux>& a, bool y = false) const;
  bool patches_listed() const { return fixed_inds_; }
  // fills sim_vals_for_each_patch from a call's grid scores (num_mov x num_patches)
  void take_patch_scores(const std::vector<float>& grid, size_type num_mov);
  std::vector<float> grid_scores_;
  // fixed-image patch variances as weights (the reference's process_mask,
  // xregImgSimMetric2DPatchNCCCPU.cpp:331-440, xregImgSimMetric2DPatchGradNCCCPU.cpp:313-400): 0 off,
  // 1 the variances of the fixed image for every direction, 2 each gradient direction the
  // variances of its own fixed gradient image
  void set_variance_weighting(int mode);
  int variance_weighting() const { return var_mode_; }
  xrh_sim_params* pp_;

 private:
  void compute_weights(const Image2D& fixed_img, const std::vector<unsigned char>* mask, const float* wimg,
                       PatchInfoList& out) const;
  std::vector<float> variance_image(const float* img, size_type rows, size_type cols,
                                    const std::vector<unsigned char>* mask) const;
  PatchIndexList draw_random_patches();
  bool compute_mean_ = false, wic_ = true, use_mask_wgt_ = true, normalize_ = true;
  bool choose_rand_ = false;
  bool save_scores_ = false;
  ListOfSimScalarLists per_patch_;
  size_type num_rand_ = 100;
  double rand_sep_ = -1.0;
  bool fixed_inds_ = false;  // set_patches_to_use in force
  PatchIndexList patch_inds_to_use_;
  WgtImgPtr wgt_img_;
  PatchInfoList patch_infos_;
  uint64_t grid_gen_ = 0;       // fixed image the grid belongs to (0: none)
  bool weights_stale_ = true;
  int var_mode_ = 0;
  PatchInfoList patch_infos_y_;  // variance mode 2: the y direction's weights
  std::shared_ptr<void> rng_;   // std::mt19937 (kept out of this header)
  // what the fixed image's device tables hold
  uint64_t applied_gen_ = 0;
  bool applied_default_ = true;
  std::vector<float> applied_w_, applied_wy_;
  PatchIndexList applied_inds_;
  bool applied_wic_ = true, applied_mean_ = false;
};

class ImgSimMetric2DPatchNCCHIP : public ImgSimMetric2DHIP, public ImgSimMetric2DPatchCommonHIP {
 public:
  explicit ImgSimMetric2DPatchNCCHIP(int device = 0)
      : ImgSimMetric2DHIP(XRH_SIM_PATCH_NCC, device), ImgSimMetric2DPatchCommonHIP(&params_) {}
  // ImgSimMetric2DPatchNCCCPU (xregImgSimMetric2DPatchNCCCPU.h:67-75): the fixed image's patch
  // variances (over the unmasked pixels when there is a mask) as the weight image
  bool use_fixed_img_patch_variances_as_wgts() const { return variance_weighting() != 0; }
  void set_use_fixed_img_patch_variances_as_wgts(bool b) { set_variance_weighting(b ? 1 : 0); }
  bool use_mov_img_patch_variances_as_wgts() const { return false; }
  void set_use_mov_img_patch_variances_as_wgts(bool b);

 protected:
  void before_sim(xrh_fixed* f, uint64_t gen) override;
  void after_sim() override;
  std::vector<float>* patch_score_sink() override { return save_all_per_patch_scores() ? &grid_scores_ : nullptr; }
};

class ImgSimMetric2DPatchGradNCCHIP : public ImgSimMetric2DHIP,
                                      public ImgSimMetric2DGradImgParamInterface,
                                      public ImgSimMetric2DPatchCommonHIP {
 public:
  explicit ImgSimMetric2DPatchGradNCCHIP(int device = 0)
      : ImgSimMetric2DHIP(XRH_SIM_PATCH_GRAD_NCC, device), ImgSimMetric2DPatchCommonHIP(&params_)
  {
    params_.smooth_ksize = 5;
  }
  size_type smooth_img_before_sobel_kernel_radius() const override { return (size_type)params_.smooth_ksize; }
  void set_smooth_img_before_sobel_kernel_radius(size_type r) override { params_.smooth_ksize = (int32_t)r; }
  // ImgSimMetric2DPatchGradNCCCPU (xregImgSimMetric2DPatchGradNCCCPU.h:62-64, :100): x and y use
  // the same patches; separate draws are not implemented and are refused at compute when
  // patches are drawn at random or listed by the caller
  bool enforce_same_patches_in_both_x_and_y() const { return same_xy_; }
  void set_enforce_same_patches_in_both_x_and_y(bool b) { same_xy_ = b; }
  // xregImgSimMetric2DPatchGradNCCCPU.h:66-76: fixed-image patch variances as weights, taken
  // from the fixed image for both directions, or with use_variances_in_grad_imgs_as_wgts from
  // each direction's own fixed gradient image
  bool use_fixed_img_patch_variances_as_wgts() const { return fixed_var_; }
  void set_use_fixed_img_patch_variances_as_wgts(bool b);
  bool use_variances_in_grad_imgs_as_wgts() const { return grad_var_; }
  void set_use_variances_in_grad_imgs_as_wgts(bool b);
  bool use_mov_img_patch_variances_as_wgts() const { return false; }
  void set_use_mov_img_patch_variances_as_wgts(bool b);

 protected:
  void before_sim(xrh_fixed* f, uint64_t gen) override;
  void after_sim() override;
  std::vector<float>* patch_score_sink() override { return save_all_per_patch_scores() ? &grid_scores_ : nullptr; }

 private:
  bool same_xy_ = true;
  bool fixed_var_ = false, grad_var_ = false;
};

// ImgSimMetric2DCombineMean (lib/regi/sim_metrics_2d/xregImgSimMetric2DCombine.cpp:79-98)
class ImgSimMetric2DCombineMean {
 public:
  void set_num_sim_metrics(size_type n) { num_sim_metrics_ = n; }
  void set_num_projs_per_sim_metric(size_type n) { num_projs_per_sim_metric_ = n; }
  void allocate_resources();
  void set_sim_metric(size_type i, ImgSimMetric2D* s);
  void compute();
  const std::vector<float>& sim_vals() const { return sim_vals_; }
  float sim_val(size_type i) const { return sim_vals_[i]; }

 private:
  size_type num_sim_metrics_ = 0, num_projs_per_sim_metric_ = 0;
  std::vector<ImgSimMetric2D*> sim_objs_;
  std::vector<float> sim_vals_;
};

// Factories with the "hip" backend (the registration point of xregRayCastProgOpts.cpp:39-70 and
// xregImgSimMetric2DProgOpts.cpp:45-94). Unknown backends throw like xregThrow does there.
std::shared_ptr<RayCaster> LineIntRayCasterFromBackend(const std::string& backend);
std::shared_ptr<ImgSimMetric2D> NCCSimMetricFromBackend(const std::string& backend);
std::shared_ptr<ImgSimMetric2D> GradNCCSimMetricFromBackend(const std::string& backend);
std::shared_ptr<ImgSimMetric2D> PatchNCCSimMetricFromBackend(const std::string& backend);
std::shared_ptr<ImgSimMetric2D> PatchGradNCCSimMetricFromBackend(const std::string& backend);

// Intensity2D3DRegi's evaluation path (xregIntensity2D3DRegi.cpp:43-133 setup, 571-696 obj_fn):
// num_projs = projs_per_view * views, view-major projections, one similarity metric per view
// bound at offset projs_per_view * view, mean over views. Several objects (volumes) are cast
// one after another into the same projections, the first with the REPLACE store and the rest
// with ACCUM; with a static volume the caster's background images (the static volumes' DRRs)
// seed the first cast.
class Intensity2D3DObjFn {
 public:
  Intensity2D3DObjFn(std::shared_ptr<RayCaster> rc, std::vector<std::shared_ptr<ImgSimMetric2D>> sims,
                     size_type projs_per_view, std::vector<size_type> vol_inds_in_ray_caster = {0});
  void setup();
  // has_a_static_vol_ (xregIntensity2D3DRegi.cpp:593-625): the caster's background images are
  // used for the first object's cast only
  void set_has_a_static_vol(bool s) { has_a_static_vol_ = s; }
  // poses: projs_per_view cam->phys transforms of the single object; returns the combined
  // similarity per pose
  std::vector<float> operator()(const FrameTransformList& poses);
  // poses_per_obj[o]: projs_per_view transforms of object o (volume vol_inds_in_ray_caster[o])
  std::vector<float> operator()(const std::vector<FrameTransformList>& poses_per_obj);
  size_type num_obj_fn_evals() const { return num_evals_; }

 private:
  std::shared_ptr<RayCaster> rc_;
  std::vector<std::shared_ptr<ImgSimMetric2D>> sims_;
  size_type projs_per_view_;
  std::vector<size_type> vol_inds_;
  bool has_a_static_vol_ = false;
  ImgSimMetric2DCombineMean combine_;
  size_type num_evals_ = 0;
  bool setup_done_ = false;
};

}  // namespace xreg
