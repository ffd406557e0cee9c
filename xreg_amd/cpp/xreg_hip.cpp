// xreg_hip.cpp -- drop-in HIP subclasses of RayCaster / ImgSimMetric2D over the C-ABI.
// Plain C++17 (g++): everything device-side happens behind include/xrh.h.
#include "xreg_hip.h"

#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <random>

namespace xreg {

// ------------------------------------------------------------------------------------------
// small value types

FrameTransform FrameTransform::Identity()
{
  FrameTransform t;
  t.m.fill(0.0f);
  t.m[0] = t.m[5] = t.m[10] = t.m[15] = 1.0f;
  return t;
}

FrameTransform FrameTransform::operator*(const FrameTransform& b) const
{
  // Affine * Affine: linear = La*Lb, translation = La*tb + ta (Eigen Transform product)
  FrameTransform r = Identity();
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      r.m[4 * i + j] = (m[4 * i] * b.m[j] + m[4 * i + 1] * b.m[4 + j]) + m[4 * i + 2] * b.m[8 + j];
    }
    r.m[4 * i + 3] = ((m[4 * i] * b.m[3] + m[4 * i + 1] * b.m[7]) + m[4 * i + 2] * b.m[11]) + m[4 * i + 3];
  }
  return r;
}

static std::array<float, 9> mat3_inverse(const std::array<float, 9>& m)
{
  // Eigen's cofactor form: entry = cofactor * (1/det)
  auto M = [&](int r, int c) { return m[3 * r + c]; };
  auto cof = [&](int i, int j) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    return M(i1, j1) * M(i2, j2) - M(i1, j2) * M(i2, j1);
  };
  const float c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
  const float det = (c00 * M(0, 0) + c10 * M(1, 0)) + c20 * M(2, 0);
  if (!(std::fabs(det) > 0.0f)) throw StringMessageException("singular intrinsic matrix");
  const float inv = 1.0f / det;
  return {c00 * inv, c10 * inv, c20 * inv, cof(0, 1) * inv, cof(1, 1) * inv, cof(2, 1) * inv,
          cof(0, 2) * inv, cof(1, 2) * inv, cof(2, 2) * inv};
}

static FrameTransform se3_inv(const FrameTransform& T)
{
  // SE3Inv (lib/transforms/xregRigidUtils.cpp): [R^T | -R^T t]
  FrameTransform r = FrameTransform::Identity();
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[4 * i + j] = T.m[4 * j + i];
  for (int i = 0; i < 3; ++i) {
    const float s = (r.m[4 * i] * T.m[3] + r.m[4 * i + 1] * T.m[7]) + r.m[4 * i + 2] * T.m[11];
    r.m[4 * i + 3] = -1.0f * s;
  }
  return r;
}

void CameraModel::setup(float f, size_type nr, size_type nc, float rs, float cs)
{
  if (!(f > 1.0e-8f) || !nr || !nc || !(rs > 1.0e-8f) || !(cs > 1.0e-8f)) {
    throw StringMessageException("invalid camera model parameters");
  }
  focal_len = f;
  num_det_rows = nr;
  num_det_cols = nc;
  det_row_spacing = rs;
  det_col_spacing = cs;
  // MakeNaiveIntrins (xregPerspectiveXform.cpp:200-228)
  intrins = {f / cs, 0, 0, 0, f / rs, 0, 0, 0, 1};
  if (coord_frame_type == kORIGIN_AT_FOCAL_PT_DET_NEG_Z) {
    intrins[0] *= -1.0f;
    intrins[4] *= -1.0f;
  }
  intrins[2] = (float)((double)(nc - 1) * 0.5);
  intrins[5] = (float)((double)(nr - 1) * 0.5);
  intrins_inv = mat3_inverse(intrins);
  extrins = FrameTransform::Identity();
  extrins_inv = FrameTransform::Identity();
  pinhole_pt = {0, 0, 0};
}

void CameraModel::setup(const std::array<float, 9>& K, const FrameTransform& E, size_type nr, size_type nc,
                        float rs, float cs)
{
  if (!nr || !nc || !(rs > 1.0e-8f) || !(cs > 1.0e-8f)) throw StringMessageException("invalid camera model");
  num_det_rows = nr;
  num_det_cols = nc;
  det_row_spacing = rs;
  det_col_spacing = cs;
  intrins = K;
  intrins_inv = mat3_inverse(K);
  // FocalLenFromIntrins (xregPerspectiveXform.cpp:186-190)
  focal_len = (std::fabs(K[0] * cs) + std::fabs(K[4] * rs)) / 2.0f;
  extrins = E;
  extrins_inv = se3_inv(E);
  if (coord_frame_type != kORIGIN_ON_DETECTOR) {
    pinhole_pt = {extrins_inv.m[3], extrins_inv.m[7], extrins_inv.m[11]};
  } else {
    for (int i = 0; i < 3; ++i) pinhole_pt[i] = extrins_inv.m[4 * i + 2] * focal_len + extrins_inv.m[4 * i + 3];
  }
}

xrh_camera CameraModel::to_xrh() const
{
  xrh_camera c;
  std::memset(&c, 0, sizeof(c));
  c.num_det_rows = (int32_t)num_det_rows;
  c.num_det_cols = (int32_t)num_det_cols;
  c.coord_frame = (int32_t)coord_frame_type;
  c.focal_len = focal_len;
  std::memcpy(c.intrins_inv, intrins_inv.data(), sizeof(c.intrins_inv));
  std::memcpy(c.extrins_inv, extrins_inv.m.data(), sizeof(c.extrins_inv));
  std::memcpy(c.pinhole, pinhole_pt.data(), sizeof(c.pinhole));
  c.det_row_spacing = det_row_spacing;
  c.det_col_spacing = det_col_spacing;
  return c;
}

FrameTransform Volume::idx_to_phys() const
{
  FrameTransform t = FrameTransform::Identity();
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) t.m[4 * r + c] = (float)(direction[3 * r + c] * spacing[c]);
    t.m[4 * r + 3] = (float)origin[r];
  }
  return t;
}

// ------------------------------------------------------------------------------------------
// context + errors

void check_xrh(xrh_status s)
{
  if (s == XRH_OK) return;
  if (s == XRH_ERR_UNSUPPORTED) throw RayCaster::UnsupportedOperationException();
  throw StringMessageException(std::string("HIP backend error: ") + xrh_last_error());
}

HIPContext& HIPContext::get(int device)
{
  static std::mutex mu;
  static std::map<int, std::unique_ptr<HIPContext>> ctxs;
  std::lock_guard<std::mutex> lk(mu);
  auto& p = ctxs[device];
  if (!p) p.reset(new HIPContext(device));
  return *p;
}

HIPContext::HIPContext(int device) { check_xrh(xrh_ctx_create(device, 0, &ctx_)); }

HIPContext::~HIPContext()
{
  if (ctx_) xrh_ctx_destroy(ctx_);
}

// ------------------------------------------------------------------------------------------
// RayCaster base (lib/ray_cast/xregRayCastInterface.cpp:30-351)

void RayCaster::set_volume(VolPtr v)
{
  vols_ = {v};
  vols_changed();
}

void RayCaster::set_volumes(const VolList& v)
{
  vols_ = v;
  vols_changed();
}

void RayCaster::set_camera_model(const CameraModel& c)
{
  camera_models_ = {c};
  camera_models_changed();
}

void RayCaster::set_camera_models(const CameraModelList& c)
{
  camera_models_ = c;
  camera_models_changed();
}

void RayCaster::set_camera_model_proj_associations(const CamModelAssocList& a)
{
  if (a.size() != num_projs_) throw StringMessageException("association list size != num_projs");
  cam_model_for_proj_ = a;
}

void RayCaster::distribute_xforms_among_cam_models(const FrameTransformList& xforms)
{
  const size_type n = xforms.size(), nc = num_camera_models();
  if (n * nc != num_projs()) throw StringMessageException("xforms * cameras != num_projs");
  // set_num_projs may have grown the count after allocate_resources (it checks the previous
  // count, as the reference's assert does): keep the pose table in bounds here and let
  // compute() report that the projection buffer is too small
  if (xforms_cam_to_itk_phys_.size() < num_projs()) xforms_cam_to_itk_phys_.resize(num_projs());
  size_type g = 0;
  for (size_type c = 0; c < nc; ++c) {
    for (size_type p = 0; p < n; ++p, ++g) {
      xforms_cam_to_itk_phys_[g] = xforms[p];
      cam_model_for_proj_[g] = c;
    }
  }
}

void RayCaster::distribute_xform_among_cam_models(const FrameTransform& x)
{
  distribute_xforms_among_cam_models(FrameTransformList(1, x));
}

void RayCaster::set_num_projs(size_type n)
{
  if (max_num_projs_ && num_projs_ > max_num_projs_) throw StringMessageException("too many projections");
  num_projs_ = n;
  cam_model_for_proj_.resize(num_projs_);
}

void RayCaster::set_xforms_cam_to_itk_phys(const FrameTransformList& x)
{
  set_num_projs(x.size());
  xforms_cam_to_itk_phys_ = x;
}

FrameTransform RayCaster::xform_img_center_to_itk_phys(size_type vol_idx) const
{
  const Volume& v = *vols_[vol_idx];
  FrameTransform x = FrameTransform::Identity();
  for (int r = 0; r < 3; ++r) {
    double p = v.origin[r];
    for (int c = 0; c < 3; ++c) p += v.direction[3 * r + c] * v.spacing[c] * (v.size[c] / 2.0);
    x.m[4 * r + 3] = (float)p;
  }
  return x;
}

FrameTransform RayCaster::xform_cam_wrt_carm_center_of_rot(size_type cam_idx) const
{
  FrameTransform x = FrameTransform::Identity();
  x.m[11] -= camera_models_[cam_idx].focal_len / 2;
  return x;
}

void RayCaster::post_multiply_all_xforms(const FrameTransform& x)
{
  for (size_type i = 0; i < num_projs_; ++i) xforms_cam_to_itk_phys_[i] = xforms_cam_to_itk_phys_[i] * x;
}

void RayCaster::pre_multiply_all_xforms(const FrameTransform& x)
{
  for (size_type i = 0; i < num_projs_; ++i) xforms_cam_to_itk_phys_[i] = x * xforms_cam_to_itk_phys_[i];
}

void RayCaster::allocate_resources()
{
  if (!num_projs_) throw StringMessageException("num_projs must be set before allocate_resources");
  if (camera_models_.empty()) throw StringMessageException("camera models must be set before allocate_resources");
  max_num_projs_ = num_projs_;
  xforms_cam_to_itk_phys_.resize(num_projs_);
  cam_model_for_proj_.resize(num_projs_, 0);
  resources_allocated_ = true;
}

void RayCaster::set_bg_proj(ImagePtr& p, bool use)
{
  bg_projs_for_each_cam_.assign(num_camera_models(), p);
  use_bg_projs_ = use;
  bg_projs_updated_ = true;
}

void RayCaster::set_bg_projs(std::vector<ImagePtr>& ps, bool use)
{
  if (ps.size() != num_camera_models()) throw StringMessageException("one background per camera");
  bg_projs_for_each_cam_ = ps;
  use_bg_projs_ = use;
  bg_projs_updated_ = true;
}

// ------------------------------------------------------------------------------------------
// RayCasterLineIntHIP

RayCasterLineIntHIP::RayCasterLineIntHIP(int device, int32_t geom, int32_t storage)
    : ctx_(HIPContext::get(device).ctx()), geom_(geom), storage_(storage)
{
}

RayCasterLineIntHIP::~RayCasterLineIntHIP()
{
  for (auto* v : dev_vols_) xrh_volume_destroy(v);
  if (bg_dev_) xrh_device_free(ctx_, bg_dev_);
}

RayCasterLineIntHIP::DeviceDrr::~DeviceDrr()
{
  if (dev) xrh_device_free(ctx, dev);
}

uint64_t RayCasterLineIntHIP::DeviceDrr::next_version()
{
  static std::atomic<uint64_t> ticket{1};
  return ticket.fetch_add(1, std::memory_order_relaxed);
}

RayCasterLineIntHIP::DeviceDrr& RayCasterLineIntHIP::out_drr(size_type need) const
{
  DeviceDrr* b = out_->buf.get();
  if (!b || need > b->capacity || xforms_cam_to_itk_phys_.size() < num_projs_) {
    // not allocated, set_num_projs grew the count after allocate_resources, or the shared
    // buffer is smaller
    throw StringMessageException("projection buffer is smaller than num_projs projections");
  }
  return *b;
}

void RayCasterLineIntHIP::vols_changed()
{
  for (auto* v : dev_vols_) check_xrh(xrh_volume_destroy(v));
  dev_vols_.clear();
  for (const auto& vp : vols_) {
    const FrameTransform i2p = vp->idx_to_phys();
    const int64_t size[3] = {(int64_t)vp->size[0], (int64_t)vp->size[1], (int64_t)vp->size[2]};
    xrh_volume* h = nullptr;
    check_xrh(xrh_volume_create(ctx_, vp->buf.data(), size, i2p.m.data(), storage_, &h));
    dev_vols_.push_back(h);
  }
}

void RayCasterLineIntHIP::set_num_projs(size_type n)
{
  RayCaster::set_num_projs(n);
  host_stale_ = true;
}

void RayCasterLineIntHIP::allocate_resources()
{
  RayCaster::allocate_resources();
  const size_type rows = camera_models_[0].num_det_rows, cols = camera_models_[0].num_det_cols;
  for (const auto& c : camera_models_) {
    // xregRayCastBaseCPU.cpp:60-70
    if (c.num_det_rows != rows) throw StringMessageException("Inconsistent number rows in camera model images!");
    if (c.num_det_cols != cols) throw StringMessageException("Inconsistent number columns in camera model images!");
  }
  const size_type need = num_projs_ * rows * cols;
  if (!own_->buf || need > own_->buf->capacity) {
    own_->buf.reset();  // a caster still reading the old allocation keeps it alive
    auto b = std::make_shared<DeviceDrr>();
    b->ctx = ctx_;
    void* p = nullptr;
    check_xrh(xrh_device_alloc(ctx_, need * sizeof(float), &p));
    b->dev = (float*)p;
    b->capacity = need;
    own_->buf = std::move(b);
  }
  if (!ext_host_buf_) host_buf_.resize(need);
  host_stale_ = true;
}

std::vector<xrh_camera> RayCasterLineIntHIP::xrh_cams() const
{
  std::vector<xrh_camera> v;
  for (const auto& c : camera_models_) v.push_back(c.to_xrh());
  return v;
}

void RayCasterLineIntHIP::compute(size_type vol_idx)
{
  if (!resources_allocated_) throw StringMessageException("allocate_resources() must be called first");
  if (vol_idx >= dev_vols_.size()) throw StringMessageException("volume index out of range");
  const size_type need = num_projs_ * camera_models_[0].num_det_rows * camera_models_[0].num_det_cols;
  DeviceDrr& out = out_drr(need);
  const auto cams = xrh_cams();
  std::vector<float> xforms(16 * num_projs_);
  for (size_type p = 0; p < num_projs_; ++p) std::memcpy(&xforms[16 * p], xforms_cam_to_itk_phys_[p].m.data(), 64);
  std::vector<int64_t> c4p(cam_model_for_proj_.begin(), cam_model_for_proj_.begin() + num_projs_);
  const float* bg = nullptr;
  if (use_bg_projs_) {
    if (bg_projs_for_each_cam_.size() != num_camera_models()) throw StringMessageException("missing background images");
    const size_type npx = camera_models_[0].num_det_rows * camera_models_[0].num_det_cols;
    if (!bg_dev_ || bg_projs_updated_) {
      if (bg_dev_) check_xrh(xrh_device_free(ctx_, bg_dev_));
      void* p = nullptr;
      check_xrh(xrh_device_alloc(ctx_, npx * num_camera_models() * sizeof(float), &p));
      bg_dev_ = (float*)p;
      std::vector<float> all(npx * num_camera_models());
      for (size_type c = 0; c < num_camera_models(); ++c) {
        std::memcpy(&all[c * npx], bg_projs_for_each_cam_[c]->buf.data(), npx * sizeof(float));
      }
      check_xrh(xrh_copy_h2d(ctx_, bg_dev_, all.data(), all.size() * sizeof(float)));
      bg_projs_updated_ = false;
    }
    bg = bg_dev_;
  }
  const int32_t store = (proj_store_meth_ == kRAY_CAST_PIXEL_ACCUM) ? XRH_STORE_ACCUM : XRH_STORE_REPLACE;
  const bool host_current = !host_stale_ && synced_buf_ == &out && synced_version_ == out.version;
  if (store == XRH_STORE_ACCUM && host_current && ext_host_buf_) {
    // an external host buffer may have been written by the caller: it is the accumulation base
    check_xrh(xrh_copy_h2d(ctx_, out.dev, ext_host_buf_, need * sizeof(float)));
  }
  const int32_t k = (kernel_id_ == kRAY_CAST_LINE_INT_MAX_KERNEL) ? XRH_LINE_INT_MAX : XRH_LINE_INT_SUM;
  if (interp_method_ == kRAY_CAST_INTERP_LINEAR && aa_fact_ == 0) {
    check_xrh(xrh_line_int(ctx_, dev_vols_[vol_idx], cams.data(), (int32_t)cams.size(), xforms.data(), c4p.data(),
                           (int64_t)num_projs_, ray_step_size_, k, store, bg, default_bg_pixel_val_, geom_, out.dev));
  } else {
    // the CPU caster's other ITK interpolators (xregRayCastLineIntCPU.cpp:126-147) and its
    // anti-aliasing (:150-183), cast with the reference's float ray set-up whatever geom_ says
    const int32_t interp = (interp_method_ == kRAY_CAST_INTERP_NN)        ? XRH_INTERP_NN
                           : (interp_method_ == kRAY_CAST_INTERP_SINC)    ? XRH_INTERP_SINC
                           : (interp_method_ == kRAY_CAST_INTERP_BSPLINE) ? XRH_INTERP_BSPLINE
                                                                          : XRH_INTERP_LINEAR;
    if (aa_fact_ > XRH_MAX_AA_RAYS) throw StringMessageException("anti-aliasing factor too large");
    uint64_t seed = aa_seed_;
    if (aa_fact_ && !aa_seed_fixed_) {
      std::random_device rd;
      seed = ((uint64_t)rd() << 32) ^ (uint64_t)rd();
    }
    check_xrh(xrh_line_int_interp(ctx_, dev_vols_[vol_idx], cams.data(), (int32_t)cams.size(), xforms.data(),
                                  c4p.data(), (int64_t)num_projs_, ray_step_size_, k, store, bg,
                                  default_bg_pixel_val_, interp, (int32_t)aa_fact_, seed, out.dev));
  }
  out.version = DeviceDrr::next_version();  // every caster sharing the buffer now sees a stale host copy
}

uint64_t RayCasterLineIntHIP::sample_count(size_type vol_idx)
{
  const size_type need = num_projs_ * camera_models_[0].num_det_rows * camera_models_[0].num_det_cols;
  out_drr(need);  // the same preconditions as compute()
  const auto cams = xrh_cams();
  std::vector<float> xforms(16 * num_projs_);
  for (size_type p = 0; p < num_projs_; ++p) std::memcpy(&xforms[16 * p], xforms_cam_to_itk_phys_[p].m.data(), 64);
  std::vector<int64_t> c4p(cam_model_for_proj_.begin(), cam_model_for_proj_.begin() + num_projs_);
  uint64_t s = 0;
  check_xrh(xrh_line_int_sample_count(ctx_, dev_vols_[vol_idx], cams.data(), (int32_t)cams.size(), xforms.data(),
                                      c4p.data(), (int64_t)num_projs_, ray_step_size_, geom_, &s, nullptr));
  return s;
}

void RayCasterLineIntHIP::sync_to_host()
{
  const size_type n = num_projs_ * camera_models_[0].num_det_rows * camera_models_[0].num_det_cols;
  DeviceDrr& out = out_drr(n);
  if (!host_stale_ && synced_buf_ == &out && synced_version_ == out.version) return;
  if (!ext_host_buf_ && host_buf_.size() < n) host_buf_.resize(n);
  float* dst = ext_host_buf_ ? ext_host_buf_ : host_buf_.data();
  check_xrh(xrh_copy_d2h(ctx_, dst, out.dev, n * sizeof(float)));
  host_stale_ = false;
  synced_buf_ = &out;
  synced_version_ = out.version;
}

ImagePtr RayCasterLineIntHIP::proj(size_type proj_idx)
{
  // a deep copy: the reference's shallow alias is invalidated by the next compute() anyway
  sync_to_host();
  const auto& cam = camera_models_[cam_model_for_proj_[proj_idx]];
  auto img = std::make_shared<Image2D>();
  img->rows = cam.num_det_rows;
  img->cols = cam.num_det_cols;
  img->row_spacing = cam.det_row_spacing;
  img->col_spacing = cam.det_col_spacing;
  const float* src = (ext_host_buf_ ? ext_host_buf_ : host_buf_.data()) + proj_idx * img->rows * img->cols;
  img->buf.assign(src, src + img->rows * img->cols);
  return img;
}

RayCaster::PixelScalar2D* RayCasterLineIntHIP::raw_host_pixel_buf()
{
  sync_to_host();
  return ext_host_buf_ ? ext_host_buf_ : host_buf_.data();
}

void RayCasterLineIntHIP::use_external_host_pixel_buf(void* buf)
{
  ext_host_buf_ = (float*)buf;
  host_stale_ = true;
}

size_type RayCasterLineIntHIP::max_num_projs_possible() const
{
  // bounded by the context budget / device memory rather than a single-allocation cap
  const size_type npx = camera_models_.empty() ? 1 : camera_models_[0].num_det_rows * camera_models_[0].num_det_cols;
  const size_type per = npx * sizeof(float);
  return (size_type)((64ull << 30) / (per ? per : 1));
}

void RayCasterLineIntHIP::use_other_proj_buf(RayCaster* other_ray_caster)
{
  // share another HIP caster's device DRR buffer (RayCasterOCL::use_other_proj_buf,
  // xregRayCastBaseOCL.cpp:258-269): its slot, so the allocation outlives either caster
  auto* other = dynamic_cast<RayCasterLineIntHIP*>(other_ray_caster);
  if (!other || other == this) throw StringMessageException("Incompatible ray caster to share buffer from!");
  out_ = other->own_;
  host_stale_ = true;
}

// ------------------------------------------------------------------------------------------
// ImgSimMetric2D

void ImgSimMetric2D::set_mov_imgs_buf_from_ray_caster(RayCaster*, size_type) { throw UnsupportedOperationException(); }

void ImgSimMetric2D::set_mask(ImageMaskPtr m)
{
  mask_updated_ = mask_updated_ || (m.get() != mask_.get());
  mask_ = m;
}

ImgSimMetric2DHIP::ImgSimMetric2DHIP(int32_t metric, int device) : ctx_(HIPContext::get(device).ctx())
{
  params_.metric = metric;
  params_.smooth_ksize = 0;
  params_.patch_radius = 5;
  params_.patch_stride = 1;
}

ImgSimMetric2DHIP::~ImgSimMetric2DHIP()
{
  if (fixed_) xrh_fixed_destroy(fixed_);
  if (mov_dev_) xrh_device_free(ctx_, mov_dev_);
}

void ImgSimMetric2DHIP::set_mov_imgs_buf_from_ray_caster(RayCaster* rc, size_type proj_offset)
{
  rc_ = rc;
  proj_off_ = proj_offset;
  host_mov_ = nullptr;
}

void ImgSimMetric2DHIP::set_mov_imgs_host_buf(Scalar* buf, size_type proj_offset)
{
  if (rc_) throw StringMessageException("moving images already bound to a ray caster");
  proj_off_ = proj_offset;
  host_mov_ = buf;
}

void ImgSimMetric2DHIP::allocate_resources()
{
  ImgSimMetric2D::allocate_resources();
  if (!fixed_img_) throw StringMessageException("fixed image must be set before allocate_resources");
  if (!rc_ && !host_mov_) throw StringMessageException("moving images must be bound before allocate_resources");
  if (fixed_) check_xrh(xrh_fixed_destroy(fixed_));
  fixed_ = nullptr;
  const unsigned char* mk = mask_ ? mask_->data() : nullptr;
  check_xrh(xrh_fixed_create(ctx_, fixed_img_->buf.data(), (int32_t)fixed_img_->rows, (int32_t)fixed_img_->cols,
                             &params_, mk, &fixed_));
  created_params_ = params_;
  mask_updated_ = false;
  ++fixed_gen_;
}

void ImgSimMetric2DHIP::compute()
{
  if (!fixed_) throw StringMessageException("allocate_resources() must be called first");
  // a new mask, or parameters changed since the fixed image was processed (patch radius,
  // stride, smoothing, use_mask_for_patch_stats), process it again
  if (mask_updated_ || std::memcmp(&params_, &created_params_, sizeof(params_)) != 0) allocate_resources();
  const size_type npx = fixed_img_->rows * fixed_img_->cols;
  const float* mov = nullptr;
  if (rc_) {
    // the moving images must lie inside the caster's device buffer and share the fixed image's
    // size: a kernel must never read past an allocation
    const auto* hip = dynamic_cast<const RayCasterLineIntHIP*>(rc_);
    const float* base = rc_->device_pixel_buf();
    if (!hip || !base || rc_->num_camera_models() == 0) {
      throw StringMessageException("the bound ray caster has no device projection buffer");
    }
    const auto& cam = rc_->camera_model(0);
    if (cam.num_det_rows != fixed_img_->rows || cam.num_det_cols != fixed_img_->cols) {
      throw StringMessageException("fixed image and projection sizes differ");
    }
    if ((proj_off_ + num_mov_imgs_) * npx > hip->device_pixel_capacity()) {
      throw StringMessageException("moving images run past the ray caster's projections");
    }
    mov = base + proj_off_ * npx;
  } else {
    const size_type need = num_mov_imgs_ * npx;
    if (need > mov_dev_capacity_) {
      if (mov_dev_) check_xrh(xrh_device_free(ctx_, mov_dev_));
      void* p = nullptr;
      check_xrh(xrh_device_alloc(ctx_, need * sizeof(float), &p));
      mov_dev_ = (float*)p;
      mov_dev_capacity_ = need;
    }
    check_xrh(xrh_copy_h2d(ctx_, mov_dev_, host_mov_ + proj_off_ * npx, need * sizeof(float)));
    mov = mov_dev_;
  }
  sim_vals_.resize(num_mov_imgs_);
  before_sim(fixed_, fixed_gen_);
  if (std::vector<float>* sink = patch_score_sink()) {
    int64_t np = 0;
    check_xrh(xrh_fixed_patches(fixed_, 0, nullptr, nullptr, nullptr, nullptr, nullptr, &np));
    sink->resize((size_t)np * num_mov_imgs_);
    check_xrh(xrh_sim_patch_scores(ctx_, fixed_, mov, (int64_t)num_mov_imgs_, sim_vals_.data(), sink->data()));
  } else {
    check_xrh(xrh_sim_compute(ctx_, fixed_, mov, (int64_t)num_mov_imgs_, nullptr, sim_vals_.data()));
  }
  after_sim();
}

// ------------------------------------------------------------------------------------------
// ImgSimMetric2DPatchCommon

ImgSimMetric2DPatchCommonHIP::ImgSimMetric2DPatchCommonHIP(xrh_sim_params* p)
    : pp_(p), rng_(std::make_shared<std::mt19937>())
{
  // SeedRNGEngWithRandDev at patch set-up (xregImgSimMetric2DPatchCommon.cpp:262-265)
  std::random_device rd;
  static_cast<std::mt19937*>(rng_.get())->seed(rd());
}

void ImgSimMetric2DPatchCommonHIP::set_use_mask_for_patch_weighting(bool b)
{
  use_mask_wgt_ = b;
  weights_stale_ = true;
}

void ImgSimMetric2DPatchCommonHIP::set_normalize_weights_as_prob(bool b)
{
  normalize_ = b;
  weights_stale_ = true;
}

size_type ImgSimMetric2DPatchCommonHIP::num_patches() const
{
  // xregImgSimMetric2DPatchCommon.cpp:180-184
  return fixed_inds_ ? patch_inds_to_use_.size() : (choose_rand_ ? num_rand_ : patch_infos_.size());
}

void ImgSimMetric2DPatchCommonHIP::set_patches_to_use(const PatchIndexList& inds)
{
  patch_inds_to_use_ = inds;
  fixed_inds_ = true;
}

void ImgSimMetric2DPatchCommonHIP::reset_patches_to_use()
{
  patch_inds_to_use_.clear();
  fixed_inds_ = false;
}

void ImgSimMetric2DPatchCommonHIP::set_wgt_img(WgtImgPtr w)
{
  wgt_img_ = w;
  weights_stale_ = true;
}

void ImgSimMetric2DPatchCommonHIP::set_variance_weighting(int mode)
{
  var_mode_ = mode;
  weights_stale_ = true;
}

void ImgSimMetric2DPatchCommonHIP::compute_weights(const Image2D& img, const std::vector<unsigned char>* mask,
                                                   const float* wimg, PatchInfoList& out) const
{
  // ImgSimMetric2DPatchCommon::compute_weights (xregImgSimMetric2DPatchCommon.cpp:306-400)
  const bool use_mask_wgts = use_mask_wgt_ && mask;
  const bool use_img_wgts = wimg != nullptr;
  for (PatchInfo& p : out) p.weight = 1.0f;
  if (!use_img_wgts && !use_mask_wgts) return;
  const size_type cols = img.cols;
  if (use_img_wgts) {
    for (PatchInfo& p : out) {
      // centre = start + (stop - start + 1) / 2 (PatchInfo::center_row_col, .cpp:33-37)
      const size_type cr = p.start_row + (p.stop_row - p.start_row + 1) / 2;
      const size_type cc = p.start_col + (p.stop_col - p.start_col + 1) / 2;
      p.weight = wimg[cr * cols + cc];
      if (use_mask_wgts && !(*mask)[cr * cols + cc]) p.weight = 0.0f;
    }
  } else {
    for (PatchInfo& p : out) {
      size_type cnt = 0;
      for (size_type r = p.start_row; r <= p.stop_row; ++r)
        for (size_type c = p.start_col; c <= p.stop_col; ++c) cnt += (*mask)[r * cols + c] ? 1 : 0;
      const size_type d = p.stop_row - p.start_row + 1;
      p.weight = (float)cnt / (float)(d * d);
    }
  }
  if (normalize_) {
    float sum = 0.0f;
    for (const PatchInfo& p : out) sum += p.weight;
    for (PatchInfo& p : out) p.weight /= sum;
  }
}

std::vector<float> ImgSimMetric2DPatchCommonHIP::variance_image(const float* img, size_type rows, size_type cols,
                                                                const std::vector<unsigned char>* mask) const
{
  // the fixed patch variances at the patch centres, zero elsewhere: ComputePatchMeanStdDev
  // (xregImgSimMetric2DPatchNCCCPU.cpp:558-619) over the unmasked pixels when there is a mask,
  // in float, squared (xregImgSimMetric2DPatchNCCCPU.cpp:419-432)
  std::vector<float> v(rows * cols, 0.0f);
  for (const PatchInfo& p : patch_infos_) {
    float mean = 0.0f;
    size_type n = 0;
    for (size_type r = p.start_row; r <= p.stop_row; ++r)
      for (size_type c = p.start_col; c <= p.stop_col; ++c)
        if (!mask || (*mask)[r * cols + c]) {
          mean += img[r * cols + c];
          ++n;
        }
    float var = 0.0f;
    if (n > 1) {
      mean /= (float)n;
      for (size_type r = p.start_row; r <= p.stop_row; ++r)
        for (size_type c = p.start_col; c <= p.stop_col; ++c)
          if (!mask || (*mask)[r * cols + c]) {
            const float t = img[r * cols + c] - mean;
            var += t * t;
          }
      var /= (float)n - 1.0f;
    }
    const float sd = std::max(1.0e-6f, std::sqrt(var));
    const size_type cr = p.start_row + (p.stop_row - p.start_row + 1) / 2;
    const size_type cc = p.start_col + (p.stop_col - p.start_col + 1) / 2;
    v[cr * cols + cc] = sd * sd;
  }
  return v;
}

PatchIndexList ImgSimMetric2DPatchCommonHIP::draw_random_patches()
{
  // ImgSimMetric2DPatchCommon::patch_indices_to_use (xregImgSimMetric2DPatchCommon.cpp:402-493):
  // draws proportional to the weights; a draw closer than the minimum separation to a patch
  // already drawn is rejected and drawn again. The reference retries without bound; here a
  // draw that cannot be completed within 1000 tries per patch throws.
  const size_type np = patch_infos_.size();
  if (num_rand_ >= np) throw StringMessageException("the number of random patches must be below the patch count");
  std::vector<double> w;
  w.reserve(np);
  double tot = 0.0;
  for (const PatchInfo& p : patch_infos_) {
    if (!(p.weight >= 0.0f) || !std::isfinite(p.weight)) {
      throw StringMessageException("random patches need finite, non-negative patch weights");
    }
    w.push_back(p.weight);
    tot += p.weight;
  }
  if (!(tot > 0.0)) throw StringMessageException("random patches need a positive weight sum");
  std::discrete_distribution<size_type> dist(w.begin(), w.end());
  auto& eng = *static_cast<std::mt19937*>(rng_.get());
  const double r = (double)pp_->patch_radius;
  const double min_sep = (rand_sep_ < 0.0) ? std::sqrt(2.0 * r * r) : rand_sep_;
  const bool check_sep = min_sep > 1.0e-8;
  PatchIndexList inds;
  inds.reserve(num_rand_);
  std::vector<std::array<double, 2>> centres;
  size_type tries = 0;
  const size_type max_tries = 1000 * (num_rand_ + 1);
  while (inds.size() < num_rand_) {
    if (++tries > max_tries) {
      throw StringMessageException("cannot place " + std::to_string(num_rand_) +
                                   " random patches at the minimum separation");
    }
    const size_type q = dist(eng);
    if (check_sep) {
      const PatchInfo& p = patch_infos_[q];
      const double cr = (double)(p.start_row + (p.stop_row - p.start_row + 1) / 2);
      const double cc = (double)(p.start_col + (p.stop_col - p.start_col + 1) / 2);
      bool ok = true;
      for (const auto& e : centres) {
        if (std::hypot(e[0] - cr, e[1] - cc) < min_sep) {
          ok = false;
          break;
        }
      }
      if (!ok) continue;
      centres.push_back({cr, cc});
    }
    inds.push_back(q);
  }
  return inds;
}

void ImgSimMetric2DPatchCommonHIP::prepare_patches(xrh_fixed* fixed, uint64_t gen, const Image2D& img,
                                                  const std::vector<unsigned char>* mask)
{
  if (gen != grid_gen_) {
    // the grid of a new fixed image (setup_patches, xregImgSimMetric2DPatchCommon.cpp:257-298)
    int64_t n = 0;
    check_xrh(xrh_fixed_patches(fixed, 0, nullptr, nullptr, nullptr, nullptr, nullptr, &n));
    std::vector<int32_t> r0((size_t)n), c0((size_t)n), r1((size_t)n), c1((size_t)n);
    check_xrh(xrh_fixed_patches(fixed, n, r0.data(), c0.data(), r1.data(), c1.data(), nullptr, &n));
    patch_infos_.assign((size_t)n, PatchInfo{});
    for (size_t q = 0; q < (size_t)n; ++q) {
      patch_infos_[q] = {(size_type)r0[q], (size_type)c0[q], (size_type)r1[q], (size_type)c1[q], 1.0f};
    }
    grid_gen_ = gen;
    weights_stale_ = true;
    std::random_device rd;
    static_cast<std::mt19937*>(rng_.get())->seed(rd());
  }
  if (weights_stale_) {
    patch_infos_y_.clear();
    if (var_mode_ == 0) {
      if (wgt_img_ && (wgt_img_->rows != img.rows || wgt_img_->cols != img.cols)) {
        throw StringMessageException("weight image and fixed image sizes differ");
      }
      compute_weights(img, mask, wgt_img_ ? wgt_img_->buf.data() : nullptr, patch_infos_);
    } else if (var_mode_ == 1) {
      // the reference installs the variances as the weight image (set_wgt_img, :435)
      const std::vector<float> v = variance_image(img.buf.data(), img.rows, img.cols, mask);
      compute_weights(img, mask, v.data(), patch_infos_);
    } else {
      std::vector<float> g(img.rows * img.cols);
      check_xrh(xrh_fixed_image(fixed, 0, g.data()));
      const std::vector<float> vx = variance_image(g.data(), img.rows, img.cols, mask);
      check_xrh(xrh_fixed_image(fixed, 1, g.data()));
      const std::vector<float> vy = variance_image(g.data(), img.rows, img.cols, mask);
      compute_weights(img, mask, vx.data(), patch_infos_);
      patch_infos_y_ = patch_infos_;
      compute_weights(img, mask, vy.data(), patch_infos_y_);
    }
    weights_stale_ = false;
  }
  const size_type np = patch_infos_.size();
  if (fixed_inds_) {
    for (size_type q : patch_inds_to_use_) {
      if (q >= np) throw StringMessageException("patch index to use out of range");
    }
  } else if (choose_rand_) {
    patch_inds_to_use_ = draw_random_patches();
  } else {
    patch_inds_to_use_.resize(np);
    for (size_type q = 0; q < np; ++q) patch_inds_to_use_[q] = q;
  }
  // the fixed image's creation weights are the reference defaults: unit, or the normalised
  // unmasked fractions of the mask
  const bool default_weights =
      var_mode_ == 0 && !wgt_img_ && (mask ? (use_mask_wgt_ && normalize_) : true);
  const bool all = !fixed_inds_ && !choose_rand_;
  const bool dflt = default_weights && all && wic_ && !compute_mean_;
  std::vector<float> w, wy;
  if (!dflt) {
    w.resize(np);
    for (size_type q = 0; q < np; ++q) w[q] = patch_infos_[q].weight;
    if (!patch_infos_y_.empty()) {
      wy.resize(np);
      for (size_type q = 0; q < np; ++q) wy[q] = patch_infos_y_[q].weight;
    }
  }
  const PatchIndexList& inds = all ? PatchIndexList() : patch_inds_to_use_;
  const bool same = applied_gen_ == gen && applied_default_ == dflt &&
                    (dflt || (applied_w_ == w && applied_wy_ == wy && applied_inds_ == inds && applied_wic_ == wic_ &&
                              applied_mean_ == compute_mean_));
  if (same) return;
  if (dflt) {
    check_xrh(xrh_fixed_set_patch_weights(fixed, nullptr, nullptr, nullptr, 0, 1, 0));
  } else {
    std::vector<int64_t> i64(inds.begin(), inds.end());
    check_xrh(xrh_fixed_set_patch_weights(fixed, w.data(), wy.empty() ? nullptr : wy.data(),
                                          all ? nullptr : i64.data(), (int64_t)i64.size(), wic_ ? 1 : 0,
                                          compute_mean_ ? 1 : 0));
  }
  applied_gen_ = gen;
  applied_default_ = dflt;
  applied_w_ = std::move(w);
  applied_wy_ = std::move(wy);
  applied_inds_ = inds;
  applied_wic_ = wic_;
  applied_mean_ = compute_mean_;
}

void ImgSimMetric2DPatchCommonHIP::record_patches(std::shared_ptr<PatchNCCSimAux>& a, bool y) const
{
  // ImgSimMetric2DPatchNCCCPU::compute with save_aux_info (xregImgSimMetric2DPatchNCCCPU.cpp:
  // 290-298): every call appends the patch list and the indices used
  if (!a) a = std::make_shared<PatchNCCSimAux>();
  a->patch_infos_per_compute_call.push_back((y && !patch_infos_y_.empty()) ? patch_infos_y_ : patch_infos_);
  a->patch_indices_per_compute_call.push_back(patch_inds_to_use_);
}

void ImgSimMetric2DPatchNCCHIP::set_use_mov_img_patch_variances_as_wgts(bool b)
{
  if (b) throw ImgSimMetric2D::UnsupportedOperationException();
}

void ImgSimMetric2DPatchGradNCCHIP::set_use_fixed_img_patch_variances_as_wgts(bool b)
{
  fixed_var_ = b;
  set_variance_weighting(fixed_var_ ? (grad_var_ ? 2 : 1) : 0);
}

void ImgSimMetric2DPatchGradNCCHIP::set_use_variances_in_grad_imgs_as_wgts(bool b)
{
  grad_var_ = b;
  set_variance_weighting(fixed_var_ ? (grad_var_ ? 2 : 1) : 0);
}

void ImgSimMetric2DPatchGradNCCHIP::set_use_mov_img_patch_variances_as_wgts(bool b)
{
  if (b) throw ImgSimMetric2D::UnsupportedOperationException();
}

void ImgSimMetric2DPatchNCCHIP::before_sim(xrh_fixed* f, uint64_t gen)
{
  prepare_patches(f, gen, *fixed_img_, mask_.get());
}

void ImgSimMetric2DPatchCommonHIP::take_patch_scores(const std::vector<float>& grid, size_type num_mov)
{
  // per local index k of the patches used, the scores of patch_inds_to_use_[k]
  // (xregImgSimMetric2DPatchNCCCPU.cpp:247-250); the reference leaves the entries of patches the
  // combine skips as they were, these hold their scores as well
  const size_type np = patch_infos_.size();
  per_patch_.assign(patch_inds_to_use_.size(), std::vector<float>(num_mov));
  for (size_type k = 0; k < patch_inds_to_use_.size(); ++k) {
    for (size_type m = 0; m < num_mov; ++m) per_patch_[k][m] = grid[m * np + patch_inds_to_use_[k]];
  }
}

void ImgSimMetric2DPatchNCCHIP::after_sim()
{
  if (save_all_per_patch_scores()) take_patch_scores(grid_scores_, num_mov_imgs_);
  if (!save_aux_info_) return;
  auto a = std::static_pointer_cast<PatchNCCSimAux>(sim_aux_);
  record_patches(a);
  sim_aux_ = a;
}

void ImgSimMetric2DPatchGradNCCHIP::before_sim(xrh_fixed* f, uint64_t gen)
{
  if (!same_xy_ && (choose_rand_patches() || patches_listed())) {
    throw ImgSimMetric2D::UnsupportedOperationException();
  }
  prepare_patches(f, gen, *fixed_img_, mask_.get());
}

void ImgSimMetric2DPatchGradNCCHIP::after_sim()
{
  if (save_all_per_patch_scores()) take_patch_scores(grid_scores_, num_mov_imgs_);
  if (!save_aux_info_) return;
  // the x and y sub-metrics keep their records (xregImgSimMetric2DPatchGradNCCCPU.cpp:246-252):
  // the same patches, and the same weights unless each direction is weighted by its own
  // gradient image's variances
  auto g = std::static_pointer_cast<PatchGradNCCSimAux>(sim_aux_);
  if (!g) g = std::make_shared<PatchGradNCCSimAux>();
  record_patches(g->sim_aux_x);
  record_patches(g->sim_aux_y, true);
  sim_aux_ = g;
}

// ------------------------------------------------------------------------------------------

void ImgSimMetric2DCombineMean::allocate_resources()
{
  sim_objs_.resize(num_sim_metrics_);
  sim_vals_.resize(num_projs_per_sim_metric_);
}

void ImgSimMetric2DCombineMean::set_sim_metric(size_type i, ImgSimMetric2D* s)
{
  sim_objs_[i] = s;
  if (s->num_moving_images() != num_projs_per_sim_metric_) throw StringMessageException("metric size mismatch");
}

void ImgSimMetric2DCombineMean::compute()
{
  sim_vals_.assign(num_projs_per_sim_metric_, 0.0f);
  for (size_type s = 0; s < num_sim_metrics_; ++s)
    for (size_type p = 0; p < num_projs_per_sim_metric_; ++p) sim_vals_[p] += sim_objs_[s]->sim_val(p);
  for (size_type p = 0; p < num_projs_per_sim_metric_; ++p) sim_vals_[p] /= num_sim_metrics_;
}

// ------------------------------------------------------------------------------------------

std::shared_ptr<RayCaster> LineIntRayCasterFromBackend(const std::string& b)
{
  if (b == "hip") return std::make_shared<RayCasterLineIntHIP>();
  throw StringMessageException("Unsupported backend for Line Int. Ray Caster: " + b);
}

template <class T>
static std::shared_ptr<ImgSimMetric2D> sim_from(const std::string& b)
{
  if (b == "hip") return std::make_shared<T>();
  throw StringMessageException("Unsupported backend for Sim Metric: " + b);
}

std::shared_ptr<ImgSimMetric2D> NCCSimMetricFromBackend(const std::string& b) { return sim_from<ImgSimMetric2DNCCHIP>(b); }
std::shared_ptr<ImgSimMetric2D> GradNCCSimMetricFromBackend(const std::string& b)
{
  return sim_from<ImgSimMetric2DGradNCCHIP>(b);
}
std::shared_ptr<ImgSimMetric2D> PatchNCCSimMetricFromBackend(const std::string& b)
{
  return sim_from<ImgSimMetric2DPatchNCCHIP>(b);
}
std::shared_ptr<ImgSimMetric2D> PatchGradNCCSimMetricFromBackend(const std::string& b)
{
  return sim_from<ImgSimMetric2DPatchGradNCCHIP>(b);
}

// ------------------------------------------------------------------------------------------

Intensity2D3DObjFn::Intensity2D3DObjFn(std::shared_ptr<RayCaster> rc, std::vector<std::shared_ptr<ImgSimMetric2D>> sims,
                                       size_type projs_per_view, std::vector<size_type> vol_inds_in_ray_caster)
    : rc_(std::move(rc)), sims_(std::move(sims)), projs_per_view_(projs_per_view),
      vol_inds_(std::move(vol_inds_in_ray_caster))
{
  if (vol_inds_.empty()) throw StringMessageException("at least one object to register");
}

void Intensity2D3DObjFn::setup()
{
  // Intensity2D3DRegi::setup (xregIntensity2D3DRegi.cpp:43-133)
  const size_type nv = sims_.size();
  if (rc_->num_camera_models() != nv) throw StringMessageException("one camera model per view");
  for (size_type vi : vol_inds_) {
    if (vi >= rc_->num_vols()) throw StringMessageException("object volume index out of range");
  }
  rc_->set_num_projs(projs_per_view_ * nv);
  rc_->allocate_resources();
  for (size_type v = 0; v < nv; ++v) {
    sims_[v]->set_num_moving_images(projs_per_view_);
    sims_[v]->set_mov_imgs_buf_from_ray_caster(rc_.get(), projs_per_view_ * v);  // view-major
  }
  for (size_type v = 0; v < nv; ++v) sims_[v]->allocate_resources();
  combine_.set_num_sim_metrics(nv);
  combine_.set_num_projs_per_sim_metric(projs_per_view_);
  combine_.allocate_resources();
  for (size_type v = 0; v < nv; ++v) combine_.set_sim_metric(v, sims_[v].get());
  num_evals_ = 0;
  setup_done_ = true;
}

std::vector<float> Intensity2D3DObjFn::operator()(const FrameTransformList& poses)
{
  return (*this)(std::vector<FrameTransformList>{poses});
}

std::vector<float> Intensity2D3DObjFn::operator()(const std::vector<FrameTransformList>& poses_per_obj)
{
  // Intensity2D3DRegi::obj_fn (xregIntensity2D3DRegi.cpp:571-696)
  if (!setup_done_) throw StringMessageException("setup() first");
  if (poses_per_obj.size() != vol_inds_.size()) throw StringMessageException("one pose list per object");
  for (const auto& poses : poses_per_obj) {
    if (poses.size() != projs_per_view_) throw StringMessageException("one pose per projection of a view");
  }
  const bool orig_use_bg = rc_->use_bg_projs();
  if (has_a_static_vol_) rc_->set_use_bg_projs(true);
  rc_->use_proj_store_replace_method();
  for (size_type o = 0; o < vol_inds_.size(); ++o) {
    rc_->distribute_xforms_among_cam_models(poses_per_obj[o]);
    rc_->compute(vol_inds_[o]);
    if (has_a_static_vol_) rc_->set_use_bg_projs(false);
    rc_->use_proj_store_accum_method();
  }
  for (auto& s : sims_) s->compute();
  combine_.compute();
  if (has_a_static_vol_) rc_->set_use_bg_projs(orig_use_bg);
  ++num_evals_;
  return combine_.sim_vals();
}

}  // namespace xreg
