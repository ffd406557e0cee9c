// dropin_features.cpp -- drives the rest of the drop-in classes' interface (xreg_amd/cpp/xreg_hip.h)
// the way xReg's callers use it:
//   * set_kernel_id(MAX) on the caster (RayCastLineIntParamInterface, xregRayCastInterface.h:581-591);
//   * nearest-neighbour, B-spline and windowed-sinc interpolation (set_interp_method,
//     xregRayCastInterface.h:165-191) and anti-aliasing (set_anti_alias_factor,
//     xregRayCastBaseCPU.h:113-130);
//   * two objects cast into the same projections, REPLACE then ACCUM, through compute(vol_idx)
//     (Intensity2D3DRegi::obj_fn, xregIntensity2D3DRegi.cpp:593-625);
//   * a static volume's DRR as the background of the first cast (has_a_static_vol_, same lines;
//     set_bg_projs, xregRayCastInterface.h:334-349);
//   * use_external_host_pixel_buf (xregRayCastInterface.h:281-287) and proj(i);
//   * a metric reading host moving images (set_mov_imgs_host_buf, xregImgSimMetric2D.h:122);
//   * fewer projections and moving images than allocated, as the exhaustive driver's last batch
//     (xregIntensity2D3DRegiExhaustive.cpp:188-196);
//   * a metric with a mask (set_mask, xregImgSimMetric2D.h:130);
//   * the patch metrics' options (ImgSimMetric2DPatchCommon, xregImgSimMetric2DPatchCommon.h:
//     67-127): a weight image, mask weighting on and off, unnormalised weights, the mean and
//     unweighted combines, a list of patches to use with its per-patch scores, random patches,
//     patch-variance weights and statistics over the unmasked pixels.
//
// usage: dropin_features <in.bin> <out.bin>
// in.bin : int32 X Y Z rows cols n_poses metric smooth
//          float32 spacing step focal px
//          float32 volA[X*Y*Z] volB[X*Y*Z] fixed[rows*cols] bg[rows*cols]
//          uint8 mask[rows*cols]
//          float32 posesA[n_poses][16] posesB[n_poses][16]
// out.bin: records, each uint32 name length, name, uint64 byte count, bytes
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "../../xreg_amd/cpp/xreg_hip.h"

using namespace xreg;

template <class T>
static void rd(std::ifstream& f, T* p, size_t n)
{
  f.read(reinterpret_cast<char*>(p), sizeof(T) * n);
  if (!f) throw std::runtime_error("short input file");
}

struct Records {
  std::ofstream out;
  explicit Records(const char* path) : out(path, std::ios::binary) {}
  void put(const std::string& name, const void* p, uint64_t bytes)
  {
    const uint32_t n = (uint32_t)name.size();
    out.write(reinterpret_cast<const char*>(&n), sizeof(n));
    out.write(name.data(), n);
    out.write(reinterpret_cast<const char*>(&bytes), sizeof(bytes));
    out.write(reinterpret_cast<const char*>(p), (std::streamsize)bytes);
  }
  void put(const std::string& name, const std::vector<float>& v) { put(name, v.data(), v.size() * sizeof(float)); }
  void put_u64(const std::string& name, uint64_t v) { put(name, &v, sizeof(v)); }
};

static std::shared_ptr<ImgSimMetric2D> make_metric(int metric, int smooth)
{
  std::shared_ptr<ImgSimMetric2D> s;
  switch (metric) {
    case 0: s = NCCSimMetricFromBackend("hip"); break;
    case 1: s = GradNCCSimMetricFromBackend("hip"); break;
    case 2: s = PatchNCCSimMetricFromBackend("hip"); break;
    default: s = PatchGradNCCSimMetricFromBackend("hip"); break;
  }
  if (auto* g = dynamic_cast<ImgSimMetric2DGradImgParamInterface*>(s.get())) {
    g->set_smooth_img_before_sobel_kernel_radius(smooth);
  }
  return s;
}

int main(int argc, char** argv)
{
  if (argc != 3) {
    std::cerr << "usage: dropin_features in.bin out.bin\n";
    return 2;
  }
  try {
    std::ifstream in(argv[1], std::ios::binary);
    int32_t hdr[8];
    rd(in, hdr, 8);
    const int X = hdr[0], Y = hdr[1], Z = hdr[2], rows = hdr[3], cols = hdr[4], np = hdr[5];
    const int metric = hdr[6], smooth = hdr[7];
    float fh[4];
    rd(in, fh, 4);
    const float spacing = fh[0], step = fh[1], focal = fh[2], px = fh[3];
    const size_t npx = (size_t)rows * cols;
    VolList vols;
    for (int v = 0; v < 2; ++v) {
      auto vol = std::make_shared<Volume>();
      vol->size = {(size_type)X, (size_type)Y, (size_type)Z};
      vol->spacing = {spacing, spacing, spacing};
      vol->buf.resize((size_t)X * Y * Z);
      rd(in, vol->buf.data(), vol->buf.size());
      vols.push_back(vol);
    }
    auto fixed = std::make_shared<Image2D>();
    fixed->rows = rows;
    fixed->cols = cols;
    fixed->buf.resize(npx);
    rd(in, fixed->buf.data(), npx);
    auto bg = std::make_shared<Image2D>();
    bg->rows = rows;
    bg->cols = cols;
    bg->buf.resize(npx);
    rd(in, bg->buf.data(), npx);
    auto mask = std::make_shared<std::vector<unsigned char>>(npx);
    rd(in, mask->data(), npx);
    FrameTransformList poses_a(np), poses_b(np);
    for (int p = 0; p < np; ++p) rd(in, poses_a[p].m.data(), 16);
    for (int p = 0; p < np; ++p) rd(in, poses_b[p].m.data(), 16);

    CameraModel cam;
    cam.setup(focal, rows, cols, px, px);
    Records rec(argv[2]);

    // one caster for everything, two volumes resident
    auto rc = std::make_shared<RayCasterLineIntHIP>();
    rc->set_volumes(vols);
    rc->set_camera_model(cam);
    rc->set_ray_step_size(step);
    rc->set_num_projs(np);
    rc->allocate_resources();
    rec.put_u64("max_num_projs_possible", rc->max_num_projs_possible());
    // interface methods kept as the reference's base class has them
    uint64_t iface_bad = 0;
    try {
      rc->to_host_buf();
      ++iface_bad;
    } catch (const RayCaster::UnsupportedOperationException&) {
    }
    try {
      rc->to_ocl_buf();
      ++iface_bad;
    } catch (const RayCaster::UnsupportedOperationException&) {
    }

    // maximum-intensity projections of object A
    rc->set_kernel_id(kRAY_CAST_LINE_INT_MAX_KERNEL);
    rc->use_proj_store_replace_method();
    rc->distribute_xforms_among_cam_models(poses_a);
    rc->compute(0);
    rec.put("max_drr", std::vector<float>(rc->raw_host_pixel_buf(), rc->raw_host_pixel_buf() + np * npx));
    rc->set_kernel_id(kRAY_CAST_LINE_INT_SUM_KERNEL);

    // the other interpolators of the CPU caster (set_interp_method, xregRayCastInterface.h:165-191)
    rc->use_nn_interp();
    rc->compute(0);
    rec.put("nn_drr", std::vector<float>(rc->raw_host_pixel_buf(), rc->raw_host_pixel_buf() + np * npx));
    rc->use_bspline_interp();
    rc->compute(0);
    rec.put("bspline_drr", std::vector<float>(rc->raw_host_pixel_buf(), rc->raw_host_pixel_buf() + np * npx));
    rc->set_interp_method(RayCaster::kRAY_CAST_INTERP_SINC);
    rc->compute(0);
    rec.put("sinc_drr", std::vector<float>(rc->raw_host_pixel_buf(), rc->raw_host_pixel_buf() + np * npx));
    rc->use_linear_interp();
    // anti-aliasing (RayCasterCPU::set_anti_alias_factor, xregRayCastBaseCPU.h:113-130) with a
    // fixed jitter seed, then back to one ray per pixel
    rc->set_anti_alias_factor(3);
    rc->set_anti_alias_seed(20240611);
    rc->compute(0);
    rec.put("aa_drr", std::vector<float>(rc->raw_host_pixel_buf(), rc->raw_host_pixel_buf() + np * npx));
    rc->set_anti_alias_factor(0);

    // line integrals of A into an external host buffer; proj(i) reads the same memory
    std::vector<float> ext(np * npx, -1.0f);
    rc->use_external_host_pixel_buf(ext.data());
    rc->compute(0);
    uint64_t ext_bad = (rc->raw_host_pixel_buf() != ext.data());
    for (int p = 0; p < np; ++p) {
      const ImagePtr img = rc->proj(p);
      ext_bad += (img->rows != (size_type)rows || img->cols != (size_type)cols);
      ext_bad += std::memcmp(img->buf.data(), ext.data() + p * npx, npx * sizeof(float)) != 0;
    }
    rec.put_u64("ext_bad", ext_bad);
    rec.put("ext_drr", ext);

    // the same DRRs scored by a metric on the caster's device buffer and by one on the host copy
    auto dev_sim = make_metric(metric, smooth);
    dev_sim->set_fixed_image(fixed);
    dev_sim->set_num_moving_images(np);
    dev_sim->set_mov_imgs_buf_from_ray_caster(rc.get(), 0);
    dev_sim->allocate_resources();
    dev_sim->compute();
    rec.put("dev_sims", dev_sim->sim_vals());
    dev_sim->set_save_aux_info(true);
    iface_bad += (dev_sim->aux_info() != nullptr);
    rec.put_u64("iface_bad", iface_bad);
    auto host_sim = make_metric(metric, smooth);
    host_sim->set_fixed_image(fixed);
    host_sim->set_num_moving_images(np);
    host_sim->set_mov_imgs_host_buf(ext.data(), 0);
    host_sim->allocate_resources();
    host_sim->compute();
    rec.put("host_sims", host_sim->sim_vals());

    // the exhaustive driver's short last batch: fewer projections and moving images than allocated
    const int ns = np - 2;
    rc->set_num_projs(ns);
    rc->distribute_xforms_among_cam_models(FrameTransformList(poses_a.begin(), poses_a.begin() + ns));
    rc->compute(0);
    dev_sim->set_num_moving_images(ns);
    dev_sim->compute();
    rec.put("short_sims", dev_sim->sim_vals());
    rc->set_num_projs(np);

    // the patch metrics' debug record (aux_info, xregImgSimMetric2D.h:134-136): save_aux_info was
    // set before the last compute, so one call's patch list is recorded (x and y for the
    // gradient form)
    {
      uint64_t calls = 0, npatch = 0, bad = 0;
      auto check_rec = [&](const std::shared_ptr<PatchNCCSimAux>& a) {
        if (!a) {
          ++bad;
          return;
        }
        calls = a->patch_infos_per_compute_call.size();
        bad += (a->patch_indices_per_compute_call.size() != calls);
        for (size_t c = 0; c < calls; ++c) {
          const auto& pl = a->patch_infos_per_compute_call[c];
          const auto& ix = a->patch_indices_per_compute_call[c];
          npatch = pl.size();
          bad += (ix.size() != pl.size());
          for (size_t q = 0; q < pl.size(); ++q) {
            bad += (ix[q] != q) + (pl[q].stop_row - pl[q].start_row != 10) + (pl[q].stop_col - pl[q].start_col != 10);
            bad += (pl[q].weight != 1.0f) + (pl[q].stop_row >= (size_type)rows) + (pl[q].stop_col >= (size_type)cols);
          }
        }
      };
      const auto aux = dev_sim->aux_info();
      if (auto g = std::dynamic_pointer_cast<PatchGradNCCSimAux>(aux)) {
        check_rec(g->sim_aux_x);
        check_rec(g->sim_aux_y);
      } else if (auto p1 = std::dynamic_pointer_cast<PatchNCCSimAux>(aux)) {
        check_rec(p1);
      } else if (aux) {
        ++bad;
      }
      rec.put_u64("aux_calls", calls);
      rec.put_u64("aux_patches", npatch);
      rec.put_u64("aux_bad", bad);
    }

    // a masked metric on the same projections (mask changes re-process the fixed image)
    auto mask_sim = make_metric(metric, smooth);
    mask_sim->set_fixed_image(fixed);
    mask_sim->set_num_moving_images(np);
    mask_sim->set_mov_imgs_buf_from_ray_caster(rc.get(), 0);
    mask_sim->allocate_resources();
    mask_sim->set_mask(mask);
    rc->distribute_xforms_among_cam_models(poses_a);
    rc->compute(0);
    mask_sim->compute();
    rec.put("mask_sims", mask_sim->sim_vals());

    // patch options, on the same projections; the weight image is 1 + ((7 r + 3 c) mod 11)
    {
      auto ps = make_metric(metric, smooth);
      auto* pc = dynamic_cast<ImgSimMetric2DPatchCommonHIP*>(ps.get());
      if (pc) {
        auto wimg = std::make_shared<Image2D>();
        wimg->rows = rows;
        wimg->cols = cols;
        wimg->buf.resize(npx);
        for (int r = 0; r < rows; ++r)
          for (int c = 0; c < cols; ++c) wimg->buf[(size_t)r * cols + c] = 1.0f + (float)((7 * r + 3 * c) % 11);
        ps->set_fixed_image(fixed);
        ps->set_num_moving_images(np);
        ps->set_mov_imgs_buf_from_ray_caster(rc.get(), 0);
        ps->allocate_resources();
        pc->set_wgt_img(wimg);
        ps->compute();
        rec.put("popt_wimg_sims", ps->sim_vals());
        std::vector<float> w;
        for (const auto& p : pc->patch_infos()) w.push_back(p.weight);
        rec.put("popt_wimg_weights", w);
        // the mean of the unweighted patch scores
        pc->set_wgt_img(nullptr);
        pc->set_compute_mean_of_patch_sims(true);
        pc->set_weight_patch_sims_in_combine(false);
        ps->compute();
        rec.put("popt_mean_sims", ps->sim_vals());
        pc->set_compute_mean_of_patch_sims(false);
        pc->set_weight_patch_sims_in_combine(true);
        // a list of patches, with repeats, weighted by the weight image
        pc->set_wgt_img(wimg);
        PatchIndexList list;
        for (size_type q = 3; q < pc->patch_infos().size(); q += 37) list.push_back(q);
        list.push_back(list[1]);
        pc->set_patches_to_use(list);
        pc->set_save_all_per_patch_scores(true);
        ps->compute();
        rec.put("popt_list_sims", ps->sim_vals());
        {
          // sim_vals_for_each_patch: per used patch, every moving image's score
          std::vector<float> flat;
          for (const auto& v : pc->sim_vals_for_each_patch()) flat.insert(flat.end(), v.begin(), v.end());
          rec.put("popt_list_scores", flat);
        }
        pc->set_save_all_per_patch_scores(false);
        rec.put("popt_list_inds", std::vector<uint64_t>(list.begin(), list.end()).data(), list.size() * 8);
        rec.put_u64("popt_list_num", pc->num_patches());
        pc->reset_patches_to_use();
        // random patches drawn by the weights; the draw is read back from the debug record
        pc->set_choose_rand_patches(true);
        pc->set_num_rand_patches(40);
        ps->set_save_aux_info(true);
        ps->compute();
        rec.put("popt_rand_sims", ps->sim_vals());
        {
          std::shared_ptr<PatchNCCSimAux> a;
          if (auto g = std::dynamic_pointer_cast<PatchGradNCCSimAux>(ps->aux_info())) {
            a = g->sim_aux_x;
          } else {
            a = std::dynamic_pointer_cast<PatchNCCSimAux>(ps->aux_info());
          }
          const PatchIndexList& ix = a->patch_indices_per_compute_call.back();
          rec.put("popt_rand_inds", std::vector<uint64_t>(ix.begin(), ix.end()).data(), ix.size() * 8);
        }
        ps->set_save_aux_info(false);
        pc->set_choose_rand_patches(false);
        // a mask and the weight image, unnormalised: zero at masked centres
        pc->set_normalize_weights_as_prob(false);
        ps->set_mask(mask);
        ps->compute();
        rec.put("popt_wimg_mask_sims", ps->sim_vals());
        // the mask without mask weighting: unit weights, correlation still over unmasked pixels
        pc->set_wgt_img(nullptr);
        pc->set_normalize_weights_as_prob(true);
        pc->set_use_mask_for_patch_weighting(false);
        ps->compute();
        rec.put("popt_nomaskwgt_sims", ps->sim_vals());
        // every option back at its default: the masked metric's values
        pc->set_use_mask_for_patch_weighting(true);
        ps->compute();
        rec.put("popt_default_sims", ps->sim_vals());
        // the fixed image's patch variances (over unmasked pixels) as the weight image; for
        // Patch-Grad-NCC also each direction's own gradient-image variances
        auto* pn = dynamic_cast<ImgSimMetric2DPatchNCCHIP*>(ps.get());
        auto* pg = dynamic_cast<ImgSimMetric2DPatchGradNCCHIP*>(ps.get());
        uint64_t mov_var_refused = 0;
        try {
          if (pn) pn->set_use_mov_img_patch_variances_as_wgts(true);
          if (pg) pg->set_use_mov_img_patch_variances_as_wgts(true);
        } catch (const ImgSimMetric2D::UnsupportedOperationException&) {
          ++mov_var_refused;
        }
        rec.put_u64("popt_mov_var_refused", mov_var_refused);
        if (pn) pn->set_use_fixed_img_patch_variances_as_wgts(true);
        if (pg) pg->set_use_fixed_img_patch_variances_as_wgts(true);
        ps->compute();
        rec.put("popt_fvar_mask_sims", ps->sim_vals());
        if (pg) {
          pg->set_use_variances_in_grad_imgs_as_wgts(true);
          ps->compute();
          rec.put("popt_gvar_mask_sims", ps->sim_vals());
          pg->set_use_variances_in_grad_imgs_as_wgts(false);
        }
        // patch statistics over the unmasked pixels, with the mask's default weights
        if (pn) pn->set_use_fixed_img_patch_variances_as_wgts(false);
        if (pg) pg->set_use_fixed_img_patch_variances_as_wgts(false);
        pc->set_use_mask_for_patch_stats(true);
        ps->compute();
        rec.put("popt_mstats_sims", ps->sim_vals());
        pc->set_use_mask_for_patch_stats(false);
        ps->compute();
        rec.put("popt_mstats_off_sims", ps->sim_vals());
      }
    }
    rc->use_external_host_pixel_buf(nullptr);

    // two objects through the objective: REPLACE with A, then ACCUM with B
    {
      std::vector<std::shared_ptr<ImgSimMetric2D>> sims{make_metric(metric, smooth)};
      sims[0]->set_fixed_image(fixed);
      Intensity2D3DObjFn obj(rc, sims, np, {0, 1});
      obj.setup();
      const std::vector<float> v = obj({poses_a, poses_b});
      rec.put("two_obj_sims", v);
      rec.put("two_obj_drr", std::vector<float>(rc->raw_host_pixel_buf(), rc->raw_host_pixel_buf() + np * npx));
    }
    // object B registered over a static volume whose DRR is the background image
    {
      rc->set_bg_proj(bg, false);
      std::vector<std::shared_ptr<ImgSimMetric2D>> sims{make_metric(metric, smooth)};
      sims[0]->set_fixed_image(fixed);
      Intensity2D3DObjFn obj(rc, sims, np, {1});
      obj.set_has_a_static_vol(true);
      obj.setup();
      const std::vector<float> v = obj(poses_b);
      rec.put("static_sims", v);
      rec.put("static_drr", std::vector<float>(rc->raw_host_pixel_buf(), rc->raw_host_pixel_buf() + np * npx));
      rec.put_u64("bg_restored", rc->use_bg_projs() ? 0 : 1);
    }
    std::cout << "dropin_features: " << np << " poses, metric " << metric << "\n";
  } catch (const StringMessageException& e) {
    std::cerr << "StringMessageException: " << e.what() << "\n";
    return 1;
  } catch (const RayCaster::UnsupportedOperationException&) {
    std::cerr << "UnsupportedOperationException\n";
    return 1;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
