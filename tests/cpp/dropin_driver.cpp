// dropin_driver.cpp -- drives the drop-in HIP classes (xreg_amd/cpp/xreg_hip.h) with the call
// sequence of Intensity2D3DRegi::setup / obj_fn (xregIntensity2D3DRegi.cpp:43-133, 571-696)
// and the factory/dynamic_cast parameter setting of the pelvis single-view app
// (apps/hip_surgery/pelvis_single_view_regi_2d_3d/...main.cpp:259-292).
//
// usage: dropin_driver <in.bin> <out.bin>
// in.bin : int32 X Y Z rows cols n_views n_poses metric smooth patch_radius geom
//          float32 spacing step focal px
//          float32 vol[X*Y*Z] fixed[n_views][rows*cols] extrins[n_views][16] poses[n_poses][16]
// out.bin: float32 sims[n_poses] drr[n_views*n_poses*rows*cols] uint64 samples
//          then (second obj_fn call with the poses reversed) float32 sims2[n_poses]
//          then uint64 shared_bad shared_nonzero (a second caster accumulating into the first
//          one's buffer: pixels != 2 x the reversed-pose DRR, and non-zero pixels)
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <vector>

#include "../../xreg_amd/cpp/xreg_hip.h"

using namespace xreg;

template <class T>
static void rd(std::ifstream& f, T* p, size_t n)
{
  f.read(reinterpret_cast<char*>(p), sizeof(T) * n);
  if (!f) throw std::runtime_error("short input file");
}

int main(int argc, char** argv)
{
  if (argc != 3) {
    std::cerr << "usage: dropin_driver in.bin out.bin\n";
    return 2;
  }
  try {
    std::ifstream in(argv[1], std::ios::binary);
    int32_t hdr[11];
    rd(in, hdr, 11);
    const int X = hdr[0], Y = hdr[1], Z = hdr[2], rows = hdr[3], cols = hdr[4], nv = hdr[5], np = hdr[6];
    const int metric = hdr[7], smooth = hdr[8], pr = hdr[9], geom = hdr[10];
    float fh[4];
    rd(in, fh, 4);
    const float spacing = fh[0], step = fh[1], focal = fh[2], px = fh[3];
    auto vol = std::make_shared<Volume>();
    vol->size = {(size_type)X, (size_type)Y, (size_type)Z};
    vol->spacing = {spacing, spacing, spacing};
    vol->buf.resize((size_t)X * Y * Z);
    rd(in, vol->buf.data(), vol->buf.size());
    std::vector<ImagePtr> fixed(nv);
    for (int v = 0; v < nv; ++v) {
      fixed[v] = std::make_shared<Image2D>();
      fixed[v]->rows = rows;
      fixed[v]->cols = cols;
      fixed[v]->buf.resize((size_t)rows * cols);
      rd(in, fixed[v]->buf.data(), fixed[v]->buf.size());
    }
    std::vector<FrameTransform> extr(nv);
    for (int v = 0; v < nv; ++v) rd(in, extr[v].m.data(), 16);
    FrameTransformList poses(np);
    for (int p = 0; p < np; ++p) rd(in, poses[p].m.data(), 16);

    // ray caster from the backend factory (LineIntRayCasterFromProgOpts with "hip")
    auto rc = LineIntRayCasterFromBackend("hip");
    if (geom) rc = std::make_shared<RayCasterLineIntHIP>(0, XRH_GEOM_FP64);
    rc->set_volume(vol);
    std::vector<CameraModel> cams(nv);
    for (int v = 0; v < nv; ++v) {
      cams[v].setup(focal, rows, cols, px, px);
      if (v > 0) cams[v].setup(cams[v].intrins, extr[v], rows, cols, px, px);
    }
    rc->set_camera_models(cams);
    rc->set_ray_step_size(step);

    std::vector<std::shared_ptr<ImgSimMetric2D>> sims;
    for (int v = 0; v < nv; ++v) {
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
      if (auto* p = dynamic_cast<ImgSimMetric2DPatchCommonHIP*>(s.get())) p->set_patch_radius(pr);
      s->set_fixed_image(fixed[v]);
      sims.push_back(s);
    }

    Intensity2D3DObjFn obj(rc, sims, np);
    obj.setup();
    const std::vector<float> vals = obj(poses);
    const size_t n_pix = (size_t)nv * np * rows * cols;
    const std::vector<float> drr(rc->raw_host_pixel_buf(), rc->raw_host_pixel_buf() + n_pix);
    const uint64_t samples = static_cast<RayCasterLineIntHIP*>(rc.get())->sample_count(0);
    FrameTransformList rev(poses.rbegin(), poses.rend());
    const std::vector<float> vals2 = obj(rev);

    // A second caster accumulating into the first one's device buffer (use_other_proj_buf with
    // the ACCUM store, xregRayCastBaseOCL.cpp:258-269): the shared buffer then holds
    // DRR + DRR == 2 DRR bit for bit, and the owner's host view sees it.
    std::vector<float> drr_rev(rc->raw_host_pixel_buf(), rc->raw_host_pixel_buf() + n_pix);
    uint64_t shared_bad = 0, shared_nonzero = 0;
    {
      RayCasterLineIntHIP rc2(0, geom ? XRH_GEOM_FP64 : XRH_GEOM_FP32);
      rc2.set_volume(vol);
      rc2.set_camera_models(cams);
      rc2.set_ray_step_size(step);
      rc2.set_num_projs(rc->num_projs());
      rc2.allocate_resources();
      rc2.set_camera_model_proj_associations(rc->camera_model_proj_associations());
      rc2.set_xforms_cam_to_itk_phys(rc->xforms_cam_to_itk_phys());
      rc2.use_other_proj_buf(rc.get());
      rc2.use_proj_store_accum_method();
      rc2.compute(0);
      const float* acc = rc->raw_host_pixel_buf();
      for (size_t i = 0; i < n_pix; ++i) {
        shared_bad += (acc[i] != 2.0f * drr_rev[i]);
        shared_nonzero += (acc[i] != 0.0f);
      }
      bool threw = false;
      try {
        rc2.use_other_proj_buf(&rc2);
      } catch (const StringMessageException&) {
        threw = true;
      }
      if (!threw) shared_bad += 1u << 30;
    }
    {
      // the shared allocation outlives its owner: a caster whose owner is destroyed keeps
      // casting into it, and its host view equals the owner's last DRR (REPLACE)
      auto make = [&]() {
        auto c = std::make_unique<RayCasterLineIntHIP>(0, geom ? XRH_GEOM_FP64 : XRH_GEOM_FP32);
        c->set_volume(vol);
        c->set_camera_models(cams);
        c->set_ray_step_size(step);
        c->set_num_projs(rc->num_projs());
        c->allocate_resources();
        c->set_camera_model_proj_associations(rc->camera_model_proj_associations());
        c->set_xforms_cam_to_itk_phys(rc->xforms_cam_to_itk_phys());
        return c;
      };
      auto owner = make();
      auto sharer = make();
      sharer->use_other_proj_buf(owner.get());
      owner.reset();
      sharer->compute(0);
      const float* px = sharer->raw_host_pixel_buf();
      for (size_t i = 0; i < n_pix; ++i) shared_bad += (px[i] != drr_rev[i]);
    }

    std::ofstream out(argv[2], std::ios::binary);
    out.write(reinterpret_cast<const char*>(vals.data()), sizeof(float) * vals.size());
    out.write(reinterpret_cast<const char*>(drr.data()), sizeof(float) * n_pix);
    out.write(reinterpret_cast<const char*>(&samples), sizeof(samples));
    out.write(reinterpret_cast<const char*>(vals2.data()), sizeof(float) * vals2.size());
    out.write(reinterpret_cast<const char*>(&shared_bad), sizeof(shared_bad));
    out.write(reinterpret_cast<const char*>(&shared_nonzero), sizeof(shared_nonzero));
    std::cout << "dropin_driver: " << np << " poses x " << nv << " views, obj_fn evals " << obj.num_obj_fn_evals()
              << "\n";
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
