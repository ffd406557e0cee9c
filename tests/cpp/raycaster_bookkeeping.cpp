// raycaster_bookkeeping.cpp -- host-only checks of RayCaster's projection bookkeeping
// (xreg_amd/cpp/xreg_hip.h), run under AddressSanitizer: a caster whose projection count grows
// after allocate_resources (set_num_projs checks the previous count, as the reference's assert
// at xregRayCastInterface.cpp:131-139 does) must keep distribute_xforms_among_cam_models
// (xregRayCastInterface.cpp:97-114) inside its pose table. No GPU is touched: the caster below
// implements the pure virtuals on the host.
#include <cstdio>
#include <stdexcept>

#include "../../xreg_amd/cpp/xreg_hip.h"

using namespace xreg;

namespace {

struct HostCaster : RayCaster {
  void compute(size_type) override {}
  ImagePtr proj(size_type) override { return nullptr; }
  PixelScalar2D* raw_host_pixel_buf() override { return nullptr; }
  void use_external_host_pixel_buf(void*) override {}
  size_type max_num_projs_possible() const override { return 1u << 20; }
  void use_other_proj_buf(RayCaster*) override {}
};

int fails = 0;

void expect(bool ok, const char* what)
{
  if (!ok) {
    std::printf("FAIL: %s\n", what);
    ++fails;
  }
}

}  // namespace

int main()
{
  HostCaster rc;
  CameraModel a, b;
  a.setup(1020.0f, 8, 8, 1.0f, 1.0f);
  b.setup(1020.0f, 8, 8, 1.0f, 1.0f);
  rc.set_camera_models({a, b});
  rc.set_num_projs(2 * 3);
  rc.allocate_resources();
  expect(rc.max_num_projs() == 6, "allocate_resources records the projection count");
  FrameTransformList three(3, FrameTransform::Identity());
  rc.distribute_xforms_among_cam_models(three);
  expect(rc.camera_model_proj_associations()[4] == 1, "view-major camera assignment");

  // grow after allocation: 2 cameras x 40 poses into a table sized for 6 projections
  rc.set_num_projs(2 * 40);
  FrameTransformList forty(40, FrameTransform::Identity());
  forty[39].m[3] = 7.0f;
  rc.distribute_xforms_among_cam_models(forty);  // ASan reports a heap overflow if unbounded
  expect(rc.xforms_cam_to_itk_phys().size() >= 80, "pose table covers the grown count");
  expect(rc.xform_cam_to_itk_phys(79).m[3] == 7.0f, "last pose of the last camera");
  expect(rc.camera_model_proj_associations()[79] == 1, "last projection on camera 1");

  // a count that does not match poses x cameras is refused
  bool threw = false;
  try {
    rc.distribute_xforms_among_cam_models(three);
  } catch (const std::exception&) {
    threw = true;
  }
  expect(threw, "xforms * cameras != num_projs is refused");
  if (fails == 0) std::printf("raycaster bookkeeping: ok\n");
  return fails == 0 ? 0 : 1;
}
