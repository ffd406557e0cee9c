// xrh_internal.h -- internal types shared by the HIP sources of libxrh (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/xrh.h"

namespace xrh {

// ------------------------------------------------------------------------------------
// errors: every C entry point catches and converts to a status + thread-local message.

void set_error(const std::string& msg);

struct Error {
  xrh_status status;
  std::string msg;
};

[[noreturn]] void fail(xrh_status s, const std::string& msg);

#define XRH_HIP_CHECK(expr)                                                             \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ::xrh::fail(XRH_ERR_HIP, std::string(#expr " failed: ") + hipGetErrorString(e_) + \
                                   " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
    }                                                                                   \
  } while (0)

#define XRH_REQUIRE(cond, msg)                                  \
  do {                                                          \
    if (!(cond)) ::xrh::fail(XRH_ERR_INVALID, std::string(msg)); \
  } while (0)

template <class F>
xrh_status guarded(F&& f)
{
  try {
    f();
    return XRH_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.status;
  } catch (const std::exception& e) {
    set_error(std::string("internal error: ") + e.what());
    return XRH_ERR_INVALID;
  } catch (...) {
    set_error("internal error: unknown exception");
    return XRH_ERR_INVALID;
  }
}

// ------------------------------------------------------------------------------------
// volume layouts in HBM. Every layout carries a one-voxel replicate apron on each side so
// that trilinear samples within one voxel of the index box [0, N-1] read edge values
// (ITK's clamp of the upper neighbour, xregRayCastLineIntCPU.cpp:273-274) without any
// branch; the gathers are buffer loads whose range check turns any stray offset into 0.
enum Layout : int32_t {
  kLayoutLinear = 0,  // padded x-fastest scalars: 8 loads of b_v bytes per sample
  kLayoutQuad = 1,    // f32 only: float4 (v[i,j,k], v[i+1,j,k], v[i,j+1,k], v[i+1,j+1,k]);
                      //   2 x 16-byte loads per sample (4x memory)
  kLayoutOct = 2,     // f16/bf16 only: all 8 corners in 16 bytes; 1 load per sample
  kLayoutCoef = 3,    // f32 only: the 8 coefficients of the cell's trilinear polynomial in
                      //   its local coordinates (32 bytes, 2 adjacent 16-byte loads per
                      //   sample, 8x memory); evaluation is 7 FMAs instead of 7 lerps
  kLayoutFace = 4,    // f32 only: bilinear coefficients of each voxel's face normal to x
                      //   (over y, z; 16 bytes, 4x memory); a cell is the faces at x and x+1,
                      //   two adjacent 16-byte loads, bilinear twice then linear along x
  kLayoutBrick = 5    // f32 only: the x-face elements stored brick by brick, bricks of 4x4x4
                      //   cells in x-fastest brick order, each holding its 5x4x4 faces (the
                      //   next brick's first x-face repeated, so a cell's two faces stay
                      //   adjacent); 5x memory. Same samples as kLayoutFace.
};
// brick geometry of kLayoutBrick: cells per brick edge and faces per brick
constexpr int32_t kBrickEdge = 4;
constexpr int32_t kBrickFaces = (kBrickEdge + 1) * kBrickEdge * kBrickEdge;

struct CamDev {
  float kinv[9];
  float einv[12];  // 3x4 row-major
  float pinhole[3];
  float det_z;     // (frame == NEG_Z ? -1 : 1) * focal
  float z_adj;     // -focal for ORIGIN_ON_DETECTOR, else 0
  int32_t rows;
  int32_t cols;
  double kinv_d[9];
  double einv_d[12];
  double pinhole_d[3];
  double det_z_d;
  double z_adj_d;
};

// Per-projection work description, produced on the device by the set-up kernel.
struct DrrItem {
  float M[12];       // phys->idx composed with cam->phys, float (reference op order)
  double Md[12];     // same in double
  CamDev cam;
  const void* vol;   // padded layout base
  uint32_t vol_bytes;
  int32_t X, Y, Z;   // unpadded index extent
  uint32_t pitch_y;  // padded row pitch in elements (X+2)
  uint32_t pitch_z;  // padded slice pitch in elements ((X+2)*(Y+2))
  float step;
  int64_t out_off;   // element offset of this projection's DRR
  const float* ncc_zf;  // zero-mean fixed image for the fused NCC epilogue, or null
  const float* bg;      // background image the first object of an item is added onto, or null
  int32_t tiles_x;
  int32_t tiles_y;
  const uint8_t* skip;  // empty-space map (see VolDesc), or null
  int32_t skip_gx, skip_gy, skip_gz, skip_shift;
};

// Host-side pose record uploaded per projection before the set-up kernel runs.
struct PoseIn {
  double pose[16];     // row-major cam->phys
  int32_t cam;         // index into the camera table
  int32_t vol;         // index into the volume table
  int64_t out_off;
  const float* ncc_zf;
  const float* bg;     // background image (first object of a multi-object item), or null
  float step;
  int32_t pad_;
};

struct VolDesc {
  double p2i_d[12];
  float p2i[12];
  const void* data;
  uint32_t bytes;
  int32_t X, Y, Z;
  uint32_t pitch_y, pitch_z;
  int32_t storage;
  int32_t layout;
  float min_spacing;  // smallest voxel edge (physical units), bounds the ray step from below
  float max_chord;    // upper bound of any chord through the index box (physical units)
  int32_t pad_;
  // empty-space map: Chebyshev distance (in bricks of 2^skip_shift voxels, capped at 255)
  // from each brick to the nearest brick holding a non-zero voxel within one voxel of it;
  // null when skipping is disabled
  const uint8_t* skip;
  int32_t skip_gx, skip_gy, skip_gz, skip_shift;
};

// Rejects steps so small that a ray would take millions of samples (keeps kernels short).
inline void check_step(float step, const VolDesc& v)
{
  if (!(step >= 1.0e-3f * v.min_spacing)) {
    throw Error{XRH_ERR_INVALID, "ray step must be at least 1e-3 of the smallest voxel spacing"};
  }
  // sample indices stay below 2^23 (the kernels count samples in float)
  if (!((double)v.max_chord / (double)step < 8388608.0)) {
    throw Error{XRH_ERR_INVALID, "ray step too small for the volume: over 2^23 samples per ray"};
  }
}

}  // namespace xrh

// ------------------------------------------------------------------------------------
// opaque objects

namespace xrh {
// How a DRR launch is laid out (results are identical under every variant):
//   reuse  1: one sample per loop iteration, loads only when a sample enters a new cell;
//          0: two samples per iteration, every sample loads; -1: chosen by launch size
//   shape  pixels per wave64 inside the 16x16 workgroup tile: 0 = 8x8, 1 = 16x4, 2 = 4x16
//   bands  detector-row bands of the XCD-aware workgroup order
struct DrrVariant {
  int32_t reuse = -1;
  int32_t shape = 0;
  int32_t bands = 8;
};
}  // namespace xrh

struct xrh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t budget = 0;
  uint64_t in_use = 0;
  std::mutex mu;
  std::unordered_map<void*, uint64_t> allocs;
  int num_cus = 0;
  // reusable staging for xrh_line_int / xrh_sim_compute
  void* scratch = nullptr;
  uint64_t scratch_bytes = 0;
  void* pinned = nullptr;
  uint64_t pinned_bytes = 0;
  // DRR variants chosen by xrh_batch_autotune, keyed by drr_tune_key
  std::map<std::string, xrh::DrrVariant> tuned;
};

namespace xrh {
// A voxel grid for the non-trilinear interpolators (float voxels, or double B-spline
// coefficients), x fastest, with an apron of `apron` elements on every side: element
// (x + apron) + py (y + apron) + pz (z + apron) belongs to voxel (x, y, z).
struct InterpGrid {
  const void* data = nullptr;
  uint32_t bytes = 0;
  uint32_t py = 0, pz = 0;
  int32_t apron = 0;
  int32_t X = 0, Y = 0, Z = 0;  // the volume's extent (voxels)
};
// grids of a volume: the voxels with a zero apron (nearest neighbour, windowed sinc) and the
// cubic B-spline coefficients with a mirrored apron
constexpr int kGridVoxels = 0;
constexpr int kGridBSpline = 1;
}  // namespace xrh

struct xrh_volume {
  xrh_ctx* ctx = nullptr;
  xrh::VolDesc desc{};
  void* data = nullptr;
  void* skip = nullptr;
  uint64_t bytes = 0;
  int64_t size[3] = {0, 0, 0};
  // built on first use by an interpolating cast (xrh_line_int_interp), freed with the volume
  xrh::InterpGrid grid[2];
  void* grid_mem[2] = {nullptr, nullptr};
  uint64_t grid_bytes[2] = {0, 0};
};

namespace xrh {

// context-budgeted allocation helpers (defined in xrh_runtime.hip)
void* ctx_alloc(xrh_ctx* ctx, uint64_t bytes);
void ctx_free(xrh_ctx* ctx, void* p);
// Fails with XRH_ERR_INVALID when [p, p + bytes) starts inside one of the context's
// allocations and runs past its end; pointers the context did not allocate pass unchecked.
void ctx_check_range(xrh_ctx* ctx, const void* p, uint64_t bytes, const char* what);
// grows ctx->scratch to at least `bytes` (device) and returns it
void* ctx_scratch(xrh_ctx* ctx, uint64_t bytes);
void* ctx_pinned(xrh_ctx* ctx, uint64_t bytes);

// host-side helpers (xrh_runtime.hip)
CamDev make_cam_dev(const xrh_camera& c);
void check_camera(const xrh_camera& c);

// DRR launch (xrh_drr.hip). `poses_dev` holds num_items PoseIn; `items_dev` receives
// DrrItem; `tile_start_dev` holds the band-major tile prefix sums built by build_tiles().
struct TilePlan {
  int32_t bands = 8;
  int64_t max_band_tiles = 0;
  std::vector<uint32_t> band_start;  // [bands][num_items + 1]
};
TilePlan build_tiles(const std::vector<int32_t>& rows, const std::vector<int32_t>& cols,
                     int32_t bands);

struct DrrLaunch {
  const PoseIn* poses_dev;
  const VolDesc* vols_dev;
  const CamDev* cams_dev;
  DrrItem* items_dev;
  const uint32_t* band_start_dev;
  int64_t num_items;
  int32_t bands;
  int64_t max_band_tiles;
  int32_t geom;
  int32_t kernel;    // SUM / MAX
  int32_t store;     // REPLACE: write op(fill, v); ACCUM: op(out, v) (bg already copied in)
  float fill;        // default background pixel value for REPLACE
  int32_t storage;   // all volumes of one launch share storage type + layout
  int32_t layout;
  float* out;
  // NCC epilogue (optional): items with a non-null ncc_zf write FP64 partial sums
  double* ncc_partials;     // [num_items][max tiles per item][4], or null
  int32_t ncc_tiles_per_item;
  uint64_t* sample_counter; // optional: count samples instead of writing pixels; [0] = all
                            // samples of the reference algorithm, [1] = samples evaluated
  int32_t* ray_steps;       // optional with sample_counter: per pixel (at out_off), the last
                            // sample index n of the reference algorithm (-1: the ray misses)
  int32_t reuse;            // DrrVariant::reuse
  int32_t wave_shape;       // DrrVariant::shape
  int32_t objs;             // objects per item (0 or 1: one); items_dev / poses_dev then hold
                            // num_items * objs entries, object-minor, cast in object order into
                            // the item's pixels (first onto its background or fill, the rest
                            // added), REPLACE only
};
// Key of the tuned-variant table: storage, layout, geometry, line kernel, the largest
// detector of the launch and its workgroups per CU (rounded down to a power of two).
std::string drr_tune_key(int32_t storage, int32_t layout, int32_t geom, int32_t kernel, int32_t max_rows,
                         int32_t max_cols, int64_t workgroups);
// The cell-reuse choice of an untuned launch of `workgroups` workgroups (reuse = -1).
bool drr_default_reuse(int64_t workgroups);
constexpr int32_t kMaxDrrBands = 16;
void launch_drr_setup(const DrrLaunch& L, hipStream_t s);
void launch_drr(const DrrLaunch& L, hipStream_t s);
// Casts with RayCasterLineIntCPU's sampling options (xrh_interp.hip). interp_prepare builds the
// grid `interp` reads, once per volume (synchronous; allocates from the volume's context; the
// linear interpolator reads the nearest-neighbour grid); launch_drr_interp runs the cast of a
// set-up launch with FP32 ray geometry (L.geom is ignored) on that grid, `aa_rays` jittered
// rays per pixel (0: one ray through the pixel index, no jitter).
const InterpGrid& interp_prepare(xrh_volume* vol, int32_t interp);
void launch_drr_interp(const DrrLaunch& L, const InterpGrid& g, int32_t interp, int32_t aa_rays, uint64_t aa_seed,
                       hipStream_t s);

}  // namespace xrh
