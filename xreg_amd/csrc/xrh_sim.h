// xrh_sim.h -- internal similarity-metric types of libxrh (not part of the ABI).
#pragma once

#include "xrh_internal.h"

namespace xrh {

// Device-side view of a processed fixed image (one per ImgSimMetric2D instance / view).
struct FixedDev {
  int32_t metric;
  int32_t rows, cols;
  int32_t smooth_ksize;
  int32_t patch_radius, patch_stride;
  int32_t ppr, ppc;          // patches per row / per column
  int64_t num_patches;
  int64_t count;             // pixels used by NCC (mask length when masked)
  const float* zf;           // NCC: zero-mean fixed image (Grad-NCC: zero-mean fixed gx)
  const float* zfy;          // Grad-NCC: zero-mean fixed gy
  double sum_zf, sum_zfy;    // FP64 sums of the float zero-mean images (rounding residue)
  double std_f, std_fy;      // fixed standard deviations (clamped at 1e-6)
  const uint8_t* mask;       // optional NCC mask
  const float* pf;           // patch metrics: fixed image (or fixed gx)
  const float* pfy;          // patch-gradient: fixed gy
  const float* pmean;        // per-patch fixed mean
  const float* pden;         // per-patch fixed std * N, divided by the patch's coefficient c
  const float* pmeany;       //   (c = 0: +inf, the patch adds 1 to the sum)
  const float* pdeny;
  const double* pg;          // per-patch sum_U(f) - |U| mean_f over the used pixels U
  const double* pgy;         //   (all pixels unmasked: the float mean's residue)
  // use_mask_for_patch_stats with a mask: patch means and deviations over the unmasked pixels
  // U, whose count per patch is pcount (as float); else null
  const float* pcount;
  const float* pden_unit;    // the unscaled std_p N tables (per-patch scores of any weighting)
  const float* pdeny_unit;
  // The patch kernels sum 1 - X_p / den_p over every patch of the grid, which with
  // den_p = std_p N / c_p is sum_p (1 - c_p X'_p); adding patch_offset = sum_p c_p - P gives
  // the reference's sum_p c_p s_p, then divided by patch_norm. Index 1: Patch-Grad-NCC's y
  // direction (its own coefficients when the directions are weighted differently).
  double patch_offset[2];
  double patch_norm[2];      // divisor of the patch-score sum (patch count, or weight sum)
};

// One moving image to score.
struct SimItem {
  const float* mov;
  const FixedDev* fixed;     // device pointer
  float* gx;                 // scratch for patch-gradient metrics
  float* gy;
  int64_t out_index;
  float* patch_scores;       // general patch kernel: every grid patch's unweighted score, or null
};

int sim_record_stride(int metric);
int sim_tiles(const FixedDev& f);     // tiles of the partial-producing kernel
int sim_img_tiles(const FixedDev& f); // 16x16 image tiles

// Launch-group key of a fixed image: the metric in the low bits; above kSimKeyShift, for
// unmasked stride-1 patch metrics whose diameter has a compiled kernel, that diameter (else
// 0); above kSimKeyBlurShift, the gradient metrics' blur radius (0 for none).
constexpr int kSimKeyShift = 4;
constexpr int kSimKeyBlurShift = 9;
int sim_group_key(const FixedDev& f);

// Queues the similarity kernels for n_items moving images that all share the group key `key`.
// partials: n_items * ptiles * record-stride doubles. If fused_ncc_done, the NCC sums were
// already written by the DRR epilogue.
void launch_sim(int key, const SimItem* items_dev, int64_t n_items, int max_img_tiles,
                int max_patch_tiles, double* partials, float* sim_out_dev, bool fused_ncc_done,
                int patch_lds_bytes, hipStream_t s);
// dynamic LDS bytes the patch kernel needs for a fixed image's patch geometry (general: the
// general kernel's, whatever the group key selects)
int sim_patch_lds_bytes(const FixedDev& f, bool general = false);

}  // namespace xrh

struct xrh_fixed {
  xrh_ctx* ctx = nullptr;
  xrh::FixedDev host{};          // host copy with device pointers
  xrh::FixedDev* dev = nullptr;  // device copy
  std::vector<void*> buffers;    // device allocations owned by this object
  // patch metrics: the weights the image was created with (mask fractions, normalised; or
  // unit), the weights last set, the unscaled per-patch denominators (host copies, one per
  // direction) and their device tables
  std::vector<float> create_weights, weights, weights_y;
  std::vector<float> base_den[2];
  const float* base_den_dev[2] = {nullptr, nullptr};
  float* scaled_den_dev[2] = {nullptr, nullptr};  // allocated on the first weighted set-up
};
