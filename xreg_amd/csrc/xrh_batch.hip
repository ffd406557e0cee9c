// xrh_batch.hip -- one optimiser generation (all poses x all views) as one submission.
//
// Per evaluation: one pinned upload of the pose records / tile plan / similarity items, the
// per-item set-up kernel (FP64 pose composition), one DRR launch over every tile of every
// item (uneven detector sizes allowed; NCC fused into the DRR epilogue), the similarity
// kernels grouped by metric, the FP64 finish, one download of the similarity values. The
// whole sequence is captured in a hipGraph the first time an item->view signature is seen
// and replayed afterwards (poses change between generations, shapes do not).
//
// A batch is sized once (views, items, detector rows x cols) and its view table can be
// replaced between evaluations: a serving loop merges whichever registrations are ready into
// one submission without reallocating device memory.
#include "xrh_internal.h"
#include "xrh_sim.h"

#include <algorithm>
#include <cstring>
#include <map>

using namespace xrh;

struct xrh_batch {
  xrh_ctx* ctx = nullptr;
  int32_t geom = XRH_GEOM_FP32;
  // capacity
  int32_t cap_views = 0;
  int64_t max_items = 0;
  int32_t cap_rows = 0, cap_cols = 0;
  int64_t cap_npix = 0;
  int cap_tiles = 0;
  // current view table
  std::vector<xrh_view> views;
  std::vector<int32_t> view_vol;   // view -> index into the volume table
  int32_t storage = XRH_F32, layout = kLayoutQuad;
  int max_img_tiles = 0, max_patch_tiles = 0, max_lds = 0;
  bool fuse_ncc = true;            // NCC sums in the DRR epilogue (off if any NCC view is masked)
  uint64_t version = 0;
  bool use_graph = true;
  bool time_kernels = false;
  // device
  VolDesc* vols_dev = nullptr;
  CamDev* cams_dev = nullptr;
  char* blob_dev = nullptr;        // poses | band plan | sim items
  DrrItem* items_dev = nullptr;
  float* drr_dev = nullptr;
  float* grad_dev = nullptr;
  double* partials_dev = nullptr;
  double* ncc_partials_dev = nullptr;
  float* sims_dev = nullptr;
  // host
  char* blob_host = nullptr;       // pinned
  float* sims_host = nullptr;      // pinned
  hipEvent_t ev_drr0 = nullptr, ev_drr1 = nullptr, ev_all0 = nullptr, ev_all1 = nullptr;
  bool timed_last = false;
  // instantiated graphs per submission signature, least recently used evicted beyond
  // kMaxCachedGraphs (a server whose item-to-view assignment changes every step would
  // otherwise keep one graph per distinct assignment)
  struct CachedGraph {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    uint64_t last_use = 0;
  };
  std::map<std::string, CachedGraph> graphs;
  uint64_t graph_clock = 0;
  int64_t last_n = 0;
  bool pending = false;
  // DRR launch variant: set explicitly / by xrh_batch_autotune, else looked up in the
  // context's tuned table, else the defaults
  bool variant_set = false;
  DrrVariant variant;
  DrrVariant last_variant;         // the variant of the last staged launch
};

namespace {

uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

struct BlobLayout {
  uint64_t off_pose, off_band, off_sim, total;
  uint64_t band_words;
};

BlobLayout blob_layout(int64_t max_items)
{
  BlobLayout L;
  L.band_words = (uint64_t)kMaxDrrBands * (max_items + 1);
  L.off_pose = 0;
  L.off_band = align256(sizeof(PoseIn) * max_items);
  L.off_sim = align256(L.off_band + sizeof(uint32_t) * L.band_words);
  L.total = align256(L.off_sim + sizeof(SimItem) * max_items);
  return L;
}

constexpr size_t kMaxCachedGraphs = 32;

void clear_graphs(xrh_batch* b)
{
  for (auto& kv : b->graphs) {
    (void)hipGraphExecDestroy(kv.second.exec);
    (void)hipGraphDestroy(kv.second.graph);
  }
  b->graphs.clear();
}

// Destroys least recently used graphs until at most `keep` remain (the stream has been
// synchronised, so none of them is executing).
void evict_graphs(xrh_batch* b, size_t keep)
{
  while (b->graphs.size() > keep) {
    auto lru = b->graphs.begin();
    for (auto it = b->graphs.begin(); it != b->graphs.end(); ++it) {
      if (it->second.last_use < lru->second.last_use) lru = it;
    }
    (void)hipGraphExecDestroy(lru->second.exec);
    (void)hipGraphDestroy(lru->second.graph);
    b->graphs.erase(lru);
  }
}

void batch_free(xrh_batch* b)
{
  if (!b) return;
  (void)hipStreamSynchronize(b->ctx->stream);
  clear_graphs(b);
  for (void* p : {(void*)b->vols_dev, (void*)b->cams_dev, (void*)b->blob_dev, (void*)b->items_dev,
                  (void*)b->drr_dev, (void*)b->grad_dev, (void*)b->partials_dev, (void*)b->ncc_partials_dev,
                  (void*)b->sims_dev}) {
    if (p) ctx_free(b->ctx, p);
  }
  if (b->blob_host) (void)hipHostFree(b->blob_host);
  if (b->sims_host) (void)hipHostFree(b->sims_host);
  for (hipEvent_t e : {b->ev_drr0, b->ev_drr1, b->ev_all0, b->ev_all1}) {
    if (e) (void)hipEventDestroy(e);
  }
  delete b;
}

void batch_alloc(xrh_ctx* ctx, int32_t max_views, int64_t max_items, int32_t max_rows, int32_t max_cols,
                 int32_t geom, xrh_batch** out)
{
  XRH_REQUIRE(ctx && out, "null argument");
  XRH_REQUIRE(max_views >= 1 && max_views <= 4096, "max_views must be in [1, 4096]");
  XRH_REQUIRE(max_items >= 1 && max_items < 65536, "max_items must be in [1, 65535]");
  XRH_REQUIRE(max_rows >= 2 && max_cols >= 2 && max_rows <= 32768 && max_cols <= 32768, "bad detector capacity");
  XRH_REQUIRE(geom == XRH_GEOM_FP32 || geom == XRH_GEOM_FP64, "bad geometry mode");
  auto* b = new xrh_batch();
  b->ctx = ctx;
  b->geom = geom;
  b->cap_views = max_views;
  b->max_items = max_items;
  b->cap_rows = max_rows;
  b->cap_cols = max_cols;
  b->cap_npix = (int64_t)max_rows * max_cols;
  b->cap_tiles = ((max_rows + 15) / 16) * ((max_cols + 15) / 16);
  try {
    b->vols_dev = (VolDesc*)ctx_alloc(ctx, sizeof(VolDesc) * max_views);
    b->cams_dev = (CamDev*)ctx_alloc(ctx, sizeof(CamDev) * max_views);
    const BlobLayout L = blob_layout(max_items);
    b->blob_dev = (char*)ctx_alloc(ctx, L.total);
    b->items_dev = (DrrItem*)ctx_alloc(ctx, sizeof(DrrItem) * max_items);
    b->drr_dev = (float*)ctx_alloc(ctx, sizeof(float) * max_items * b->cap_npix);
    b->grad_dev = (float*)ctx_alloc(ctx, sizeof(float) * 2 * max_items * b->cap_npix);
    b->partials_dev = (double*)ctx_alloc(ctx, sizeof(double) * 8 * b->cap_tiles * max_items);
    b->ncc_partials_dev = (double*)ctx_alloc(ctx, sizeof(double) * 4 * b->cap_tiles * max_items);
    b->sims_dev = (float*)ctx_alloc(ctx, sizeof(float) * max_items);
    XRH_HIP_CHECK(hipHostMalloc(&b->blob_host, L.total, hipHostMallocDefault));
    XRH_HIP_CHECK(hipHostMalloc(&b->sims_host, sizeof(float) * max_items, hipHostMallocDefault));
    XRH_HIP_CHECK(hipEventCreate(&b->ev_drr0));
    XRH_HIP_CHECK(hipEventCreate(&b->ev_drr1));
    XRH_HIP_CHECK(hipEventCreate(&b->ev_all0));
    XRH_HIP_CHECK(hipEventCreate(&b->ev_all1));
  } catch (...) {
    batch_free(b);
    throw;
  }
  *out = b;
}

void set_views(xrh_batch* b, const xrh_view* views, int32_t num_views)
{
  XRH_REQUIRE(b && views, "null argument");
  XRH_REQUIRE(!b->pending, "batch has a launch in flight");
  XRH_REQUIRE(num_views >= 1 && num_views <= b->cap_views, "view count exceeds the batch capacity");
  std::vector<CamDev> cams;
  std::vector<VolDesc> vds;
  std::vector<const xrh_volume*> vols;
  std::vector<int32_t> view_vol;
  int max_img = 0, max_patch = 0, max_lds = 0;
  bool fuse = true;
  for (int32_t v = 0; v < num_views; ++v) {
    const xrh_view& w = views[v];
    XRH_REQUIRE(w.vol && w.fixed, "view needs a volume and a fixed image");
    XRH_REQUIRE(w.vol->ctx == b->ctx && w.fixed->ctx == b->ctx, "view objects belong to another context");
    XRH_REQUIRE(w.kernel == XRH_LINE_INT_SUM || w.kernel == XRH_LINE_INT_MAX, "bad line-integral kernel");
    XRH_REQUIRE(w.step > 0.0f, "ray step must be positive");
    check_step(w.step, w.vol->desc);
    XRH_REQUIRE(w.cam.num_det_rows == w.fixed->host.rows && w.cam.num_det_cols == w.fixed->host.cols,
                "fixed image size differs from the view's detector");
    XRH_REQUIRE(w.cam.num_det_rows <= b->cap_rows && w.cam.num_det_cols <= b->cap_cols,
                "view detector exceeds the batch capacity");
    XRH_REQUIRE(w.kernel == views[0].kernel, "all views of one batch must share the line-integral kernel");
    cams.push_back(make_cam_dev(w.cam));
    auto it = std::find(vols.begin(), vols.end(), w.vol);
    if (it == vols.end()) {
      view_vol.push_back((int32_t)vols.size());
      vols.push_back(w.vol);
      vds.push_back(w.vol->desc);
    } else {
      view_vol.push_back((int32_t)(it - vols.begin()));
    }
    const FixedDev& f = w.fixed->host;
    max_img = std::max(max_img, sim_img_tiles(f));
    max_patch = std::max(max_patch, sim_tiles(f));
    max_lds = std::max(max_lds, sim_patch_lds_bytes(f));
    if (f.metric == XRH_SIM_NCC && f.mask) fuse = false;
  }
  for (const VolDesc& d : vds) {
    XRH_REQUIRE(d.storage == vds[0].storage && d.layout == vds[0].layout,
                "all volumes of one batch must share storage type and layout");
  }
  hipStream_t s = b->ctx->stream;
  XRH_HIP_CHECK(hipStreamSynchronize(s));
  XRH_HIP_CHECK(hipMemcpyAsync(b->vols_dev, vds.data(), sizeof(VolDesc) * vds.size(), hipMemcpyHostToDevice, s));
  XRH_HIP_CHECK(hipMemcpyAsync(b->cams_dev, cams.data(), sizeof(CamDev) * cams.size(), hipMemcpyHostToDevice, s));
  XRH_HIP_CHECK(hipStreamSynchronize(s));
  b->views.assign(views, views + num_views);
  b->view_vol = view_vol;
  b->storage = vds[0].storage;
  b->layout = vds[0].layout;
  b->max_img_tiles = max_img;
  b->max_patch_tiles = max_patch;
  b->max_lds = max_lds;
  b->fuse_ncc = fuse;
  ++b->version;
  clear_graphs(b);
}

void batch_create(xrh_ctx* ctx, const xrh_view* views, int32_t num_views, int64_t max_items, int32_t geom,
                  xrh_batch** out)
{
  XRH_REQUIRE(views && num_views >= 1, "need at least one view");
  int32_t rows = 2, cols = 2;
  for (int32_t v = 0; v < num_views; ++v) {
    rows = std::max(rows, views[v].cam.num_det_rows);
    cols = std::max(cols, views[v].cam.num_det_cols);
  }
  xrh_batch* b = nullptr;
  batch_alloc(ctx, num_views, max_items, rows, cols, geom, &b);
  try {
    set_views(b, views, num_views);
  } catch (...) {
    batch_free(b);
    throw;
  }
  *out = b;
}

// Tuned-variant key of a launch over items of these detector sizes.
std::string tune_key(const xrh_batch* b, const std::vector<int32_t>& rows, const std::vector<int32_t>& cols)
{
  int32_t mr = 0, mc = 0;
  int64_t wgs = 0;
  for (size_t i = 0; i < rows.size(); ++i) {
    mr = std::max(mr, rows[i]);
    mc = std::max(mc, cols[i]);
    wgs += (int64_t)((rows[i] + 15) / 16) * ((cols[i] + 15) / 16);
  }
  return drr_tune_key(b->storage, b->layout, b->geom, b->views[0].kernel, mr, mc, wgs);
}

// Fills the pinned staging for n items of `objs` entries each (entry i * objs + k is object k of
// item i; one entry per item in the single-object case); returns the DRR launch description.
// bg: NULL, or background images of the views back to back in view order, the first object of
// an item being cast onto its view's image.
DrrLaunch stage(xrh_batch* b, const int32_t* item_view, const double* item_pose, int64_t n, float* drr_out,
                std::vector<std::pair<int, std::vector<int64_t>>>& groups, int32_t objs = 1,
                const float* bg = nullptr)
{
  XRH_REQUIRE(!b->views.empty(), "batch has no views");
  XRH_REQUIRE(objs >= 1 && n * objs <= b->max_items, "items x objects exceed the batch capacity (max_items)");
  const BlobLayout L = blob_layout(b->max_items);
  PoseIn* pin = (PoseIn*)(b->blob_host + L.off_pose);
  SimItem* sin = (SimItem*)(b->blob_host + L.off_sim);
  float* drr = drr_out ? drr_out : b->drr_dev;
  std::vector<int32_t> rows((size_t)n), cols((size_t)n);
  std::vector<int64_t> bg_off;
  if (bg) {
    // view v's image starts after those of views 0..v-1; the buffer must reach the end of the
    // last image an item's first object uses
    int64_t o = 0, need = 0;
    for (const xrh_view& w : b->views) {
      bg_off.push_back(o);
      o += (int64_t)w.cam.num_det_rows * w.cam.num_det_cols;
    }
    for (int64_t i = 0; i < n; ++i) {
      const int32_t v = item_view[i * objs];
      XRH_REQUIRE(v >= 0 && v < (int32_t)b->views.size(), "item view index out of range");
      need = std::max<int64_t>(need, bg_off[(size_t)v] + (int64_t)b->views[v].cam.num_det_rows *
                                                             b->views[v].cam.num_det_cols);
    }
    ctx_check_range(b->ctx, bg, (uint64_t)need * 4, "background images");
  }
  for (int64_t e = 0; e < n * objs; ++e) {
    const int32_t v = item_view[e];
    XRH_REQUIRE(v >= 0 && v < (int32_t)b->views.size(), "item view index out of range");
    const xrh_view& w = b->views[v];
    const xrh_view& w0 = b->views[item_view[e - e % objs]];
    XRH_REQUIRE(w.cam.num_det_rows == w0.cam.num_det_rows && w.cam.num_det_cols == w0.cam.num_det_cols,
                "the objects of one item must share the detector size");
  }
  int64_t off = 0;
  groups.clear();
  for (int64_t i = 0; i < n; ++i) {
    const int32_t v = item_view[i * objs];
    const xrh_view& w = b->views[v];
    const FixedDev& f = w.fixed->host;
    for (int32_t k = 0; k < objs; ++k) {
      const int64_t e = i * objs + k;
      const int32_t vk = item_view[e];
      PoseIn& p = pin[e];
      std::memcpy(p.pose, item_pose + 16 * e, sizeof(double) * 16);
      p.cam = vk;
      p.vol = b->view_vol[vk];
      p.out_off = off;
      p.ncc_zf = (k == 0 && f.metric == XRH_SIM_NCC && b->fuse_ncc) ? f.zf : nullptr;
      p.bg = (k == 0 && bg) ? bg + bg_off[(size_t)v] : nullptr;
      p.step = b->views[vk].step;
      p.pad_ = 0;
    }
    rows[i] = w.cam.num_det_rows;
    cols[i] = w.cam.num_det_cols;
    SimItem& s = sin[i];
    s.mov = drr + off;
    s.fixed = w.fixed->dev;
    s.gx = b->grad_dev + 2 * i * b->cap_npix;
    s.gy = b->grad_dev + (2 * i + 1) * b->cap_npix;
    s.out_index = i;
    s.patch_scores = nullptr;
    off += (int64_t)rows[i] * cols[i];
  }
  XRH_REQUIRE(drr_out || off <= b->max_items * b->cap_npix, "DRR arena overflow");
  if (drr_out) ctx_check_range(b->ctx, drr_out, (uint64_t)off * 4, "DRR output");
  // items grouped by launch key (metric, then kernel variant), in item order within a group
  std::vector<int> keys;
  for (int64_t i = 0; i < n; ++i) keys.push_back(sim_group_key(b->views[item_view[i * objs]].fixed->host));
  std::vector<int> distinct(keys);
  std::sort(distinct.begin(), distinct.end());
  distinct.erase(std::unique(distinct.begin(), distinct.end()), distinct.end());
  for (int key : distinct) {
    std::vector<int64_t> idx;
    for (int64_t i = 0; i < n; ++i) {
      if (keys[(size_t)i] == key) idx.push_back(i);
    }
    groups.emplace_back(key, std::move(idx));
  }
  // SimItem table laid out group by group
  std::vector<SimItem> tmp(sin, sin + n);
  int64_t k = 0;
  for (auto& g : groups) {
    for (int64_t i : g.second) sin[k++] = tmp[(size_t)i];
  }
  DrrVariant var = b->variant;
  if (!b->variant_set) {
    var = DrrVariant();
    const std::string key = tune_key(b, rows, cols);
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    auto it = b->ctx->tuned.find(key);
    if (it != b->ctx->tuned.end()) var = it->second;
  }
  b->last_variant = var;
  const TilePlan plan = build_tiles(rows, cols, var.bands);
  XRH_REQUIRE(plan.band_start.size() <= L.band_words, "band plan exceeds its staging");
  std::memcpy(b->blob_host + L.off_band, plan.band_start.data(), sizeof(uint32_t) * plan.band_start.size());

  DrrLaunch D;
  std::memset(&D, 0, sizeof(D));
  D.poses_dev = (const PoseIn*)(b->blob_dev + L.off_pose);
  D.vols_dev = b->vols_dev;
  D.cams_dev = b->cams_dev;
  D.items_dev = b->items_dev;
  D.band_start_dev = (const uint32_t*)(b->blob_dev + L.off_band);
  D.num_items = n;
  D.bands = plan.bands;
  D.max_band_tiles = plan.max_band_tiles;
  D.geom = b->geom;
  D.kernel = b->views[0].kernel;
  D.store = XRH_STORE_REPLACE;
  D.fill = 0.0f;
  D.storage = b->storage;
  D.layout = b->layout;
  D.out = drr;
  D.ncc_partials = b->ncc_partials_dev;
  D.ncc_tiles_per_item = b->max_img_tiles;
  D.reuse = var.reuse;
  D.wave_shape = var.shape;
  D.objs = objs;
  return D;
}

void enqueue(xrh_batch* b, const DrrLaunch& D, const std::vector<std::pair<int, std::vector<int64_t>>>& groups,
             int64_t n, bool timed)
{
  hipStream_t s = b->ctx->stream;
  const BlobLayout L = blob_layout(b->max_items);
  const uint64_t up = L.off_sim + sizeof(SimItem) * n;
  XRH_HIP_CHECK(hipMemcpyAsync(b->blob_dev, b->blob_host, up, hipMemcpyHostToDevice, s));
  launch_drr_setup(D, s);
  if (timed) XRH_HIP_CHECK(hipEventRecord(b->ev_drr0, s));
  launch_drr(D, s);
  if (timed) XRH_HIP_CHECK(hipEventRecord(b->ev_drr1, s));
  const SimItem* sim_items = (const SimItem*)(b->blob_dev + L.off_sim);
  int64_t k = 0;
  for (auto& g : groups) {
    const int64_t cnt = (int64_t)g.second.size();
    // NCC sums were written by the DRR epilogue (records keyed by item index)
    const bool fused = ((g.first & ((1 << kSimKeyShift) - 1)) == XRH_SIM_NCC) && b->fuse_ncc;
    launch_sim(g.first, sim_items + k, cnt, b->max_img_tiles, b->max_patch_tiles,
               fused ? b->ncc_partials_dev : b->partials_dev, b->sims_dev, fused, b->max_lds, s);
    k += cnt;
  }
  XRH_HIP_CHECK(hipMemcpyAsync(b->sims_host, b->sims_dev, sizeof(float) * n, hipMemcpyDeviceToHost, s));
}

std::string signature(const xrh_batch* b, const int32_t* item_view, int64_t n, const float* drr_out,
                      int32_t objs, const float* bg)
{
  std::string key((const char*)item_view, (size_t)(n * objs) * sizeof(int32_t));
  key.append((const char*)&objs, sizeof(objs));
  key.append((const char*)&bg, sizeof(bg));
  key.append((const char*)&drr_out, sizeof(drr_out));
  key.append((const char*)&b->version, sizeof(b->version));
  key.append((const char*)&b->last_variant, sizeof(b->last_variant));
  return key;
}

void launch(xrh_batch* b, const int32_t* item_view, const double* item_pose, int64_t n, float* drr_out,
            int32_t objs = 1, const float* bg = nullptr)
{
  XRH_REQUIRE(b && item_view && item_pose, "null argument");
  XRH_REQUIRE(n >= 1 && n <= b->max_items, "item count must be in [1, max_items]");
  XRH_REQUIRE(!b->pending, "previous launch not yet waited for");
  hipStream_t s = b->ctx->stream;
  XRH_HIP_CHECK(hipStreamSynchronize(s));  // pinned staging is reused
  std::vector<std::pair<int, std::vector<int64_t>>> groups;
  const DrrLaunch D = stage(b, item_view, item_pose, n, drr_out, groups, objs, bg);
  const bool timed = b->time_kernels || !b->use_graph;
  b->timed_last = timed;
  if (timed) XRH_HIP_CHECK(hipEventRecord(b->ev_all0, s));
  if (b->use_graph && !b->time_kernels) {
    const std::string key = signature(b, item_view, n, drr_out, objs, bg);
    auto it = b->graphs.find(key);
    if (it == b->graphs.end()) {
      evict_graphs(b, kMaxCachedGraphs - 1);
      hipGraph_t g = nullptr;
      XRH_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        enqueue(b, D, groups, n, false);
      } catch (...) {
        hipGraph_t junk = nullptr;
        (void)hipStreamEndCapture(s, &junk);
        if (junk) (void)hipGraphDestroy(junk);
        throw;
      }
      XRH_HIP_CHECK(hipStreamEndCapture(s, &g));
      hipGraphExec_t ge = nullptr;
      const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      if (ie != hipSuccess) {
        (void)hipGraphDestroy(g);
        XRH_HIP_CHECK(ie);
      }
      xrh_batch::CachedGraph cg;
      cg.exec = ge;
      cg.graph = g;
      it = b->graphs.emplace(key, cg).first;
    }
    it->second.last_use = ++b->graph_clock;
    XRH_HIP_CHECK(hipGraphLaunch(it->second.exec, s));
  } else {
    enqueue(b, D, groups, n, timed);
  }
  if (timed) XRH_HIP_CHECK(hipEventRecord(b->ev_all1, s));
  b->last_n = n;
  b->pending = true;
}

void wait(xrh_batch* b, float* sim_host_out)
{
  XRH_REQUIRE(b, "null batch");
  XRH_REQUIRE(b->pending, "no launch to wait for");
  XRH_HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
  b->pending = false;
  if (sim_host_out) std::memcpy(sim_host_out, b->sims_host, sizeof(float) * b->last_n);
}

// Times the DRR launch of this submission under every variant (HIP events around the DRR
// kernel, eager submissions, the minimum over `reps` runs after one untimed run) and keeps
// the fastest for this batch and, in the context, for launches of the same key. The default
// is timed first and last; a variant replaces it only if it is 1% faster than the better of
// the two.
void autotune(xrh_batch* b, const int32_t* item_view, const double* item_pose, int64_t n, int32_t reps,
              xrh_tune_result* out)
{
  XRH_REQUIRE(b && item_view && item_pose && out, "null argument");
  XRH_REQUIRE(reps >= 1 && reps <= 50, "reps must be in [1, 50]");
  XRH_REQUIRE(!b->pending, "batch has a launch in flight");
  XRH_REQUIRE(n >= 1 && n <= b->max_items, "item count must be in [1, max_items]");
  const bool saved_graph = b->use_graph, saved_time = b->time_kernels, saved_set = b->variant_set;
  const DrrVariant saved_var = b->variant;
  auto restore = [&] {
    b->use_graph = saved_graph;
    b->time_kernels = saved_time;
    b->variant_set = saved_set;
    b->variant = saved_var;
  };
  b->use_graph = false;
  b->time_kernels = true;
  auto time_variant = [&](const DrrVariant& v) {
    b->variant = v;
    b->variant_set = true;
    float best = 0.0f;
    for (int r = 0; r <= reps; ++r) {
      launch(b, item_view, item_pose, n, nullptr);
      wait(b, nullptr);
      float ms = 0.0f;
      XRH_HIP_CHECK(hipEventElapsedTime(&ms, b->ev_drr0, b->ev_drr1));
      if (r == 1 || (r > 1 && ms < best)) best = ms;
    }
    return best;
  };
  DrrVariant best_v;
  float best_ms = 0.0f, def_ms = 0.0f;
  int tried = 0;
  std::vector<int32_t> rows, cols;
  try {
    def_ms = time_variant(DrrVariant());
    best_ms = def_ms;
    tried = 1;
    for (int reuse = 0; reuse <= 1; ++reuse) {
      for (int shape = 0; shape <= 2; ++shape) {
        for (int bands : {4, 8, 16}) {
          DrrVariant v;
          v.reuse = reuse;
          v.shape = shape;
          v.bands = bands;
          const float ms = time_variant(v);
          ++tried;
          if (ms < best_ms) {
            best_ms = ms;
            best_v = v;
          }
        }
      }
    }
    // the default once more (the first timing runs on colder caches and clocks); a variant
    // must beat the better of the two by 1%
    def_ms = std::min(def_ms, time_variant(DrrVariant()));
    ++tried;
    if (!(best_ms < 0.99f * def_ms)) {
      best_v = DrrVariant();
      best_ms = def_ms;
    }
    for (int64_t i = 0; i < n; ++i) {
      rows.push_back(b->views[item_view[i]].cam.num_det_rows);
      cols.push_back(b->views[item_view[i]].cam.num_det_cols);
    }
  } catch (...) {
    restore();
    throw;
  }
  restore();
  b->variant = best_v;
  b->variant_set = true;
  int64_t wgs = 0;
  for (size_t i = 0; i < rows.size(); ++i) wgs += (int64_t)((rows[i] + 15) / 16) * ((cols[i] + 15) / 16);
  {
    const std::string key = tune_key(b, rows, cols);
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    b->ctx->tuned[key] = best_v;
  }
  out->chosen.cell_reuse = best_v.reuse >= 0 ? best_v.reuse : (drr_default_reuse(wgs) ? 1 : 0);
  out->chosen.wave_shape = best_v.shape;
  out->chosen.bands = best_v.bands;
  out->chosen_ms = best_ms;
  out->default_ms = def_ms;
  out->variants_timed = tried;
}

// The sample-counting pass of the DRR kernel over one submission: totals of the reference
// algorithm's samples and of those evaluated after skipping, and optionally every ray's last
// sample index written to steps_dev (item order, rows * cols per item, like the DRRs).
void count_samples(xrh_batch* b, const int32_t* item_view, const double* item_pose, int64_t n,
                   uint64_t* samples_out, uint64_t* evaluated_out, int32_t* steps_dev)
{
  XRH_REQUIRE(b && item_view && item_pose, "null argument");
  XRH_REQUIRE(!b->pending, "batch has a launch in flight");
  XRH_REQUIRE(n >= 1 && n <= b->max_items, "item count must be in [1, max_items]");
  hipStream_t s = b->ctx->stream;
  XRH_HIP_CHECK(hipStreamSynchronize(s));
  std::vector<std::pair<int, std::vector<int64_t>>> groups;
  // staged as if writing DRRs to steps_dev, so every item's out_off indexes that buffer
  DrrLaunch D = stage(b, item_view, item_pose, n, (float*)steps_dev, groups);
  const BlobLayout L = blob_layout(b->max_items);
  XRH_HIP_CHECK(hipMemcpyAsync(b->blob_dev, b->blob_host, L.off_sim, hipMemcpyHostToDevice, s));
  uint64_t* cnt = (uint64_t*)ctx_scratch(b->ctx, 16);
  uint64_t both[2] = {0, 0};
  XRH_HIP_CHECK(hipMemcpyAsync(cnt, both, 16, hipMemcpyHostToDevice, s));
  D.sample_counter = cnt;
  D.ray_steps = steps_dev;
  launch_drr_setup(D, s);
  launch_drr(D, s);
  XRH_HIP_CHECK(hipMemcpyAsync(both, cnt, 16, hipMemcpyDeviceToHost, s));
  XRH_HIP_CHECK(hipStreamSynchronize(s));
  if (samples_out) *samples_out = both[0];
  if (evaluated_out) *evaluated_out = both[1];
}

}  // namespace

extern "C" {

xrh_status xrh_batch_create(xrh_ctx* ctx, const xrh_view* views, int32_t num_views, int64_t max_items,
                            int32_t geom, xrh_batch** out)
{
  return guarded([&] { batch_create(ctx, views, num_views, max_items, geom, out); });
}

xrh_status xrh_batch_create_empty(xrh_ctx* ctx, int32_t max_views, int64_t max_items, int32_t max_rows,
                                  int32_t max_cols, int32_t geom, xrh_batch** out)
{
  return guarded([&] { batch_alloc(ctx, max_views, max_items, max_rows, max_cols, geom, out); });
}

xrh_status xrh_batch_set_views(xrh_batch* batch, const xrh_view* views, int32_t num_views)
{
  return guarded([&] { set_views(batch, views, num_views); });
}

xrh_status xrh_batch_destroy(xrh_batch* batch)
{
  return guarded([&] { batch_free(batch); });
}

xrh_status xrh_batch_autotune(xrh_batch* batch, const int32_t* item_view, const double* item_pose, int64_t n_items,
                              int32_t reps, xrh_tune_result* out)
{
  return guarded([&] { autotune(batch, item_view, item_pose, n_items, reps, out); });
}

xrh_status xrh_batch_set_drr_variant(xrh_batch* batch, const xrh_drr_variant* v)
{
  return guarded([&] {
    XRH_REQUIRE(batch, "null batch");
    XRH_REQUIRE(!batch->pending, "batch has a launch in flight");
    if (!v) {
      batch->variant_set = false;
      batch->variant = DrrVariant();
      return;
    }
    XRH_REQUIRE(v->cell_reuse >= -1 && v->cell_reuse <= 1, "cell_reuse must be -1, 0 or 1");
    XRH_REQUIRE(v->wave_shape >= 0 && v->wave_shape <= 2, "wave_shape must be 0, 1 or 2");
    XRH_REQUIRE(v->bands >= 1 && v->bands <= kMaxDrrBands, "bands must be in [1, 16]");
    batch->variant.reuse = v->cell_reuse;
    batch->variant.shape = v->wave_shape;
    batch->variant.bands = v->bands;
    batch->variant_set = true;
  });
}

xrh_status xrh_batch_set_options(xrh_batch* batch, int32_t use_graph, int32_t time_kernels)
{
  return guarded([&] {
    XRH_REQUIRE(batch, "null batch");
    XRH_REQUIRE(!batch->pending, "batch has a launch in flight");
    batch->use_graph = use_graph != 0;
    batch->time_kernels = time_kernels != 0;
  });
}

xrh_status xrh_batch_eval(xrh_batch* batch, const int32_t* item_view, const double* item_pose, int64_t n,
                          float* sim_host_out, float* drr_dev_out)
{
  return guarded([&] {
    launch(batch, item_view, item_pose, n, drr_dev_out);
    wait(batch, sim_host_out);
  });
}

xrh_status xrh_batch_eval_objects(xrh_batch* batch, const int32_t* entry_view, const double* entry_pose,
                                  int64_t n_items, int32_t num_objs, const float* bg_dev, float* sim_host_out,
                                  float* drr_dev_out)
{
  return guarded([&] {
    XRH_REQUIRE(num_objs >= 1 && num_objs <= 64, "objects per item must be in [1, 64]");
    XRH_REQUIRE(n_items >= 1, "need at least one item");
    launch(batch, entry_view, entry_pose, n_items, drr_dev_out, num_objs, bg_dev);
    wait(batch, sim_host_out);
  });
}

xrh_status xrh_batch_launch(xrh_batch* batch, const int32_t* item_view, const double* item_pose, int64_t n)
{
  return guarded([&] { launch(batch, item_view, item_pose, n, nullptr); });
}

xrh_status xrh_batch_wait(xrh_batch* batch, float* sim_host_out)
{
  return guarded([&] { wait(batch, sim_host_out); });
}

xrh_status xrh_batch_last_timing(const xrh_batch* batch, float* drr_kernel_ms, float* total_ms)
{
  return guarded([&] {
    XRH_REQUIRE(batch, "null batch");
    XRH_REQUIRE(!batch->pending, "launch still in flight");
    XRH_REQUIRE(batch->timed_last, "last launch was not timed (enable time_kernels)");
    float a = 0.0f, t = 0.0f;
    if (drr_kernel_ms) {
      XRH_HIP_CHECK(hipEventElapsedTime(&a, batch->ev_drr0, batch->ev_drr1));
      *drr_kernel_ms = a;
    }
    if (total_ms) {
      XRH_HIP_CHECK(hipEventElapsedTime(&t, batch->ev_all0, batch->ev_all1));
      *total_ms = t;
    }
  });
}

xrh_status xrh_batch_sample_count(xrh_batch* batch, const int32_t* item_view, const double* item_pose,
                                  int64_t n, uint64_t* samples_out, uint64_t* evaluated_out)
{
  return guarded([&] {
    XRH_REQUIRE(samples_out, "null argument");
    count_samples(batch, item_view, item_pose, n, samples_out, evaluated_out, nullptr);
  });
}

xrh_status xrh_batch_ray_steps(xrh_batch* batch, const int32_t* item_view, const double* item_pose, int64_t n,
                               int32_t* steps_dev_out)
{
  return guarded([&] {
    XRH_REQUIRE(steps_dev_out, "null argument");
    count_samples(batch, item_view, item_pose, n, nullptr, nullptr, steps_dev_out);
  });
}

}  // extern "C"
