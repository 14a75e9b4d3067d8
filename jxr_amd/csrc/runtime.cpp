// runtime.cpp — host runtime of the gfx950 reconstruction backend: plans,
// sessions, uploads, submissions and batches behind the C ABI of
// include/jxr_hip.h.
//
// Reference behaviour kept (crates/jxr-cuda/src):
//   plan validation before any device resource exists   plan.rs:91-354
//   immutable coefficient uploads cached by identity      upload_cache.rs:15-153
//   exact-size scratch pool                               buffer_pool.rs:9-129
//   status word per submission, first failing stage       submission.rs:141-154
//   several streams, independent images round-robin       session.rs:112-123
//   route threshold of 16,384 coefficients                route.rs:8-50
// New here: one persistent launch per shape bucket over a whole batch, and
// 64-bit sizes throughout.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/jxr_hip.h"
#include "recon_abi.h"

namespace jxr {
hipError_t launch_tiles(const KernelArgs &args, uint32_t grid, hipStream_t stream);
hipError_t launch_clear_rows(uint8_t *base, uint64_t pitch, uint32_t row_bytes, uint32_t rows, hipStream_t stream);
hipError_t launch_expand(const Slot *slots, uint32_t slot_count, uint32_t item_count, uint32_t *item_slot,
                         uint32_t *counters, hipStream_t stream);
hipError_t kernel_occupancy(uint32_t storage, uint32_t reduced, size_t dynamic_lds, int *groups_per_cu);
hipError_t launch_scale_terms(const int32_t *pictures, uint32_t length, uint32_t levels, int64_t span,
                              int64_t *terms, uint32_t *status, hipStream_t stream);
void flat_probe_coefficients(int32_t level, uint32_t bc, uint32_t br, bool doubled, uint32_t overlap,
                             const std::vector<uint32_t> &cols, const std::vector<uint32_t> &rows,
                             std::vector<int32_t> &out);
hipError_t set_kernel_lds_limit(size_t bytes);
hipError_t launch_valu_chains(uint32_t *out, uint32_t groups, uint32_t iterations, hipStream_t stream);
uint32_t valu_per_iteration();
hipError_t static_scratch_address(int device, uint8_t **out);
}  // namespace jxr

using namespace jxr;

// ---- errors -----------------------------------------------------------------
static thread_local std::string g_error;
extern "C" void jxrhip_set_error(const char *message) { g_error = message ? message : ""; }
extern "C" const char *jxrhip_last_error(void) { return g_error.c_str(); }
extern "C" uint32_t jxrhip_abi_version(void) { return JXRHIP_ABI_VERSION; }

static jxrhip_status fail(jxrhip_status code, const std::string &message) {
  g_error = message;
  return code;
}
static jxrhip_status hip_fail(hipError_t e, const char *what) {
  g_error = std::string(what) + ": " + hipGetErrorString(e);
  return JXRHIP_ERR_DEVICE;
}
#define HIP_TRY(expr, what)                   \
  do {                                        \
    hipError_t e_ = (expr);                   \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

// ---- routing (route.rs:8-50) ------------------------------------------------
static constexpr uint64_t kAutoThreshold = 16384;
extern "C" uint64_t jxrhip_route_threshold(void) { return kAutoThreshold; }
extern "C" jxrhip_status jxrhip_route(uint64_t coefficients, int32_t strict, int32_t available,
                                      int32_t *use_device) {
  if (!use_device) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null output");
  if (strict) {
    if (!available) return fail(JXRHIP_ERR_UNAVAILABLE, "strict device route without a usable device");
    *use_device = 1;
    return JXRHIP_OK;
  }
  *use_device = available && coefficients >= kAutoThreshold ? 1 : 0;
  return JXRHIP_OK;
}

extern "C" int32_t jxrhip_is_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return 0;
  }
  for (int d = 0; d < n; d++) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d) == hipSuccess && std::strncmp(prop.gcnArchName, "gfx950", 6) == 0)
      return 1;
  }
  return 0;
}

// ---- plans ---------------------------------------------------------------
struct PlanPlane {
  uint32_t arena;  // 0 primary, 1 separate alpha
  uint32_t bc, br;
  bool scale_first = false, alpha = false;
  uint32_t overlap = 0;
  std::vector<MbRec> recs;                 // raster order
  std::vector<uint32_t> colmask, rowmask;  // empty for soft tiles
  std::vector<uint32_t> regions_x, regions_y;  // region boundaries (macroblocks), 0 .. extent
  // reduced decode: samples across / down a macroblock, the sample scale
  // (ReduceMode, shift) and, under overlap, the probe set whose terms it
  // reads and, for each reduced column and row, the probe sample placed
  // like it (sample_scale.rs:405-417)
  uint32_t ux = 0, uy = 0;
  uint32_t reduce_mode = kReduceShift, reduce_shift = 0;
  int32_t probe = -1;
  std::vector<uint32_t> probe_cols, probe_rows;
};

// The flat probe images of one plane configuration's sample scale
// (sample_scale.rs:157-246): the levels +-0x10001 .. +-0x10004 analysed at
// unit steps over a probe of one region for each region extent of the plane
// (five or more folding onto five), hard tiles at the probe's regions.
constexpr uint32_t kProbeLevels = 4;
constexpr uint32_t kProbeExtent = 5;
struct ProbeSet {
  uint32_t arena, bc, br, overlap;
  bool scale_first, alpha;
  uint32_t mbx, mby;                     // probe macroblocks
  std::vector<uint32_t> bounds_x, bounds_y;  // probe regions (macroblocks), 0 .. extent
  std::vector<uint32_t> colmask, rowmask;
  std::vector<MbRec> recs;               // shared by every probe image
  uint32_t stride, length;               // probe picture: samples across, samples
  int64_t span, gain;
  // what the terms depend on: block shape, doubling, overlap, decode scale
  // and the probe's regions; a session computes them once per key
  std::vector<uint32_t> key;
};

static std::atomic<uint64_t> g_plan_ids{1};

struct jxrhip_plan {
  uint64_t id = 0;
  const jxrhip_arena *arenas[2] = {nullptr, nullptr};
  uint32_t arena_count = 0;
  std::vector<PlanPlane> planes;
  DevImage dev{};  // the whole image's output policy; device pointers filled per session at upload
  // Plane groups: one descriptor per group of planes a work item holds in
  // LDS at once. Every image is one group, except N-component images of
  // more than four colour planes (see jxrhip_plan_create).
  std::vector<DevImage> groups;
  std::vector<std::vector<uint32_t>> group_planes;  // plan plane of each group plane
  uint32_t mbx = 0, mby = 0;
  uint64_t output_bytes = 0, recon_coefs = 0, pixels = 0, alg_bytes = 0;
  uint64_t content_bytes = 0;  // bytes of the surface the output stage writes
  uint64_t descriptor_bytes = 0;
  uint64_t coefficient_bytes = 0;
  uint64_t geometry_key = 0;  // bucket key: plane geometries
  jxrhip_surface_layout surface{};  // the validated output layout
  uint32_t scale = JXRHIP_SCALE_FULL;
  std::vector<ProbeSet> probes;     // reduced decode under overlap
};

static int32_t bias_of(uint32_t depth, uint32_t shift_bits, bool *bad) {
  int32_t base = 0;
  uint32_t shift = 0;
  switch (depth) {
    case JXRHIP_DEPTH_RGB555: base = 16; break;
    case JXRHIP_DEPTH_RGB565: base = 32; break;
    case JXRHIP_DEPTH_U8: base = 128; break;
    case JXRHIP_DEPTH_U10:
    case JXRHIP_DEPTH_RGB101010: base = 512; break;
    case JXRHIP_DEPTH_U16: base = 32768; shift = shift_bits; break;
    case JXRHIP_DEPTH_I16:
    case JXRHIP_DEPTH_I32: shift = shift_bits; break;
    default: break;
  }
  if (shift >= 32) {
    *bad = true;
    return 0;
  }
  return base >> shift;
}

// ChannelScale::resolve (crates/jxr-native/src/output_format/scaling.rs:42-97)
static ChanScale make_scale(uint32_t component, uint32_t output_color, uint32_t depth,
                            uint32_t shift_bits, uint32_t scaled) {
  ChanScale s{};
  bool bad = false;
  int32_t bias = bias_of(depth, shift_bits, &bad);
  if (bad) {
    s.flags = 1;
    return s;
  }
  int32_t cb = bias;
  if (output_color == JXRHIP_COLOR_CMYK) cb = component < 3 ? (bias >> 1) : -(bias >> 1);
  else if (output_color == JXRHIP_COLOR_RGBE) cb = 0;
  uint32_t scale = scaled ? 3 : 0;
  int64_t sh = (int64_t)cb * ((int64_t)1 << scale);
  if (sh > INT32_MAX || sh < INT32_MIN) {
    s.flags = 1;
    return s;
  }
  s.shifted_bias = (int32_t)sh;
  s.rounding = scaled ? ((depth == JXRHIP_DEPTH_BIT1_WHITE || depth == JXRHIP_DEPTH_BIT1_BLACK ||
                          depth == JXRHIP_DEPTH_U16)
                             ? 4
                             : 3)
                      : 0;
  s.shift = (depth == JXRHIP_DEPTH_RGB565 && component != 1) ? scale + 1 : scale;
  if (depth == JXRHIP_DEPTH_U16 || depth == JXRHIP_DEPTH_I16 || depth == JXRHIP_DEPTH_I32) {
    s.flags |= 4;
    s.post_shift = shift_bits;
  }
  return s;
}

static bool layout_has_alpha(uint32_t l) {
  return l == JXRHIP_LAYOUT_LUMA_ALPHA || l == JXRHIP_LAYOUT_YUVA || l == JXRHIP_LAYOUT_RGBA ||
         l == JXRHIP_LAYOUT_BGRA || l == JXRHIP_LAYOUT_CMYKA || l == JXRHIP_LAYOUT_NCOMPONENT_ALPHA;
}

// Clipped tile boundaries of one axis (crates/jxr-core/src/device_plan/
// overlap.rs:102-123) as a bitmask over macroblock positions.
static bool tile_mask(const uint32_t *sizes, uint32_t n, uint32_t origin, uint32_t extent,
                      std::vector<uint32_t> &mask) {
  mask.assign(extent / 32 + 2, 0u);
  uint64_t end = (uint64_t)origin + extent, pos = 0;
  for (uint32_t i = 0; i < n; i++) {
    pos += sizes[i];
    if (pos > origin && pos < end) {
      uint64_t b = pos - origin;
      mask[b >> 5] |= 1u << (b & 31);
    }
  }
  return pos >= end;
}

static jxrhip_status build_plane(jxrhip_plan *plan, const jxrhip_arena *a, uint32_t arena_index,
                                 uint32_t index, bool scale, bool alpha, uint32_t *ox, uint32_t *oy) {
  if (index >= a->plane_count) return fail(JXRHIP_ERR_INVALID_PLAN, "coefficient plane index is absent");
  const jxrhip_coefficient_plane &cp = a->planes[index];
  PlanPlane pp;
  pp.arena = arena_index;
  pp.bc = cp.block_columns;
  pp.br = cp.block_rows;
  pp.scale_first = scale;
  pp.alpha = alpha;
  pp.overlap = a->overlap;
  if (!((pp.bc == 4 && pp.br == 4) || (pp.bc == 2 && pp.br == 4) || (pp.bc == 2 && pp.br == 2)))
    return fail(JXRHIP_ERR_INVALID_PLAN, "unsupported component macroblock geometry");
  uint64_t mo = cp.macroblock_offset, mc = cp.macroblock_count;
  if (mc == 0 || mo + mc > a->macroblock_count || mo + mc < mo)
    return fail(JXRHIP_ERR_INVALID_PLAN, "macroblock plane range");
  // checked addition, as check_range does (jxr-core plan.rs): a huge offset
  // must not wrap into range
  const uint64_t co = cp.coefficient_offset, cc = cp.coefficient_count;
  if (co + cc < co || co + cc > a->coefficient_count)
    return fail(JXRHIP_ERR_INVALID_PLAN, "coefficient plane range");
  uint32_t x0 = UINT32_MAX, y0 = UINT32_MAX, x1 = 0, y1 = 0;
  for (uint64_t i = mo; i < mo + mc; i++) {
    x0 = std::min(x0, a->coded_x[i]);
    y0 = std::min(y0, a->coded_y[i]);
    x1 = std::max(x1, a->coded_x[i]);
    y1 = std::max(y1, a->coded_y[i]);
  }
  const uint32_t mbx = x1 - x0 + 1, mby = y1 - y0 + 1;
  if ((uint64_t)mbx * mby != mc)
    return fail(JXRHIP_ERR_INVALID_PLAN, "coefficient plane reconstruction window is not rectangular");
  if (plan->mbx == 0) {
    plan->mbx = mbx;
    plan->mby = mby;
    *ox = x0;
    *oy = y0;
  } else if (plan->mbx != mbx || plan->mby != mby || *ox != x0 || *oy != y0) {
    return fail(JXRHIP_ERR_INVALID_PLAN, "planes do not share one macroblock window");
  }
  if ((uint64_t)mbx * 16 > INT32_MAX / 2 || (uint64_t)mby * 16 > INT32_MAX / 2)
    return fail(JXRHIP_ERR_INVALID_PLAN, "sample coordinates exceed signed indexing");
  const uint32_t blocks = pp.bc * pp.br;
  pp.recs.assign(mc, MbRec{UINT32_MAX, 0, 0, 0, 0, 0, 0, 0});
  std::vector<uint8_t> seen(mc, 0);
  for (uint64_t i = mo; i < mo + mc; i++) {
    uint64_t r = (uint64_t)(a->coded_y[i] - y0) * mbx + (a->coded_x[i] - x0);
    if (seen[r]) return fail(JXRHIP_ERR_INVALID_PLAN, "coefficient plane repeats a macroblock coordinate");
    seen[r] = 1;
    uint32_t bands = a->bands[i];
    if (bands > 3) return fail(JXRHIP_ERR_INVALID_PLAN, "band presence code");
    uint64_t need = bands == 0 ? 1 : bands == 1 ? blocks : (uint64_t)blocks * 16;
    if ((uint64_t)a->coefficient_offsets[i] + need > a->coefficient_count)
      return fail(JXRHIP_ERR_INVALID_PLAN, "macroblock coefficient span exceeds the uploaded arena");
    uint32_t pred = a->hp_predictions[i];
    if (pred > JXRHIP_PRED_FROM_TOP) return fail(JXRHIP_ERR_INVALID_PLAN, "top-left high-pass prediction is invalid");
    const uint32_t *q = a->quantizers + 3 * i;
    if (q[0] == 0 || q[1] == 0 || q[2] == 0) return fail(JXRHIP_ERR_INVALID_PLAN, "zero quantizer step");
    MbRec &rec = pp.recs[r];
    rec.coef_off = a->coefficient_offsets[i];
    rec.qdc = q[0];
    rec.qlp = q[1];
    rec.qhp = q[2];
    rec.bands = bands;
    rec.pred = bands >= JXRHIP_BANDS_NO_FLEXBITS ? pred : JXRHIP_PRED_NONE;
  }
  if (a->hard_tiles) {
    if (!tile_mask(a->tile_column_widths, a->tile_column_count, x0, mbx, pp.colmask) ||
        !tile_mask(a->tile_row_heights, a->tile_row_count, y0, mby, pp.rowmask))
      return fail(JXRHIP_ERR_INVALID_PLAN, "tile partition does not cover reconstruction window");
  }
  // regions no overlap operator crosses: the window, split at hard tile
  // edges (device_plan/overlap.rs:102-123)
  pp.regions_x.push_back(0);
  for (uint32_t m = 1; m < mbx && !pp.colmask.empty(); m++)
    if ((pp.colmask[m >> 5] >> (m & 31)) & 1u) pp.regions_x.push_back(m);
  pp.regions_x.push_back(mbx);
  pp.regions_y.push_back(0);
  for (uint32_t m = 1; m < mby && !pp.rowmask.empty(); m++)
    if ((pp.rowmask[m >> 5] >> (m & 31)) & 1u) pp.regions_y.push_back(m);
  pp.regions_y.push_back(mby);
  if (a->overlap > 2) return fail(JXRHIP_ERR_INVALID_PLAN, "overlap mode");
  plan->planes.push_back(std::move(pp));
  return JXRHIP_OK;
}

// ProbeAxis (sample_scale.rs:336-381): the probe's region boundaries along
// one axis, and the probe macroblock placed in its region as each
// macroblock of the plane is placed in its own (probe_offset,
// sample_scale.rs:405-417).
static void probe_axis(const std::vector<uint32_t> &regions, std::vector<uint32_t> &bounds,
                       std::vector<uint32_t> &pos) {
  bounds.assign(1, 0);
  pos.clear();
  int64_t starts[kProbeExtent + 1];
  std::fill(starts, starts + kProbeExtent + 1, int64_t(-1));
  uint32_t mbs = 0;
  const uint32_t ends = kProbeExtent / 2;
  for (size_t i = 0; i + 1 < regions.size(); i++) {
    const uint32_t extent = regions[i + 1] - regions[i];  // positive: regions advance
    const uint32_t pe = std::min(extent, kProbeExtent);
    if (starts[pe] < 0) {
      starts[pe] = mbs;
      mbs += pe;
      bounds.push_back(mbs);
    }
    for (uint32_t off = 0; off < extent; off++) {
      const uint32_t o = (extent <= kProbeExtent || off < ends) ? off
                         : extent - off <= ends                 ? kProbeExtent - (extent - off)
                                                                : ends;
      pos.push_back((uint32_t)starts[pe] + o);
    }
  }
}

static void probe_mask(const std::vector<uint32_t> &bounds, std::vector<uint32_t> &mask) {
  mask.assign(bounds.back() / 32 + 2, 0u);
  for (size_t i = 1; i + 1 < bounds.size(); i++) mask[bounds[i] >> 5] |= 1u << (bounds[i] & 31);
}

// Sample scale of every plane of a reduced decode (pipeline.rs:84-112,
// sample_scale.rs:157-246): without overlap the gain of the transform
// stages left out (two bits for the block transform, and at a sixteenth two
// more for a full macroblock's first level, one for subsampled chroma); with
// it, the flat probe images of the plane's configuration, whose reduced
// pictures the device turns into the scale's terms.
static jxrhip_status build_reduction(jxrhip_plan *plan) {
  const uint32_t scale = plan->scale;
  for (PlanPlane &pp : plan->planes) {
    pp.ux = reduced_units(pp.bc, scale);
    pp.uy = reduced_units(pp.br, scale);
    const uint32_t mb_gain = pp.bc == 4 && pp.br == 4 ? 2 : 1;
    pp.reduce_shift = scale == JXRHIP_SCALE_QUARTER ? 2 : mb_gain + 2;
    if (pp.overlap == 0) {
      pp.reduce_mode = kReduceShift;
      continue;
    }
    pp.reduce_mode = kReduceFlat;
    std::vector<uint32_t> bx, px, by, py;
    probe_axis(pp.regions_x, bx, px);
    probe_axis(pp.regions_y, by, py);
    // planes of one arena share its regions: one probe set serves every
    // plane of the same configuration
    int32_t found = -1;
    for (size_t k = 0; k < plan->probes.size(); k++) {
      const ProbeSet &q = plan->probes[k];
      if (q.arena == pp.arena && q.bc == pp.bc && q.br == pp.br && q.overlap == pp.overlap &&
          q.scale_first == pp.scale_first)
        found = (int32_t)k;
    }
    if (found < 0) {
      ProbeSet q{};
      q.arena = pp.arena;
      q.bc = pp.bc;
      q.br = pp.br;
      q.overlap = pp.overlap;
      q.scale_first = pp.scale_first;
      q.alpha = pp.alpha;
      q.mbx = bx.back();
      q.mby = by.back();
      probe_mask(bx, q.colmask);
      probe_mask(by, q.rowmask);
      q.bounds_x = bx;
      q.bounds_y = by;
      const uint32_t blocks = pp.bc * pp.br;
      q.recs.resize((size_t)q.mbx * q.mby);
      for (uint32_t m = 0; m < q.mbx * q.mby; m++)
        q.recs[m] = MbRec{m * blocks, 1, 1, 1, JXRHIP_BANDS_NO_HIGH_PASS, JXRHIP_PRED_NONE, 0, 0};
      q.span = 0;
      for (uint32_t k = 0; k < kProbeLevels; k++) q.span += 2 * (int64_t)(0x10001 + (int32_t)k);
      q.key = {pp.bc, pp.br, pp.scale_first ? 1u : 0u, pp.overlap, scale, (uint32_t)bx.size()};
      q.key.insert(q.key.end(), bx.begin(), bx.end());
      q.key.insert(q.key.end(), by.begin(), by.end());
      q.gain = 2 * (int64_t)kProbeLevels * q.span;
      q.stride = q.mbx * pp.ux;
      q.length = q.stride * q.mby * pp.uy;
      plan->probes.push_back(std::move(q));
      found = (int32_t)plan->probes.size() - 1;
    }
    pp.probe = found;
    pp.probe_cols.resize((size_t)plan->mbx * pp.ux);
    for (uint32_t x = 0; x < plan->mbx * pp.ux; x++) pp.probe_cols[x] = px[x / pp.ux] * pp.ux + x % pp.ux;
    pp.probe_rows.resize((size_t)plan->mby * pp.uy);
    for (uint32_t y = 0; y < plan->mby * pp.uy; y++) pp.probe_rows[y] = py[y / pp.uy] * pp.uy + y % pp.uy;
  }
  return JXRHIP_OK;
}

static jxrhip_status validate_arena(const jxrhip_arena *a) {
  if (!a || !a->coefficients || !a->coefficient_offsets || !a->quantizers || !a->bands ||
      !a->hp_predictions || !a->coded_x || !a->coded_y || !a->planes)
    return fail(JXRHIP_ERR_INVALID_PLAN, "coefficient arena has a null array");
  if (a->coefficient_count > UINT32_MAX)
    return fail(JXRHIP_ERR_INVALID_PLAN, "coefficient arena exceeds 2^32 coefficients");
  if (a->hard_tiles && (!a->tile_column_widths || !a->tile_row_heights))
    return fail(JXRHIP_ERR_INVALID_PLAN, "hard tiles without a tile grid");
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_plan_create(const jxrhip_image_desc *d, jxrhip_plan **out) {
  if (!d || !out) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (d->abi_version != JXRHIP_ABI_VERSION) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "ABI version mismatch");
  std::unique_ptr<jxrhip_plan> plan(new (std::nothrow) jxrhip_plan());
  if (!plan) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "out of host memory");
  plan->id = g_plan_ids.fetch_add(1);
  jxrhip_status st = validate_arena(d->primary);
  if (st) return st;
  const uint32_t cf = d->color_format;
  const uint32_t colors = d->component_count;
  const uint32_t expect = cf == JXRHIP_COLOR_LUMA ? 1
                          : (cf == JXRHIP_COLOR_YUV420 || cf == JXRHIP_COLOR_YUV422 ||
                             cf == JXRHIP_COLOR_YUV444 || cf == JXRHIP_COLOR_RGB || cf == JXRHIP_COLOR_RGBE)
                              ? 3
                          : (cf == JXRHIP_COLOR_CMYK || cf == JXRHIP_COLOR_CMYK_DIRECT || cf == JXRHIP_COLOR_YUVK)
                              ? 4
                              : colors;
  if (colors == 0 || colors > kMaxColors || colors != expect)
    return fail(JXRHIP_ERR_INVALID_PLAN, "primary component count does not match the colour format");
  const bool integrated = d->alpha_mode == JXRHIP_ALPHA_INTEGRATED;
  if (d->primary->plane_count != colors + (integrated ? 1 : 0))
    return fail(JXRHIP_ERR_INVALID_PLAN, "primary coefficient arena component count does not match the image");
  plan->arenas[0] = d->primary;
  plan->arena_count = 1;
  uint32_t ox = 0, oy = 0;
  for (uint32_t i = 0; i < colors + (integrated ? 1 : 0); i++) {
    const bool alpha = integrated && i == colors;
    st = build_plane(plan.get(), d->primary, 0, i, !alpha && d->scaled && i != 0, alpha, &ox, &oy);
    if (st) return st;
  }
  int32_t alpha_plane = integrated ? (int32_t)colors : -1;
  if (d->alpha_mode == JXRHIP_ALPHA_SEPARATE) {
    st = validate_arena(d->separate_alpha);
    if (st) return st;
    if (d->separate_alpha->plane_count != 1)
      return fail(JXRHIP_ERR_INVALID_PLAN, "separate alpha arena is not one independent luma plane");
    plan->arenas[1] = d->separate_alpha;
    plan->arena_count = 2;
    st = build_plane(plan.get(), d->separate_alpha, 1, 0, false, true, &ox, &oy);
    if (st) return st;
    alpha_plane = (int32_t)plan->planes.size() - 1;
  } else if (d->alpha_mode > JXRHIP_ALPHA_SEPARATE) {
    return fail(JXRHIP_ERR_INVALID_PLAN, "alpha mode");
  }
  for (uint32_t i = 0; i < colors; i++) {
    const PlanPlane &p = plan->planes[i];
    const bool chroma = i == 1 || i == 2;
    const bool ok = (cf == JXRHIP_COLOR_YUV420 && chroma)   ? (p.bc == 2 && p.br == 2)
                    : (cf == JXRHIP_COLOR_YUV422 && chroma) ? (p.bc == 2 && p.br == 4)
                                                           : (p.bc == 4 && p.br == 4);
    if (!ok) return fail(JXRHIP_ERR_INVALID_PLAN, "component macroblock geometry does not match the colour format");
  }
  // ---- decode scale ----
  if (d->scale > JXRHIP_SCALE_SIXTEENTH) return fail(JXRHIP_ERR_INVALID_PLAN, "decode scale");
  plan->scale = d->scale;
  if (plan->scale != JXRHIP_SCALE_FULL) {
    st = build_reduction(plan.get());
    if (st) return st;
  }
  const uint32_t mb_px = plan->scale == JXRHIP_SCALE_QUARTER ? 4u : plan->scale == JXRHIP_SCALE_SIXTEENTH ? 1u : 16u;

  // ---- output policy (device_plan/output.rs:102-295) ----
  const jxrhip_output_format &f = d->output;
  DevImage &dv = plan->dev;
  std::memset(&dv, 0, sizeof dv);
  dv.plane_count = (uint32_t)plan->planes.size();
  dv.colors = colors;
  dv.alpha_plane = alpha_plane;
  dv.mbx = plan->mbx;
  dv.mby = plan->mby;
  // a sixteenth has one chroma sample per macroblock, as luma does: nothing
  // to interpolate (metadata.rs:253-263, decode/output.rs:464-482)
  dv.upsample = plan->scale == JXRHIP_SCALE_SIXTEENTH ? 0
                : cf == JXRHIP_COLOR_YUV420           ? 1
                : cf == JXRHIP_COLOR_YUV422           ? 2
                                                      : 0;
  dv.scale = plan->scale;
  dv.mb_px = mb_px;
  dv.raw = 0;
  const uint32_t l = f.channel_layout;
  uint32_t primary = 0;
  switch (l) {
    case JXRHIP_LAYOUT_LUMA:
    case JXRHIP_LAYOUT_LUMA_ALPHA: primary = 1; break;
    case JXRHIP_LAYOUT_RGB:
    case JXRHIP_LAYOUT_BGR:
    case JXRHIP_LAYOUT_RGBA:
    case JXRHIP_LAYOUT_BGRA: primary = 3; break;
    case JXRHIP_LAYOUT_RGBX:
    case JXRHIP_LAYOUT_BGRX:
    case JXRHIP_LAYOUT_CMYK:
    case JXRHIP_LAYOUT_CMYKA: primary = 4; break;
    case JXRHIP_LAYOUT_NCOMPONENT:
    case JXRHIP_LAYOUT_NCOMPONENT_ALPHA: primary = colors; break;
    case JXRHIP_LAYOUT_YUV:
    case JXRHIP_LAYOUT_YUVA: primary = 3; break;
    default: return fail(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "channel layout not supported");
  }
  dv.primary_count = primary;
  dv.has_alpha = layout_has_alpha(l) ? 1 : 0;
  dv.reverse = (l == JXRHIP_LAYOUT_BGR || l == JXRHIP_LAYOUT_BGRX || l == JXRHIP_LAYOUT_BGRA) ? 1 : 0;
  dv.channels = primary + dv.has_alpha;
  if (dv.channels != f.channel_count) return fail(JXRHIP_ERR_INVALID_PLAN, "channel count differs from the layout");
  if (dv.has_alpha && alpha_plane < 0)
    return fail(JXRHIP_ERR_INVALID_PLAN, "alpha output layout has no prepared alpha plane");
  const uint32_t dep = f.bit_depth;
  const uint32_t adep = f.has_alpha_format ? f.alpha_bit_depth : dep;
  const uint32_t oc = f.output_color;
  // storage, colour, layout and depth combinations of the packed, one-bit
  // and planar YUV outputs (output_format/validate.rs:137-255)
  dv.out_kind = kOutTyped;
  dv.out_depth = dep;
  if (f.storage == JXRHIP_STORE_BITS) {
    if (dep != JXRHIP_DEPTH_BIT1_WHITE && dep != JXRHIP_DEPTH_BIT1_BLACK)
      return fail(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "bit depth does not match the storage kind");
    if (l != JXRHIP_LAYOUT_LUMA || oc != JXRHIP_COLOR_LUMA)
      return fail(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "one-bit output stores luma only");
    dv.out_kind = kOutBits;
  } else if (f.storage == JXRHIP_STORE_PACKED16 || f.storage == JXRHIP_STORE_PACKED32) {
    if (l != JXRHIP_LAYOUT_RGB) return fail(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "packed output takes the RGB layout");
    if (f.storage == JXRHIP_STORE_PACKED16 && (dep == JXRHIP_DEPTH_RGB555 || dep == JXRHIP_DEPTH_RGB565) &&
        oc == JXRHIP_COLOR_RGB)
      dv.out_kind = kOutPacked16;
    else if (f.storage == JXRHIP_STORE_PACKED32 && dep == JXRHIP_DEPTH_RGB101010 && oc == JXRHIP_COLOR_RGB)
      dv.out_kind = kOutRgb101010;
    else if (f.storage == JXRHIP_STORE_PACKED32 && dep == JXRHIP_DEPTH_U8 && oc == JXRHIP_COLOR_RGBE)
      dv.out_kind = kOutRgbe;
    else
      return fail(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "bit depth does not match the storage kind");
  } else if (l == JXRHIP_LAYOUT_YUV || l == JXRHIP_LAYOUT_YUVA) {
    if (oc != cf || (oc != JXRHIP_COLOR_YUV420 && oc != JXRHIP_COLOR_YUV422) || f.internal_color != oc)
      return fail(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "planar YUV output keeps the coded 4:2:0 or 4:2:2 sampling");
    if (f.premultiply_alpha) return fail(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "premultiplied YUV output");
    if (!d->surface.planar) return fail(JXRHIP_ERR_INVALID_PLAN, "YUV output is planar");
    if (plan->scale == JXRHIP_SCALE_SIXTEENTH)  // prepare.rs:185-192
      return fail(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "sixteenth-resolution subsampled planar YUV output");
    if (f.storage != JXRHIP_STORE_U8 && f.storage != JXRHIP_STORE_U16 && f.storage != JXRHIP_STORE_I16)
      return fail(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "T.832 output colour and bit depth matrix");
    dv.out_kind = kOutYuv;
    dv.chroma_sub_y = oc == JXRHIP_COLOR_YUV420 ? 2 : 1;
    dv.upsample = 0;  // chroma is stored at its coded resolution
  } else if (dep == JXRHIP_DEPTH_BIT1_WHITE || dep == JXRHIP_DEPTH_BIT1_BLACK || dep == JXRHIP_DEPTH_RGB555 ||
             dep == JXRHIP_DEPTH_RGB565 || dep == JXRHIP_DEPTH_RGB101010) {
    return fail(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "bit depth does not match the storage kind");
  }
  // packed RGB words of an image whose red and blue are swapped store them
  // swapped back (output_format/color.rs:24-26); the packed layout is RGB,
  // so `reverse` is free to carry it
  if ((dv.out_kind == kOutPacked16 || dv.out_kind == kOutRgb101010) && !f.red_blue_not_swapped) dv.reverse = 1;
  if (dv.out_kind == kOutBits) {
    dv.bytes = 1;
  } else if (dv.out_kind == kOutPacked16 || dv.out_kind == kOutRgb101010 || dv.out_kind == kOutRgbe) {
    dv.bytes = f.storage == JXRHIP_STORE_PACKED16 ? 2 : 4;
  } else {
    bool depth_ok = true;
    switch (f.storage) {
      case JXRHIP_STORE_U8: dv.bytes = 1; depth_ok = dep == JXRHIP_DEPTH_U8; break;
      case JXRHIP_STORE_U16: dv.bytes = 2; depth_ok = dep == JXRHIP_DEPTH_U16 || dep == JXRHIP_DEPTH_U10; break;
      case JXRHIP_STORE_I16: dv.bytes = 2; depth_ok = dep == JXRHIP_DEPTH_I16; break;
      case JXRHIP_STORE_I32: dv.bytes = 4; depth_ok = dep == JXRHIP_DEPTH_I32; break;
      case JXRHIP_STORE_F16: dv.bytes = 2; depth_ok = dep == JXRHIP_DEPTH_F16; break;
      case JXRHIP_STORE_F32:
        dv.bytes = 4;
        depth_ok = dep == JXRHIP_DEPTH_F32 && (!dv.has_alpha || adep == JXRHIP_DEPTH_F32);
        if (depth_ok && (f.mantissa_length > 23 || (dv.has_alpha && f.has_alpha_format && f.alpha_mantissa_length > 23)))
          return fail(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "BD32F mantissa length exceeds 23 bits");
        break;
      default: return fail(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "storage kind not supported");
    }
    if (!depth_ok) return fail(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "bit depth does not match the storage kind");
  }
  dv.storage = f.storage;
  dv.u16_max = dep == JXRHIP_DEPTH_U10 ? 1023 : 65535;
  {
    const bool yuv = cf == JXRHIP_COLOR_YUV420 || cf == JXRHIP_COLOR_YUV422 || cf == JXRHIP_COLOR_YUV444;
    if (dv.out_kind == kOutYuv) dv.conv = kConvDirect;  // planes stored as they are
    else if (oc == JXRHIP_COLOR_NCOMPONENT && cf == JXRHIP_COLOR_NCOMPONENT) dv.conv = kConvDirect;
    else if (cf == JXRHIP_COLOR_LUMA && oc == JXRHIP_COLOR_LUMA) dv.conv = kConvLuma;
    else if (cf == JXRHIP_COLOR_LUMA && oc == JXRHIP_COLOR_RGB) dv.conv = kConvLumaRgb;
    else if (yuv && (oc == JXRHIP_COLOR_RGB || oc == JXRHIP_COLOR_RGBE)) dv.conv = kConvYuvRgb;
    else if (cf == JXRHIP_COLOR_YUVK && oc == JXRHIP_COLOR_CMYK) dv.conv = kConvYuvkCmyk;
    else if (cf == JXRHIP_COLOR_YUVK && oc == JXRHIP_COLOR_CMYK_DIRECT) dv.conv = kConvYuvkDirect;
    else return fail(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "internal-to-output colour conversion");
    if (dv.upsample && (f.chroma_centering_x > 4 || (dv.upsample == 1 && f.chroma_centering_y > 4)))
      return fail(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "unknown chroma centering has no exact reconstruction filter");
    for (uint32_t c = 0; c < 8; c++) {
      dv.color_scale[c] = make_scale(c, oc, dep, f.shift_bits, f.scaled);
      if ((l == JXRHIP_LAYOUT_RGBX || l == JXRHIP_LAYOUT_BGRX) && c == 3) dv.color_scale[c].flags = 2;
    }
    dv.alpha_scale = make_scale(0, JXRHIP_COLOR_LUMA, adep, f.has_alpha_format ? f.alpha_shift_bits : f.shift_bits,
                                f.has_alpha_format ? f.alpha_scaled : f.scaled);
  }
  {
    // output arithmetic needs no overflow checks for samples within 2^27
    // when every bias and rounding constant is below 2^29 in magnitude
    bool small = true;
    auto check = [&](const ChanScale &c) {
      small &= (int64_t)std::llabs((long long)c.shifted_bias) + std::llabs((long long)c.rounding) < (1ll << 29);
    };
    for (const ChanScale &c : dv.color_scale) check(c);
    check(dv.alpha_scale);
    dv.out_fast = small ? 1 : 0;
  }
  {
    // specialised output recipes (plain-arithmetic tiles only)
    // a channel qualifies when it is neither padding nor invalid and its
    // post-shift (if any) is zero
    auto plain = [](const ChanScale &c) { return (c.flags & 3u) == 0 && (!(c.flags & 4u) || c.post_shift == 0); };
    const ChanScale &c0 = dv.color_scale[0];
    bool same = plain(c0);
    for (uint32_t c = 1; c < dv.channels && c < 8; c++) {
      const ChanScale &c1 = dv.color_scale[c];
      same &= plain(c1) && c1.shifted_bias == c0.shifted_bias && c1.rounding == c0.rounding && c1.shift == c0.shift;
    }
    const bool integer = f.storage == JXRHIP_STORE_U8 || f.storage == JXRHIP_STORE_U16 ||
                         f.storage == JXRHIP_STORE_I16 || f.storage == JXRHIP_STORE_I32;
    dv.recipe = kRecipeGeneric;
    if (dv.out_kind == kOutTyped && dv.out_fast && same && integer && !dv.has_alpha && !f.premultiply_alpha) {
      if (dv.conv == kConvYuvRgb && dv.channels == 3) dv.recipe = kRecipeRgb3;
      else if (dv.conv == kConvLuma && dv.channels == 1) dv.recipe = kRecipeLuma1;
    }
    if (dv.out_kind == kOutTyped && f.storage == JXRHIP_STORE_F32 && !f.premultiply_alpha) {
      // every stored channel a BD32F sample, colour first, alpha last (these
      // conversions have at most three colour planes: one plane group)
      const uint32_t primaries = dv.conv == kConvLuma ? 1u : 3u;
      bool ok = dv.conv == kConvYuvRgb || dv.conv == kConvLumaRgb || dv.conv == kConvLuma;
      ok &= dv.primary_count == primaries && dv.channels == primaries + dv.has_alpha;
      for (uint32_t c = 0; c < primaries && c < 8; c++) ok &= (dv.color_scale[c].flags & 2u) == 0;
      if (dv.has_alpha) ok &= (dv.alpha_scale.flags & 2u) == 0;
      const int64_t limit = 1ll << 23;
      const int64_t bias = f.exponent_bias;
      const int64_t abias = f.has_alpha_format ? f.alpha_exponent_bias : f.exponent_bias;
      ok &= bias >= -limit && bias <= limit && abias >= -limit && abias <= limit;
      if (ok) dv.recipe = kRecipeF32;
    }
    dv.recipe_add = c0.shifted_bias + c0.rounding;  // both below 2^29 when out_fast
    dv.recipe_shift = c0.shift;
    switch (f.storage) {
      case JXRHIP_STORE_U8: dv.recipe_lo = 0; dv.recipe_hi = 255; break;
      case JXRHIP_STORE_U16: dv.recipe_lo = 0; dv.recipe_hi = dv.u16_max; break;
      case JXRHIP_STORE_I16: dv.recipe_lo = -32768; dv.recipe_hi = 32767; break;
      default: dv.recipe_lo = INT32_MIN; dv.recipe_hi = INT32_MAX; break;
    }
  }
  dv.premultiply = f.premultiply_alpha ? 1 : 0;
  dv.centering_x = f.chroma_centering_x;
  dv.centering_y = f.chroma_centering_y;
  dv.f_scaled = f.scaled ? 1 : 0;
  dv.f_alpha_scaled = (f.has_alpha_format ? f.alpha_scaled : f.scaled) ? 1 : 0;
  dv.f_mant = f.mantissa_length;
  dv.f_bias = f.exponent_bias;
  dv.f_alpha_mant = f.has_alpha_format ? f.alpha_mantissa_length : f.mantissa_length;
  dv.f_alpha_bias = f.has_alpha_format ? f.alpha_exponent_bias : f.exponent_bias;
  {
    // crop in the coded plane (at the decode scale) -> relative to the
    // window's sample origin
    const uint64_t wx = (uint64_t)ox * mb_px, wy = (uint64_t)oy * mb_px;
    const jxrhip_surface_layout &s = d->surface;
    if (f.crop_width == 0 || f.crop_height == 0 || s.width != f.crop_width || s.height != f.crop_height)
      return fail(JXRHIP_ERR_INVALID_PLAN, "output layout and prepared request differ");
    if (f.crop_x < wx || f.crop_y < wy || (uint64_t)f.crop_x + f.crop_width > wx + (uint64_t)plan->mbx * mb_px ||
        (uint64_t)f.crop_y + f.crop_height > wy + (uint64_t)plan->mby * mb_px)
      return fail(JXRHIP_ERR_INVALID_PLAN, "output crop is outside a sample plane");
    dv.crop_x = (uint32_t)(f.crop_x - wx);
    dv.crop_y = (uint32_t)(f.crop_y - wy);
    dv.crop_w = f.crop_width;
    dv.crop_h = f.crop_height;
    const bool word = dv.out_kind == kOutPacked16 || dv.out_kind == kOutRgb101010 || dv.out_kind == kOutRgbe;
    if ((word || dv.out_kind == kOutBits) && s.planar)
      return fail(JXRHIP_ERR_INVALID_PLAN, "packed and one-bit outputs are interleaved");
    const uint32_t need = s.planar ? dv.channels : 1;
    if (s.plane_count != need || need > kMaxPlanes) return fail(JXRHIP_ERR_INVALID_PLAN, "surface plane count");
    dv.planar = s.planar ? 1 : 0;
    uint64_t content = 0;  // bytes the output stage writes
    for (uint32_t i = 0; i < need; i++) {
      const jxrhip_surface_plane &sp = s.planes[i];
      const uint32_t ch = s.planar ? 1 : dv.channels;
      uint32_t pw = s.width, ph = s.height;
      if (dv.out_kind == kOutYuv && (i == 1 || i == 2)) {
        // chroma planes at the coded sampling (planar.rs:77-106)
        pw = s.width / 2;
        ph = s.height / dv.chroma_sub_y;
        if (pw == 0 || ph == 0) return fail(JXRHIP_ERR_INVALID_PLAN, "planar chroma plane is empty");
      }
      if (sp.width != pw || sp.height != ph || sp.channels != ch)
        return fail(JXRHIP_ERR_INVALID_PLAN, "surface plane geometry");
      uint64_t row = (uint64_t)sp.width * ch * dv.bytes;
      if (dv.out_kind == kOutBits) row = ((uint64_t)sp.width + 7) / 8;  // format.rs:216
      if (word) row = (uint64_t)sp.width * dv.bytes;                     // one word a pixel
      content += row * sp.height;
      if (sp.row_stride_bytes < row) return fail(JXRHIP_ERR_INVALID_PLAN, "surface row stride");
      if (sp.byte_offset + (uint64_t)(sp.height - 1) * sp.row_stride_bytes + row > s.byte_len)
        return fail(JXRHIP_ERR_INVALID_PLAN, "surface plane exceeds its byte length");
      if (sp.byte_offset % dv.bytes || sp.row_stride_bytes % dv.bytes)
        return fail(JXRHIP_ERR_INVALID_PLAN, "surface plane is not aligned to its sample size");
      dv.plane_offset[i] = sp.byte_offset;
      dv.row_stride[i] = sp.row_stride_bytes;
    }
    plan->output_bytes = s.byte_len;
    plan->surface = s;
    plan->content_bytes = content;
  }
  // ---- work and footprint figures ----
  plan->pixels = (uint64_t)f.crop_width * f.crop_height;
  for (const PlanPlane &p : plan->planes) {
    // coefficients a decode at the plan's scale reconstructs: all of them,
    // the DC and low-pass of each block, or each macroblock's DC
    const uint64_t per_mb = plan->scale == JXRHIP_SCALE_QUARTER     ? p.bc * p.br
                            : plan->scale == JXRHIP_SCALE_SIXTEENTH ? 1
                                                                    : (uint64_t)p.bc * p.br * 16;
    plan->recon_coefs += (uint64_t)p.recs.size() * per_mb;
    plan->descriptor_bytes += p.recs.size() * sizeof(MbRec) + (p.colmask.size() + p.rowmask.size()) * 4 +
                              (p.probe_cols.size() + p.probe_rows.size()) * 4;
  }

  for (uint32_t a = 0; a < plan->arena_count; a++) plan->coefficient_bytes += plan->arenas[a]->coefficient_count * 4;
  {
    // algorithmic bytes: 4 B per stored coefficient read, 32 B per
    // macroblock record per plane, output bytes written
    // (a reduced decode reads the stored DC and low-pass values of each
    // macroblock, or its DC alone)
    uint64_t stored = 0, mbs = 0;
    for (const PlanPlane &p : plan->planes) {
      mbs += p.recs.size();
      for (const MbRec &r : p.recs) {
        const uint64_t coded = r.bands == 0 ? 1 : r.bands == 1 ? p.bc * p.br : (uint64_t)p.bc * p.br * 16;
        const uint64_t low = r.bands == 0 ? 1 : p.bc * p.br;
        stored += plan->scale == JXRHIP_SCALE_QUARTER ? low : plan->scale == JXRHIP_SCALE_SIXTEENTH ? 1 : coded;
      }
    }
    const uint64_t out_px_bytes = plan->content_bytes;
    plan->alg_bytes = stored * 4 + mbs * sizeof(MbRec) + out_px_bytes;
  }
  // ---- plane groups ----
  // A work item holds up to kMaxPlanes planes of its tile in LDS, and its
  // output stage reads up to four colour planes. An N-component image with
  // more colour planes runs as groups of up to four colour
  // planes, each a descriptor of its own that writes its channels of the
  // interleaved pixel; the group with the last colour planes also stores
  // alpha, and with premultiplied output every group carries the alpha
  // plane, since each multiplies its channels by it. On a planar surface
  // (at most JXRHIP_MAX_SURFACE_PLANES channels) a group writes its planes.
  dv.pixel_channels = dv.channels;
  dv.channel_base = 0;
  dv.store_channels = dv.channels;
  if (colors <= 4) {  // at most four colour planes and alpha: planes.size() <= kMaxPlanes
    plan->groups.push_back(dv);
    plan->group_planes.emplace_back();
    for (uint32_t i = 0; i < plan->planes.size(); i++) plan->group_planes.back().push_back(i);
  } else {
    if (dv.conv != kConvDirect || dv.out_kind != kOutTyped) return fail(JXRHIP_ERR_INVALID_PLAN, "too many planes");
    for (uint32_t g0 = 0; g0 < colors; g0 += 4) {
      const uint32_t n = std::min<uint32_t>(4, colors - g0);
      const bool last = g0 + n == colors;
      DevImage gd = dv;
      std::vector<uint32_t> idx;
      for (uint32_t k = 0; k < n; k++) idx.push_back(g0 + k);
      gd.colors = n;
      gd.primary_count = n;
      gd.has_alpha = 0;
      gd.alpha_plane = -1;
      if (dv.has_alpha && (last || dv.premultiply)) {
        gd.alpha_plane = (int32_t)idx.size();
        idx.push_back((uint32_t)alpha_plane);
        gd.has_alpha = 1;
      }
      gd.plane_count = (uint32_t)idx.size();
      gd.channels = n + gd.has_alpha;
      gd.store_channels = n + (dv.has_alpha && last ? 1u : 0u);
      gd.channel_base = g0;
      if (dv.planar) {
        // planar: the group's channels are surface planes g0 .. g0 + n (+ alpha)
        for (uint32_t c = 0; c < kMaxPlanes; c++) {
          gd.plane_offset[c] = g0 + c < kMaxPlanes ? dv.plane_offset[g0 + c] : 0;
          gd.row_stride[c] = g0 + c < kMaxPlanes ? dv.row_stride[g0 + c] : 0;
        }
        gd.channel_base = 0;
      }
      plan->groups.push_back(gd);
      plan->group_planes.push_back(std::move(idx));
    }
  }
  plan->descriptor_bytes += sizeof(DevImage) * plan->groups.size();
  // bucket key: plane geometries (of the first group, the largest) and the
  // storage kind (one kernel instantiation per storage kind)
  uint64_t key = plan->group_planes[0].size();
  for (uint32_t i : plan->group_planes[0]) key = key * 16 + plan->planes[i].bc * 4 + plan->planes[i].br;
  plan->geometry_key = key * 16 + dv.storage;
  // reduced decodes run the reduced kernels: buckets of their own (the
  // plane key stays far below bit 62)
  plan->geometry_key |= (uint64_t)plan->scale << 62;
  *out = plan.release();
  return JXRHIP_OK;
}

extern "C" void jxrhip_plan_destroy(jxrhip_plan *plan) { delete plan; }
extern "C" uint64_t jxrhip_plan_output_bytes(const jxrhip_plan *p) { return p ? p->output_bytes : 0; }
extern "C" uint64_t jxrhip_plan_reconstructed_coefficients(const jxrhip_plan *p) { return p ? p->recon_coefs : 0; }
extern "C" uint64_t jxrhip_plan_output_pixels(const jxrhip_plan *p) { return p ? p->pixels : 0; }
extern "C" uint64_t jxrhip_plan_device_footprint(const jxrhip_plan *p) {
  return p ? p->coefficient_bytes + p->descriptor_bytes + p->output_bytes : 0;
}
extern "C" uint64_t jxrhip_plan_algorithmic_bytes(const jxrhip_plan *p) { return p ? p->alg_bytes : 0; }

// ---- device memory pool (exact sizes, buffer_pool.rs:9-129) -----------
struct DeviceBuffer {
  void *ptr = nullptr;
  size_t bytes = 0;
};

class Pool {
 public:
  ~Pool() { trim(0); }
  hipError_t get(size_t bytes, DeviceBuffer &out, uint64_t &hits, uint64_t &misses) {
    bytes = (bytes + 255) & ~size_t(255);
    if (bytes == 0) bytes = 256;
    auto it = free_.find(bytes);
    if (it != free_.end() && !it->second.empty()) {
      out.ptr = it->second.back();
      it->second.pop_back();
      out.bytes = bytes;
      retained_ -= bytes;
      live_ += bytes;
      hits++;
      return hipSuccess;
    }
    misses++;
    hipError_t e = hipMalloc(&out.ptr, bytes);
    if (e != hipSuccess) {
      trim(0);
      e = hipMalloc(&out.ptr, bytes);
      if (e != hipSuccess) return e;
    }
    out.bytes = bytes;
    live_ += bytes;
    high_water_ = std::max(high_water_, live_ + retained_);
    return hipSuccess;
  }
  void put(DeviceBuffer &b) {
    if (!b.ptr) return;
    live_ -= b.bytes;
    if (retained_ + b.bytes > limit_) {
      (void)hipFree(b.ptr);
    } else {
      free_[b.bytes].push_back(b.ptr);
      retained_ += b.bytes;
    }
    b.ptr = nullptr;
    b.bytes = 0;
  }
  void trim(size_t keep) {
    for (auto &kv : free_) {
      while (!kv.second.empty() && retained_ > keep) {
        (void)hipFree(kv.second.back());
        kv.second.pop_back();
        retained_ -= kv.first;
      }
    }
  }
  uint64_t high_water() const { return high_water_; }

 private:
  std::map<size_t, std::vector<void *>> free_;
  size_t retained_ = 0, live_ = 0, high_water_ = 0;
  size_t limit_ = size_t(8) << 30;
};

// ---- sessions -----------------------------------------------------------
static constexpr uint32_t kStreams = 4;  // runtime.rs:17

struct CacheEntry {
  DeviceBuffer buf;
  uint64_t last_use = 0;
  const void *deps[2] = {nullptr, nullptr};  // coefficient buffers a descriptor points at
  uint32_t pins = 0;  // prepared batches and graphs that read it; never evicted while pinned
  // coefficient uploads are asynchronous: `ready` is recorded on `stream`
  // after the copy that filled the entry, and work on any other stream that
  // reads the entry before the copy has finished waits for it on the device
  hipEvent_t ready = nullptr;
  hipStream_t stream = nullptr;
};

struct StageBlock {
  uint8_t *p = nullptr;
  size_t bytes = 0;
  hipEvent_t done = nullptr;  // recorded after the copy that reads the block
};

struct jxrhip_session {
  int device = 0;
  uint32_t cus = 0;
  size_t max_lds = 65536;
  hipStream_t streams[kStreams] = {};
  hipEvent_t t0 = nullptr, t1 = nullptr;
  hipEvent_t fork = nullptr, join[kStreams] = {};  // fan-out of a batch over the streams
  std::mutex lock;  // one submitter at a time (encode.rs:83-88)
  Pool pool;
  // coefficient uploads by arena identity; descriptor uploads by plan id
  std::unordered_map<uint64_t, CacheEntry> coef_cache;
  std::unordered_map<uint64_t, CacheEntry> desc_cache;
  uint64_t cache_bytes = 0, cache_budget = uint64_t(64) << 30, tick = 0;
  uint32_t next_stream = 0;
  uint32_t tile_x = 0, tile_y = 0, groups_per_cu = 0;
  std::map<uint64_t, std::pair<uint32_t, uint32_t>> tuned;  // bucket key -> tile shape
  // bucket key -> workgroups per CU the autotuner chose below the most that
  // fit (absent: as many as fit)
  std::map<uint64_t, uint32_t> tuned_groups;
  uint32_t live_graphs = 0, live_servers = 0;  // objects that must be destroyed before the session
  // owners of the session: the caller's handle, plus every submission,
  // result image and batch not yet released. The last one frees it, so an
  // image may outlive jxrhip_session_destroy (a resident image keeps its
  // device context alive, resident.rs:8-12).
  std::atomic<uint32_t> refs{1};
  // device destinations from jxrhip_allocate_destination; each one is an
  // owner too, until jxrhip_free_destination
  std::unordered_map<void *, uint64_t> destinations;
  std::unordered_set<void *> host_blocks;  // jxrhip_host_alloc blocks
  double last_ms = 0;
  uint32_t last_launches = 0;
  jxrhip_diagnostics diag{};
  // status words of in-flight per-image submissions (device + pinned host
  // mirror) and reusable completion events, so a submission allocates nothing
  uint32_t *status_dev = nullptr;
  uint32_t *status_host = nullptr;
  // the device's static scratch array (allocation-free batches)
  uint8_t *scratch = nullptr;
  std::vector<uint32_t> free_status;
  std::vector<hipEvent_t> free_events;
  // the sample-scale terms of reduced decodes, by probe key: computed on the
  // device once per key and kept until the session ends (descriptors point
  // at them; each is a few kilobytes)
  std::map<std::vector<uint32_t>, DeviceBuffer> probe_terms;
  // page-locked staging of descriptor uploads: a copy from pageable memory
  // would wait for the DMA already queued (the coefficient uploads of the
  // same submission), so descriptors are built in pinned blocks, copied
  // asynchronously, and a block is reused once its copy's event completes
  std::vector<StageBlock> stage_free, stage_busy;
  size_t stage_free_bytes = 0;
};

static constexpr uint32_t kStatusSlots = 16384;

struct jxrhip_image {
  jxrhip_session *session = nullptr;
  DeviceBuffer buf;
  uint64_t bytes = 0;
  // a serving loop's result: the bytes it counts against that loop's budget
  // until the caller destroys the image
  std::shared_ptr<std::atomic<uint64_t>> held;
};

struct jxrhip_submission {
  jxrhip_session *session = nullptr;
  jxrhip_surface_layout layout{};  // the plan's output layout
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  uint32_t slot = UINT32_MAX;  // status word index
  DeviceBuffer output;    // owned output (submit), empty for submit_into
  uint64_t output_bytes = 0;
  std::vector<DeviceBuffer> transient;
  bool waited = false;
};

static void make_current(jxrhip_session *s) { (void)hipSetDevice(s->device); }

static void session_free(jxrhip_session *s);
static void session_ref(jxrhip_session *s) { s->refs.fetch_add(1, std::memory_order_relaxed); }
static void session_unref(jxrhip_session *s) {
  if (s->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) session_free(s);
}

// An event for one-shot device ordering (completion of a submission, an
// upload): reused from the session's free list when one is there.
static hipError_t take_event(jxrhip_session *s, hipEvent_t *e) {
  if (!s->free_events.empty()) {
    *e = s->free_events.back();
    s->free_events.pop_back();
    return hipSuccess;
  }
  return hipEventCreateWithFlags(e, hipEventDisableTiming);
}

// A pinned staging block of at least `bytes`: the smallest free one, after
// moving blocks whose copies have completed to the free list, or a new one
// (64 KiB or the next power of two). Blocks beyond 256 MiB of free staging
// are released.
static constexpr size_t kStageKeepBytes = size_t(256) << 20;
static hipError_t stage_get(jxrhip_session *s, size_t bytes, StageBlock &out) {
  for (size_t i = 0; i < s->stage_busy.size();) {
    if (hipEventQuery(s->stage_busy[i].done) == hipSuccess) {
      s->stage_free_bytes += s->stage_busy[i].bytes;
      s->stage_free.push_back(s->stage_busy[i]);
      s->stage_busy[i] = s->stage_busy.back();
      s->stage_busy.pop_back();
    } else {
      i++;
    }
  }
  size_t best = SIZE_MAX;
  for (size_t i = 0; i < s->stage_free.size(); i++)
    if (s->stage_free[i].bytes >= bytes && (best == SIZE_MAX || s->stage_free[i].bytes < s->stage_free[best].bytes))
      best = i;
  if (best != SIZE_MAX) {
    out = s->stage_free[best];
    s->stage_free_bytes -= out.bytes;
    s->stage_free[best] = s->stage_free.back();
    s->stage_free.pop_back();
    return hipSuccess;
  }
  while (s->stage_free_bytes > kStageKeepBytes && !s->stage_free.empty()) {
    StageBlock &b = s->stage_free.back();
    s->stage_free_bytes -= b.bytes;
    (void)hipHostFree(b.p);
    (void)hipEventDestroy(b.done);
    s->stage_free.pop_back();
  }
  StageBlock b;
  b.bytes = size_t(64) << 10;
  while (b.bytes < bytes) b.bytes <<= 1;
  hipError_t e = hipHostMalloc((void **)&b.p, b.bytes, hipHostMallocDefault);
  if (e != hipSuccess) return e;
  e = hipEventCreateWithFlags(&b.done, hipEventDisableTiming);
  if (e != hipSuccess) {
    (void)hipHostFree(b.p);
    return e;
  }
  out = b;
  return hipSuccess;
}
// The block is read by a copy queued on `stream`: busy until that point.
static hipError_t stage_release(jxrhip_session *s, StageBlock &b, hipStream_t stream) {
  const hipError_t e = hipEventRecord(b.done, stream);
  if (e != hipSuccess) {
    // without its event the copy's end is unknown: wait for the stream
    (void)hipStreamSynchronize(stream);
    s->stage_free_bytes += b.bytes;
    s->stage_free.push_back(b);
    return e;
  }
  s->stage_busy.push_back(b);
  return hipSuccess;
}

// Returns a cache entry's memory to the pool (the caller has made sure no
// queued work reads it) and its upload event to the free list.
static void drop_entry(jxrhip_session *s, CacheEntry &ce) {
  s->cache_bytes -= ce.buf.bytes;
  s->pool.put(ce.buf);
  if (ce.ready) s->free_events.push_back(ce.ready);
  ce.ready = nullptr;
}

extern "C" jxrhip_status jxrhip_session_create(int32_t ordinal, jxrhip_session **out) {
  if (!out) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null output");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return fail(JXRHIP_ERR_UNAVAILABLE, "no HIP device is available");
  }
  if (ordinal < 0 || ordinal >= n) return fail(JXRHIP_ERR_UNAVAILABLE, "device ordinal out of range");
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, ordinal), "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(JXRHIP_ERR_UNAVAILABLE, std::string("device is ") + prop.gcnArchName + ", kernels are built for gfx950");
  std::unique_ptr<jxrhip_session> s(new (std::nothrow) jxrhip_session());
  if (!s) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "out of host memory");
  s->device = ordinal;
  s->cus = (uint32_t)prop.multiProcessorCount;
  HIP_TRY(hipSetDevice(ordinal), "hipSetDevice");
  for (uint32_t i = 0; i < kStreams; i++)
    HIP_TRY(hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking), "hipStreamCreate");
  HIP_TRY(hipEventCreate(&s->t0), "hipEventCreate");
  HIP_TRY(hipEventCreate(&s->t1), "hipEventCreate");
  HIP_TRY(hipEventCreateWithFlags(&s->fork, hipEventDisableTiming), "hipEventCreate");
  for (uint32_t i = 0; i < kStreams; i++)
    HIP_TRY(hipEventCreateWithFlags(&s->join[i], hipEventDisableTiming), "hipEventCreate");
  size_t lds_limit = prop.sharedMemPerBlockOptin ? prop.sharedMemPerBlockOptin : prop.sharedMemPerBlock;
  s->max_lds = std::min<size_t>(lds_limit, 160 * 1024);
  if (s->max_lds > 65536 && set_kernel_lds_limit(s->max_lds) != hipSuccess) {
    (void)hipGetLastError();
    s->max_lds = 65536;
  }
  HIP_TRY(static_scratch_address(ordinal, &s->scratch), "static scratch address");
  HIP_TRY(hipMalloc((void **)&s->status_dev, kStatusSlots * sizeof(uint32_t)), "status words");
  HIP_TRY(hipHostMalloc((void **)&s->status_host, kStatusSlots * sizeof(uint32_t), hipHostMallocDefault),
          "status words");
  s->free_status.reserve(kStatusSlots);
  for (uint32_t i = kStatusSlots; i-- > 0;) s->free_status.push_back(i);
  s->diag.device_ordinal = ordinal;
  s->diag.stream_count = kStreams;
  s->diag.compute_units = s->cus;
  *out = s.release();
  return JXRHIP_OK;
}

extern "C" void jxrhip_session_destroy(jxrhip_session *s) {
  if (!s) return;
  {
    // graphs and servers point into the session: destroying it under them
    // would leave them dangling, so it is refused (and the session leaks)
    std::lock_guard<std::mutex> g(s->lock);
    if (s->live_graphs || s->live_servers) {
      (void)fail(JXRHIP_ERR_INVALID_ARGUMENT, "session still has captured batches or servers");
      return;
    }
  }
  // submissions, images and batches still alive keep the session until the
  // last of them is released
  session_unref(s);
}

static void session_free(jxrhip_session *s) {
  make_current(s);
  for (uint32_t i = 0; i < kStreams; i++)
    if (s->streams[i]) (void)hipStreamSynchronize(s->streams[i]);
  for (auto &kv : s->coef_cache) drop_entry(s, kv.second);
  for (auto &kv : s->desc_cache) drop_entry(s, kv.second);
  for (auto &kv : s->probe_terms) s->pool.put(kv.second);
  if (s->status_dev) (void)hipFree(s->status_dev);
  if (s->status_host) (void)hipHostFree(s->status_host);
  for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
  for (auto *list : {&s->stage_free, &s->stage_busy})
    for (StageBlock &b : *list) {
      (void)hipEventSynchronize(b.done);
      (void)hipHostFree(b.p);
      (void)hipEventDestroy(b.done);
    }
  for (uint32_t i = 0; i < kStreams; i++)
    if (s->streams[i]) (void)hipStreamDestroy(s->streams[i]);
  if (s->t0) (void)hipEventDestroy(s->t0);
  if (s->t1) (void)hipEventDestroy(s->t1);
  if (s->fork) (void)hipEventDestroy(s->fork);
  for (uint32_t i = 0; i < kStreams; i++)
    if (s->join[i]) (void)hipEventDestroy(s->join[i]);
  delete s;
}

extern "C" jxrhip_status jxrhip_session_diagnostics(jxrhip_session *s, jxrhip_diagnostics *out) {
  if (!s || !out) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(s->lock);
  s->diag.upload_resident_bytes = s->cache_bytes;
  s->diag.scratch_high_water_bytes = s->pool.high_water();
  *out = s->diag;
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_session_set_upload_budget(jxrhip_session *s, uint64_t bytes) {
  if (!s) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null session");
  std::lock_guard<std::mutex> g(s->lock);
  s->cache_budget = bytes;
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_session_set_tile(jxrhip_session *s, uint32_t tx, uint32_t ty) {
  if (!s) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null session");
  if ((tx == 0) != (ty == 0) || tx > 16 || ty > 16) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "tile shape");
  std::lock_guard<std::mutex> g(s->lock);
  s->tile_x = tx;
  s->tile_y = ty;
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_session_set_occupancy(jxrhip_session *s, uint32_t groups) {
  if (!s) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null session");
  std::lock_guard<std::mutex> g(s->lock);
  s->groups_per_cu = groups;
  return JXRHIP_OK;
}

static void evict_until(jxrhip_session *s, uint64_t need) {
  while (s->cache_bytes + need > s->cache_budget) {
    // least recently used of both caches
    std::unordered_map<uint64_t, CacheEntry> *victim_map = nullptr;
    uint64_t victim = 0, oldest = UINT64_MAX;
    for (auto *m : {&s->coef_cache, &s->desc_cache})
      for (auto &kv : *m)
        if (!kv.second.pins && kv.second.last_use < oldest) {
          oldest = kv.second.last_use;
          victim = kv.first;
          victim_map = m;
        }
    if (!victim_map) return;
    auto it = victim_map->find(victim);
    // a buffer still read by queued work must not be recycled early; work
    // of a serving loop reads only pinned entries
    for (uint32_t i = 0; i < kStreams; i++) (void)hipStreamSynchronize(s->streams[i]);
    if (it->second.ready) (void)hipEventSynchronize(it->second.ready);
    drop_entry(s, it->second);
    victim_map->erase(it);
    s->diag.upload_evictions++;
  }
}

// Orders work queued next on `stream` after the copy that filled a cached
// coefficient entry, if that copy may still be running on another stream.
static hipError_t await_upload(jxrhip_session *s, CacheEntry &ce, hipStream_t stream) {
  if (!ce.ready) return hipSuccess;
  const hipError_t q = hipEventQuery(ce.ready);
  if (q == hipSuccess) {
    s->free_events.push_back(ce.ready);
    ce.ready = nullptr;
    return hipSuccess;
  }
  if (q != hipErrorNotReady) return q;
  return ce.stream == stream ? hipSuccess : hipStreamWaitEvent(stream, ce.ready, 0);
}

// Device-resident coefficients of arena `a` (cached by identity when set).
static jxrhip_status device_coefficients(jxrhip_session *s, const jxrhip_arena *a, hipStream_t stream,
                                         const int32_t **out, jxrhip_submission *transient_owner) {
  const size_t bytes = (size_t)a->coefficient_count * 4;
  if (a->identity) {
    auto it = s->coef_cache.find(a->identity);
    if (it != s->coef_cache.end()) {
      CacheEntry &ce = it->second;
      HIP_TRY(await_upload(s, ce, stream), "upload dependency");
      ce.last_use = ++s->tick;
      s->diag.upload_hits++;
      *out = (const int32_t *)ce.buf.ptr;
      return JXRHIP_OK;
    }
  }
  s->diag.upload_misses++;
  DeviceBuffer b;
  if (a->identity) evict_until(s, bytes);
  HIP_TRY(s->pool.get(bytes + 64, b, s->diag.scratch_hits, s->diag.scratch_misses), "coefficient buffer");
  HIP_TRY(hipMemcpyAsync(b.ptr, a->coefficients, bytes, hipMemcpyHostToDevice, stream), "coefficient upload");
  if (a->identity) {
    CacheEntry ce{b, ++s->tick, {nullptr, nullptr}};
    hipError_t e = take_event(s, &ce.ready);
    if (e == hipSuccess) e = hipEventRecord(ce.ready, stream);
    if (e != hipSuccess) {
      if (ce.ready) s->free_events.push_back(ce.ready);
      (void)hipStreamSynchronize(stream);
      s->pool.put(b);
      return hip_fail(e, "upload event");
    }
    ce.stream = stream;
    s->cache_bytes += b.bytes;
    s->coef_cache[a->identity] = ce;
  } else if (transient_owner) {
    transient_owner->transient.push_back(b);
  } else {
    HIP_TRY(hipStreamSynchronize(stream), "upload");
    s->pool.put(b);
    return fail(JXRHIP_ERR_INVALID_ARGUMENT, "an arena without identity cannot be uploaded ahead of a submission");
  }
  *out = (const int32_t *)b.ptr;
  return JXRHIP_OK;
}

// The terms of probe sets `which` of a plan's sample scale
// (sample_scale.rs:157-246), computed on the device into `terms[k]`: the
// forward analysis gives each set's eight flat probe images (+-level, four
// levels) on the host, the reduced kernel decodes them into raw reduced
// pictures, and jxr_scale_terms fits the line through them for every probe
// sample. Waits for the device; an overflow in a probe, or a term whose
// divisor is not positive, fails.
static jxrhip_status run_probes(jxrhip_session *s, const jxrhip_plan *p, const std::vector<size_t> &which,
                                const std::vector<int64_t *> &terms, hipStream_t stream) {
  constexpr uint32_t kImages = 2 * kProbeLevels;
  // the flat probe images' DC and low-pass coefficients, +L1, -L1, +L2, ...
  std::vector<std::vector<int32_t>> coef(which.size() * kImages);
  for (size_t w = 0; w < which.size(); w++) {
    const ProbeSet &q = p->probes[which[w]];
    for (uint32_t k = 0; k < kProbeLevels; k++) {
      const int32_t level = 0x10001 + (int32_t)k;
      flat_probe_coefficients(level, q.bc, q.br, q.scale_first, q.overlap, q.bounds_x, q.bounds_y,
                              coef[w * kImages + 2 * k]);
      flat_probe_coefficients(-level, q.bc, q.br, q.scale_first, q.overlap, q.bounds_x, q.bounds_y,
                              coef[w * kImages + 2 * k + 1]);
    }
  }
  // layout: status word | per set: records, masks, sample scale, pictures,
  // then each image's coefficients and descriptor
  size_t bytes = 256;
  for (size_t w = 0; w < which.size(); w++) {
    const ProbeSet &q = p->probes[which[w]];
    bytes += q.recs.size() * sizeof(MbRec) + (q.colmask.size() + q.rowmask.size()) * 4 + sizeof(DevReduce) +
             (size_t)kImages * q.length * 4 + 5 * 64;
    for (uint32_t j = 0; j < kImages; j++) bytes += coef[w * kImages + j].size() * 4 + sizeof(DevImage) + 2 * 64;
  }
  std::vector<uint8_t> host(bytes, 0);
  DeviceBuffer b;
  HIP_TRY(s->pool.get(bytes, b, s->diag.scratch_hits, s->diag.scratch_misses), "probe buffer");
  uint8_t *dev = (uint8_t *)b.ptr;
  size_t off = 256;  // the status word is at 0
  auto place = [&](const void *src, size_t n) -> uint8_t * {
    off = (off + 63) & ~size_t(63);
    if (src && n) std::memcpy(host.data() + off, src, n);
    uint8_t *d = dev + off;
    off += n;
    return d;
  };
  struct ProbeLaunch {
    const DevImage *img;
    int32_t *pictures;
    const ProbeSet *set;
  };
  std::vector<ProbeLaunch> launches;
  std::vector<int32_t *> pictures;
  for (size_t w = 0; w < which.size(); w++) {
    const ProbeSet &q = p->probes[which[w]];
    const MbRec *recs = (const MbRec *)place(q.recs.data(), q.recs.size() * sizeof(MbRec));
    const uint32_t *cm = (const uint32_t *)place(q.colmask.data(), q.colmask.size() * 4);
    const uint32_t *rm = (const uint32_t *)place(q.rowmask.data(), q.rowmask.size() * 4);
    DevReduce r;
    std::memset(&r, 0, sizeof r);
    r.mode = kReduceNone;
    const DevReduce *rd = (const DevReduce *)place(&r, sizeof r);
    int32_t *pic = (int32_t *)place(nullptr, (size_t)kImages * q.length * 4);
    pictures.push_back(pic);
    for (uint32_t j = 0; j < kImages; j++) {
      const std::vector<int32_t> &c = coef[w * kImages + j];
      const int32_t *cd = (const int32_t *)place(c.data(), c.size() * 4);
      DevImage im;
      std::memset(&im, 0, sizeof im);
      DevPlane &pl = im.planes[0];
      pl.coef = cd;
      pl.mbs = recs;
      pl.colmask = cm;
      pl.rowmask = rm;
      pl.bc = q.bc;
      pl.br = q.br;
      pl.scale_first = q.scale_first ? 1 : 0;
      pl.overlap = q.overlap;
      pl.alpha = q.alpha ? 1 : 0;
      im.plane_count = 1;
      im.colors = 1;
      im.alpha_plane = -1;
      im.mbx = q.mbx;
      im.mby = q.mby;
      im.scale = p->scale;
      im.mb_px = p->groups[0].mb_px;
      im.raw = 1;
      im.reduce = rd;
      im.storage = JXRHIP_STORE_I32;
      im.bytes = 4;
      im.plane_offset[0] = (uint64_t)j * q.length * 4;
      im.row_stride[0] = (uint64_t)q.stride * 4;
      launches.push_back({(const DevImage *)place(&im, sizeof im), pic, &q});
    }
  }
  auto undo = [&](jxrhip_status st) {
    (void)hipStreamSynchronize(stream);
    s->pool.put(b);
    return st;
  };
  hipError_t e = hipMemcpyAsync(dev, host.data(), off, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return undo(hip_fail(e, "probe upload"));
  for (const ProbeLaunch &L : launches) {
    const ProbeSet &q = *L.set;
    const uint32_t ux = reduced_units(q.bc, p->scale), uy = reduced_units(q.br, p->scale);
    const uint32_t tx = 4, ty = 4;
    KernelArgs args{};
    args.inline_slot.img = L.img;
    args.inline_slot.out = (uint8_t *)L.pictures;
    args.inline_slot.status = (uint32_t *)dev;
    args.inline_slot.tiles_x = (q.mbx + tx - 1) / tx;
    args.persistent = 0;
    args.tile_x = tx;
    args.tile_y = ty;
    args.lds_words = reduced_sample_words(ux, uy, tx, ty) + low_region_words(q.bc, q.br, q.overlap, tx, ty) +
                     mb_table_words(tx, ty);
    args.storage = JXRHIP_STORE_I32;
    args.reduced = 1;
    const uint32_t tiles = args.inline_slot.tiles_x * ((q.mby + ty - 1) / ty);
    args.item_count = tiles;
    if ((e = launch_tiles(args, (tiles + 7u) & ~7u, stream)) != hipSuccess) return undo(hip_fail(e, "probe launch"));
    s->diag.kernel_launches++;
  }
  for (size_t w = 0; w < which.size(); w++) {
    const ProbeSet &q = p->probes[which[w]];
    if ((e = launch_scale_terms(pictures[w], q.length, kProbeLevels, q.span, terms[which[w]], (uint32_t *)dev,
                                stream)) != hipSuccess)
      return undo(hip_fail(e, "sample scale terms"));
    s->diag.kernel_launches++;
  }
  uint32_t status = 0;
  e = hipMemcpyAsync(&status, dev, 4, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return undo(hip_fail(e, "probe status"));
  s->pool.put(b);
  if (status) return fail(JXRHIP_ERR_KERNEL_ARITHMETIC, "reduced-resolution sample scale");
  return JXRHIP_OK;
}

// The device terms of every probe set of a plan, from the session's table,
// computing the sets it does not hold yet.
static jxrhip_status sample_scale_terms(jxrhip_session *s, const jxrhip_plan *p, hipStream_t stream,
                                       std::vector<int64_t *> &terms) {
  terms.assign(p->probes.size(), nullptr);
  std::vector<size_t> missing;
  std::vector<DeviceBuffer> fresh;
  for (size_t k = 0; k < p->probes.size(); k++) {
    auto it = s->probe_terms.find(p->probes[k].key);
    if (it != s->probe_terms.end()) {
      terms[k] = (int64_t *)it->second.ptr;
      continue;
    }
    // two sets of one plan may share a key: compute it once
    bool dup = false;
    for (size_t m : missing)
      if (p->probes[m].key == p->probes[k].key) {
        terms[k] = terms[m];
        dup = true;
      }
    if (dup) continue;
    DeviceBuffer b;
    const hipError_t e = s->pool.get((size_t)p->probes[k].length * 16, b, s->diag.scratch_hits, s->diag.scratch_misses);
    if (e != hipSuccess) {
      for (auto &f : fresh) s->pool.put(f);
      return hip_fail(e, "probe terms");
    }
    terms[k] = (int64_t *)b.ptr;
    missing.push_back(k);
    fresh.push_back(b);
  }
  if (missing.empty()) return JXRHIP_OK;
  const jxrhip_status st = run_probes(s, p, missing, terms, stream);
  if (st) {
    for (auto &f : fresh) s->pool.put(f);
    return st;
  }
  for (size_t i = 0; i < missing.size(); i++) s->probe_terms[p->probes[missing[i]].key] = fresh[i];
  return JXRHIP_OK;
}

// Device DevImage for a plan (descriptors cached per plan id). The
// coefficient pointers inside depend on the cached uploads, so the
// descriptor is rebuilt whenever an upload changes.
static jxrhip_status device_image(jxrhip_session *s, const jxrhip_plan *p, hipStream_t stream,
                                  const DevImage **out, jxrhip_submission *owner) {
  const int32_t *coef[2] = {nullptr, nullptr};
  for (uint32_t a = 0; a < p->arena_count; a++) {
    jxrhip_status st = device_coefficients(s, p->arenas[a], stream, &coef[a], owner);
    if (st) return st;
  }
  const bool cacheable = p->arenas[0]->identity && (p->arena_count < 2 || p->arenas[1]->identity);
  auto it = s->desc_cache.find(p->id);
  if (cacheable && it != s->desc_cache.end()) {
    // valid while the coefficient buffers it points at are the same; its
    // copy may still be running on the stream that made it
    if (it->second.deps[0] == coef[0] && it->second.deps[1] == coef[1]) {
      HIP_TRY(await_upload(s, it->second, stream), "descriptor dependency");
      it->second.last_use = ++s->tick;
      *out = (const DevImage *)it->second.buf.ptr;
      return JXRHIP_OK;
    }
  }
  // layout: DevImage of each plane group | recs of each plane | masks, and
  // for a reduced decode: the sample scale of each group plane | each
  // plane's probe maps | the terms of each probe set (written on the device)
  const size_t ngroups = p->groups.size();
  size_t bytes = sizeof(DevImage) * ngroups;
  for (const PlanPlane &pp : p->planes)
    bytes += pp.recs.size() * sizeof(MbRec) + (pp.colmask.size() + pp.rowmask.size()) * 4 + 64 +
             (pp.probe_cols.size() + pp.probe_rows.size()) * 4;
  if (p->scale != JXRHIP_SCALE_FULL)
    for (const auto &gp : p->group_planes) bytes += gp.size() * sizeof(DevReduce) + 64;
  // the sample-scale terms come first: the descriptors point at them
  std::vector<int64_t *> terms;
  if (!p->probes.empty()) {
    const jxrhip_status pst = sample_scale_terms(s, p, stream, terms);
    if (pst) return pst;
  }
  DeviceBuffer b;
  if (cacheable) evict_until(s, bytes);
  HIP_TRY(s->pool.get(bytes, b, s->diag.scratch_hits, s->diag.scratch_misses), "descriptor buffer");
  StageBlock stage;
  {
    const hipError_t se = stage_get(s, bytes, stage);
    if (se != hipSuccess) {
      s->pool.put(b);
      return hip_fail(se, "descriptor staging");
    }
  }
  // every byte the kernels read is written below; alignment gaps between
  // the parts are never read, so the block is not cleared first
  uint8_t *const host = stage.p;
  std::vector<DevPlane> dp(p->planes.size());
  size_t off = sizeof(DevImage) * ngroups;
  for (size_t i = 0; i < p->planes.size(); i++) {
    const PlanPlane &pp = p->planes[i];
    DevPlane &d = dp[i];
    d.coef = coef[pp.arena];
    off = (off + 63) & ~size_t(63);
    std::memcpy(host + off, pp.recs.data(), pp.recs.size() * sizeof(MbRec));
    d.mbs = (const MbRec *)((uint8_t *)b.ptr + off);
    off += pp.recs.size() * sizeof(MbRec);
    if (!pp.colmask.empty()) {
      std::memcpy(host + off, pp.colmask.data(), pp.colmask.size() * 4);
      d.colmask = (const uint32_t *)((uint8_t *)b.ptr + off);
      off += pp.colmask.size() * 4;
      std::memcpy(host + off, pp.rowmask.data(), pp.rowmask.size() * 4);
      d.rowmask = (const uint32_t *)((uint8_t *)b.ptr + off);
      off += pp.rowmask.size() * 4;
    } else {
      d.colmask = d.rowmask = nullptr;
    }
    d.bc = pp.bc;
    d.br = pp.br;
    d.scale_first = pp.scale_first ? 1 : 0;
    d.overlap = pp.overlap;
    d.alpha = pp.alpha ? 1 : 0;
  }
  // reduced decode: each plane's sample scale
  std::vector<DevReduce> dr(p->planes.size());
  if (p->scale != JXRHIP_SCALE_FULL) {
    for (size_t i = 0; i < p->planes.size(); i++) {
      const PlanPlane &pp = p->planes[i];
      DevReduce &r = dr[i];
      std::memset(&r, 0, sizeof r);
      r.mode = pp.reduce_mode;
      r.shift = pp.reduce_shift;
      if (pp.reduce_mode == kReduceFlat) {
        const ProbeSet &q = p->probes[pp.probe];
        r.terms = terms[pp.probe];
        r.gain = q.gain;
        r.stride = q.stride;
        std::memcpy(host + off, pp.probe_cols.data(), pp.probe_cols.size() * 4);
        r.cols = (const uint32_t *)((uint8_t *)b.ptr + off);
        off += pp.probe_cols.size() * 4;
        std::memcpy(host + off, pp.probe_rows.data(), pp.probe_rows.size() * 4);
        r.rows = (const uint32_t *)((uint8_t *)b.ptr + off);
        off += pp.probe_rows.size() * 4;
      }
    }
  }
  for (size_t g = 0; g < ngroups; g++) {
    DevImage dv = p->groups[g];
    for (size_t k = 0; k < p->group_planes[g].size(); k++) dv.planes[k] = dp[p->group_planes[g][k]];
    if (p->scale != JXRHIP_SCALE_FULL) {
      off = (off + 63) & ~size_t(63);
      dv.reduce = (const DevReduce *)((uint8_t *)b.ptr + off);
      for (size_t k = 0; k < p->group_planes[g].size(); k++) {
        std::memcpy(host + off, &dr[p->group_planes[g][k]], sizeof(DevReduce));
        off += sizeof(DevReduce);
      }
    }
    std::memcpy(host + g * sizeof(DevImage), &dv, sizeof dv);
  }
  // stream-ordered ahead of the kernels that read it, from pinned memory, so
  // nothing waits here for the coefficient uploads queued before it
  {
    const hipError_t ce = hipMemcpyAsync(b.ptr, host, bytes, hipMemcpyHostToDevice, stream);
    const hipError_t re = stage_release(s, stage, stream);
    if (ce != hipSuccess || re != hipSuccess) {
      (void)hipStreamSynchronize(stream);
      s->pool.put(b);
      return hip_fail(ce != hipSuccess ? ce : re, "descriptor upload");
    }
  }
  if (cacheable) {
    // the eviction above may have removed the stale entry already
    it = s->desc_cache.find(p->id);
    uint32_t pins = 0;
    if (it != s->desc_cache.end()) {
      drop_entry(s, it->second);
      pins = it->second.pins;
    }
    CacheEntry ce{b, ++s->tick, {coef[0], coef[1]}, pins};
    // the copy is asynchronous: work on other streams that finds this entry
    // waits for it (await_upload)
    hipError_t e = take_event(s, &ce.ready);
    if (e == hipSuccess) e = hipEventRecord(ce.ready, stream);
    if (e != hipSuccess) {
      if (ce.ready) s->free_events.push_back(ce.ready);
      ce.ready = nullptr;
      (void)hipStreamSynchronize(stream);
    }
    ce.stream = stream;
    s->cache_bytes += b.bytes;
    s->desc_cache[p->id] = ce;
  } else {
    owner->transient.push_back(b);
  }
  *out = (const DevImage *)b.ptr;
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_session_upload(jxrhip_session *s, const jxrhip_plan *p) {
  if (!s || !p) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(s->lock);
  make_current(s);
  for (uint32_t a = 0; a < p->arena_count; a++)
    if (!p->arenas[a]->identity)
      return fail(JXRHIP_ERR_INVALID_ARGUMENT, "an arena without identity cannot be uploaded ahead of a submission");
  const DevImage *dimg;
  jxrhip_status st = device_image(s, p, s->streams[0], &dimg, nullptr);
  if (st) return st;
  HIP_TRY(hipStreamSynchronize(s->streams[0]), "upload");
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_session_clear_cache(jxrhip_session *s) {
  if (!s) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null session");
  std::lock_guard<std::mutex> g(s->lock);
  make_current(s);
  for (uint32_t i = 0; i < kStreams; i++) (void)hipStreamSynchronize(s->streams[i]);
  // entries that a prepared batch or a graph still reads stay
  for (auto *m : {&s->coef_cache, &s->desc_cache})
    for (auto it = m->begin(); it != m->end();) {
      if (it->second.pins) {
        ++it;
        continue;
      }
      drop_entry(s, it->second);
      it = m->erase(it);
    }
  s->pool.trim(0);
  return JXRHIP_OK;
}

// LDS words of a tile of tx x ty macroblocks: a sample region and a low
// region for every plane.
static uint32_t tile_words(const jxrhip_plan *p, uint32_t tx, uint32_t ty) {
  uint32_t most = 0;
  for (const auto &gp : p->group_planes) {
    uint32_t w = 0;
    for (uint32_t i : gp) {
      const PlanPlane &pp = p->planes[i];
      const uint32_t samples = p->scale != JXRHIP_SCALE_FULL ? reduced_sample_words(pp.ux, pp.uy, tx, ty)
                                                             : sample_region_words(pp.bc, pp.br, tx, ty);
      w += samples + low_region_words(pp.bc, pp.br, pp.overlap, tx, ty) + mb_table_words(tx, ty);
    }
    most = std::max(most, w);
  }
  return most;
}

static constexpr uint32_t kDefaultTileBudgetWords = 48 * 1024 / 4;

// Tile shape of a plan: `force` when given, else the session's fixed shape,
// else the tuned shape of the plan's bucket, else the largest default whose
// LDS fits 48 KiB (which keeps four workgroups on a compute unit). A fixed
// shape that does not fit is halved along its longer side until it does.
static void choose_tile(const jxrhip_session *s, const jxrhip_plan *p, uint32_t &tx, uint32_t &ty, uint32_t &words,
                        const uint32_t *force = nullptr) {
  if (force) {
    tx = force[0];
    ty = force[1];
    words = tile_words(p, tx, ty);
    return;
  }
  if (!s->tile_x) {
    auto it = s->tuned.find(p->geometry_key);
    if (it != s->tuned.end() && (size_t)tile_words(p, it->second.first, it->second.second) * 4 <= s->max_lds) {
      tx = it->second.first;
      ty = it->second.second;
      words = tile_words(p, tx, ty);
      return;
    }
  }
  if (s->tile_x) {
    tx = s->tile_x;
    ty = s->tile_y;
    while (tile_words(p, tx, ty) > kDefaultTileBudgetWords && (tx > 1 || ty > 1)) {
      if (tx >= ty) tx = (tx + 1) / 2;
      else ty = (ty + 1) / 2;
    }
    words = tile_words(p, tx, ty);
    return;
  }
  static const uint32_t cand[][2] = {{4, 4}, {4, 2}, {2, 2}, {2, 1}, {1, 1}};
  for (auto &c : cand) {
    if (tile_words(p, c[0], c[1]) <= kDefaultTileBudgetWords) {
      tx = c[0];
      ty = c[1];
      words = tile_words(p, tx, ty);
      return;
    }
  }
  tx = ty = 1;
  words = tile_words(p, 1, 1);
}

// One-bit output: the kernel ORs each tile's bits into cleared rows (a byte
// can hold pixels of two tiles), so the rows' bytes are zeroed first; bytes
// between rows are left as they are.
static hipError_t clear_bit_rows(const jxrhip_plan *p, void *dst, hipStream_t stream) {
  if (p->dev.out_kind != kOutBits) return hipSuccess;
  const uint32_t row = (p->dev.crop_w + 7) / 8;
  return launch_clear_rows((uint8_t *)dst + p->dev.plane_offset[0], p->dev.row_stride[0], row, p->dev.crop_h, stream);
}

static jxrhip_status submit_impl(jxrhip_session *s, const jxrhip_plan *p, void *dst, uint64_t dst_bytes,
                                 jxrhip_submission **out) {
  if (!s || !p || !out) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (dst && dst_bytes < p->output_bytes) return fail(JXRHIP_ERR_BUFFER_TOO_SMALL, "destination smaller than the plan's output");
  std::lock_guard<std::mutex> g(s->lock);
  make_current(s);
  std::unique_ptr<jxrhip_submission> sub(new (std::nothrow) jxrhip_submission());
  if (!sub) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "out of host memory");
  sub->session = s;
  sub->layout = p->surface;
  sub->stream = s->streams[s->next_stream++ % kStreams];
  if (s->free_status.empty()) return fail(JXRHIP_ERR_BUDGET, "too many submissions in flight");
  // a failure once work may be queued waits for the stream, then returns
  // the status word, the event and every buffer the submission took
  auto undo = [&](jxrhip_status st) {
    (void)hipStreamSynchronize(sub->stream);
    for (auto &b : sub->transient) s->pool.put(b);
    sub->transient.clear();
    if (sub->output.ptr) s->pool.put(sub->output);
    if (sub->slot != UINT32_MAX) s->free_status.push_back(sub->slot);
    if (sub->done) s->free_events.push_back(sub->done);
    return st;
  };
  hipError_t e = take_event(s, &sub->done);
  if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
  sub->slot = s->free_status.back();
  s->free_status.pop_back();
  uint32_t *status_word = s->status_dev + sub->slot;
  s->status_host[sub->slot] = 0;
  if (!dst) {
    e = s->pool.get(p->output_bytes, sub->output, s->diag.scratch_hits, s->diag.scratch_misses);
    if (e != hipSuccess) return undo(hip_fail(e, "output"));
    dst = sub->output.ptr;
    sub->output_bytes = p->output_bytes;
  }
  const DevImage *dimg;
  jxrhip_status st = device_image(s, p, sub->stream, &dimg, sub.get());
  if (st) return undo(st);
  uint32_t tx, ty, words;
  choose_tile(s, p, tx, ty, words);
  if ((size_t)words * 4 > s->max_lds) return undo(fail(JXRHIP_ERR_INVALID_PLAN, "tile does not fit in LDS"));
  KernelArgs args{};
  args.inline_slot.out = (uint8_t *)dst;
  args.inline_slot.status = status_word;
  args.inline_slot.tiles_x = (p->mbx + tx - 1) / tx;
  args.inline_slot.tile_start = 0;
  args.persistent = 0;
  args.tile_x = tx;
  args.tile_y = ty;
  args.lds_words = words;
  args.storage = p->dev.storage;
  args.reduced = p->scale != JXRHIP_SCALE_FULL ? 1 : 0;
  const uint32_t tiles = args.inline_slot.tiles_x * ((p->mby + ty - 1) / ty);
  args.item_count = tiles;
  if ((e = hipMemsetAsync(status_word, 0, 4, sub->stream)) != hipSuccess) return undo(hip_fail(e, "status reset"));
  if ((e = clear_bit_rows(p, dst, sub->stream)) != hipSuccess) return undo(hip_fail(e, "one-bit rows"));
  // one item per workgroup; a multiple of 8 workgroups lets the kernel give
  // each XCD a contiguous share of the tiles. One launch per plane group.
  for (size_t g = 0; g < p->groups.size(); g++) {
    args.inline_slot.img = dimg + g;
    if ((e = launch_tiles(args, (tiles + 7u) & ~7u, sub->stream)) != hipSuccess)
      return undo(hip_fail(e, "reconstruction launch"));
    s->diag.kernel_launches++;
  }
  if ((e = hipMemcpyAsync(s->status_host + sub->slot, status_word, 4, hipMemcpyDeviceToHost, sub->stream)) !=
      hipSuccess)
    return undo(hip_fail(e, "status readback"));
  if ((e = hipEventRecord(sub->done, sub->stream)) != hipSuccess) return undo(hip_fail(e, "hipEventRecord"));
  s->diag.submissions++;
  session_ref(s);
  *out = sub.release();
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_submit(jxrhip_session *s, const jxrhip_plan *p, jxrhip_submission **out) {
  return submit_impl(s, p, nullptr, 0, out);
}
extern "C" jxrhip_status jxrhip_submit_into(jxrhip_session *s, const jxrhip_plan *p, void *dst, uint64_t bytes,
                                            jxrhip_submission **out) {
  if (!dst) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null destination");
  return submit_impl(s, p, dst, bytes, out);
}

extern "C" int32_t jxrhip_submission_is_complete(const jxrhip_submission *sub) {
  if (!sub) return 1;
  return hipEventQuery(sub->done) == hipSuccess ? 1 : 0;
}

static void release_submission(jxrhip_submission *sub) {
  jxrhip_session *s = sub->session;
  {
    std::lock_guard<std::mutex> g(s->lock);
    make_current(s);
    for (auto &b : sub->transient) s->pool.put(b);
    sub->transient.clear();
    if (sub->output.ptr) s->pool.put(sub->output);
    if (sub->slot != UINT32_MAX) s->free_status.push_back(sub->slot);
    if (sub->done) s->free_events.push_back(sub->done);
  }
  delete sub;
  session_unref(s);
}

extern "C" jxrhip_status jxrhip_submission_wait(jxrhip_submission *sub, jxrhip_image **image, uint32_t *stage) {
  if (!sub) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null submission");
  if (image) *image = nullptr;
  if (stage) *stage = 0;
  make_current(sub->session);
  hipError_t e = hipEventSynchronize(sub->done);
  if (e != hipSuccess) {
    release_submission(sub);
    return hip_fail(e, "hipEventSynchronize");
  }
  const uint32_t status = sub->session->status_host[sub->slot];
  if (status != 0) {
    if (stage) *stage = status;
    release_submission(sub);
    return fail(JXRHIP_ERR_KERNEL_ARITHMETIC, "reconstruction arithmetic overflow (stage " + std::to_string(status) + ")");
  }
  if (image && sub->output.ptr) {
    jxrhip_image *im = new (std::nothrow) jxrhip_image();
    if (!im) {
      release_submission(sub);
      return fail(JXRHIP_ERR_INVALID_ARGUMENT, "out of host memory");
    }
    im->session = sub->session;
    im->buf = sub->output;
    im->bytes = sub->output_bytes;
    sub->output = DeviceBuffer{};
    session_ref(im->session);
    *image = im;
  }
  release_submission(sub);
  return JXRHIP_OK;
}

extern "C" void jxrhip_submission_destroy(jxrhip_submission *sub) {
  if (!sub) return;
  make_current(sub->session);
  (void)hipEventSynchronize(sub->done);
  release_submission(sub);
}

// ---- consumers of a pending submission (submission.rs:50-93) ---------------
extern "C" jxrhip_status jxrhip_submission_enqueue_wait(const jxrhip_submission *sub, void *stream) {
  if (!sub) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null submission");
  make_current(sub->session);
  HIP_TRY(hipStreamWaitEvent(static_cast<hipStream_t>(stream), sub->done, 0), "hipStreamWaitEvent");
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_submission_output_layout(const jxrhip_submission *sub, jxrhip_surface_layout *out) {
  if (!sub || !out) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  *out = sub->layout;
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_submission_pending_device_ptr(const jxrhip_submission *sub, void **ptr,
                                                              uint64_t *bytes) {
  if (!sub || !ptr) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  *ptr = nullptr;
  if (bytes) *bytes = 0;
  if (!sub->output.ptr)
    return fail(JXRHIP_ERR_INVALID_ARGUMENT, "the submission writes a caller-owned destination");
  *ptr = sub->output.ptr;
  if (bytes) *bytes = sub->output_bytes;
  return JXRHIP_OK;
}

// DecodeReport of a device reconstruction (crates/jxr-cuda/src/plan.rs:
// 214-253): the stages before reconstruction ran on the CPU, the eight
// reconstruction stages on the device, then an optional host readback.
extern "C" jxrhip_status jxrhip_plan_decode_report(const jxrhip_plan *p, uint32_t requested, int32_t host_readback,
                                                   jxrhip_decode_report *out) {
  if (!p || !out) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  std::memset(out, 0, sizeof *out);
  out->requested = requested;
  out->selected = JXRHIP_BACKEND_HIP;
  out->fallback = JXRHIP_FALLBACK_NONE;
  uint32_t n = 0;
  for (uint32_t st = JXRHIP_DECODE_STAGE_PARSE; st <= JXRHIP_DECODE_STAGE_DC_LOW_PASS_PREDICTION; st++)
    out->stages[n++] = jxrhip_stage_report{st, JXRHIP_EXECUTOR_CPU_SCALAR};
  for (uint32_t st = JXRHIP_DECODE_STAGE_DEQUANTIZE_AND_FIRST_INVERSE_TRANSFORM;
       st <= JXRHIP_DECODE_STAGE_CROP_CLIP_AND_PACK; st++)
    out->stages[n++] = jxrhip_stage_report{st, JXRHIP_EXECUTOR_HIP};
  if (host_readback) out->stages[n++] = jxrhip_stage_report{JXRHIP_DECODE_STAGE_HOST_READBACK, JXRHIP_EXECUTOR_CPU_SCALAR};
  out->stage_count = n;
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_batch_enqueue_wait(const jxrhip_batch *b, void *stream);

extern "C" jxrhip_status jxrhip_host_alloc(jxrhip_session *s, uint64_t bytes, void **ptr) {
  if (!s || !ptr) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  *ptr = nullptr;
  make_current(s);
  HIP_TRY(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault), "hipHostMalloc");
  {
    std::lock_guard<std::mutex> g(s->lock);
    s->host_blocks.insert(*ptr);
  }
  session_ref(s);  // a block keeps the session, as a destination does
  return JXRHIP_OK;
}
extern "C" jxrhip_status jxrhip_host_free(jxrhip_session *s, void *ptr) {
  if (!s) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null session");
  if (!ptr) return JXRHIP_OK;
  {
    std::lock_guard<std::mutex> g(s->lock);
    auto it = s->host_blocks.find(ptr);
    if (it == s->host_blocks.end()) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "not a host block of this session");
    s->host_blocks.erase(it);
  }
  make_current(s);
  // uploads from the block may still be in flight on the session's streams
  jxrhip_status st = JXRHIP_OK;
  for (uint32_t i = 0; i < kStreams && !st; i++) {
    const hipError_t e = hipStreamSynchronize(s->streams[i]);
    if (e != hipSuccess) st = hip_fail(e, "hipStreamSynchronize");
  }
  const hipError_t e = hipHostFree(ptr);
  if (e != hipSuccess && !st) st = hip_fail(e, "hipHostFree");
  session_unref(s);
  return st;
}

extern "C" jxrhip_status jxrhip_allocate_destination(jxrhip_session *s, uint64_t bytes, void **ptr) {
  if (!s || !ptr) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  *ptr = nullptr;
  make_current(s);
  HIP_TRY(hipMalloc(ptr, bytes ? bytes : 1), "hipMalloc");
  {
    std::lock_guard<std::mutex> g(s->lock);
    s->destinations[*ptr] = bytes;
  }
  session_ref(s);
  return JXRHIP_OK;
}
extern "C" jxrhip_status jxrhip_free_destination(jxrhip_session *s, void *ptr) {
  if (!s) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null session");
  if (!ptr) return JXRHIP_OK;
  make_current(s);
  {
    std::lock_guard<std::mutex> g(s->lock);
    auto it = s->destinations.find(ptr);
    if (it == s->destinations.end())
      return fail(JXRHIP_ERR_INVALID_ARGUMENT, "not a destination allocated by this session");
    s->destinations.erase(it);
  }
  jxrhip_status st = JXRHIP_OK;
  for (uint32_t i = 0; i < kStreams && !st; i++) {
    const hipError_t e = hipStreamSynchronize(s->streams[i]);
    if (e != hipSuccess) st = hip_fail(e, "hipStreamSynchronize");
  }
  const hipError_t e = hipFree(ptr);
  if (e != hipSuccess && !st) st = hip_fail(e, "hipFree");
  session_unref(s);
  return st;
}
extern "C" jxrhip_status jxrhip_fill_destination(jxrhip_session *s, void *ptr, uint8_t value, uint64_t bytes) {
  if (!s || !ptr) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  make_current(s);
  HIP_TRY(hipMemsetAsync(ptr, value, bytes, s->streams[0]), "hipMemsetAsync");
  HIP_TRY(hipStreamSynchronize(s->streams[0]), "hipStreamSynchronize");
  return JXRHIP_OK;
}
extern "C" jxrhip_status jxrhip_copy_to_host(jxrhip_session *s, const void *ptr, void *host, uint64_t bytes) {
  if (!s || !ptr || !host) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  make_current(s);
  for (uint32_t i = 0; i < kStreams; i++) HIP_TRY(hipStreamSynchronize(s->streams[i]), "hipStreamSynchronize");
  HIP_TRY(hipMemcpy(host, ptr, bytes, hipMemcpyDeviceToHost), "hipMemcpy");
  return JXRHIP_OK;
}
extern "C" jxrhip_status jxrhip_copy_device(jxrhip_session *s, void *dst, const void *src, uint64_t bytes) {
  if (!s || !dst || !src) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  make_current(s);
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s->streams[0]), "hipMemcpyAsync");
  HIP_TRY(hipStreamSynchronize(s->streams[0]), "hipStreamSynchronize");
  return JXRHIP_OK;
}
extern "C" jxrhip_status jxrhip_calibrate_valu(jxrhip_session *s, uint32_t iterations, jxrhip_valu_rate *out) {
  if (!s || !out || iterations == 0) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument or no iterations");
  make_current(s);
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, s->device), "hipGetDeviceProperties");
  const uint32_t groups = s->cus * 4;  // sixteen waves per compute unit, as the reconstruction kernel runs
  hipEvent_t a = nullptr, b = nullptr;
  DeviceBuffer buf;
  float ms = 0.f;
  hipError_t e;
  {
    std::lock_guard<std::mutex> g(s->lock);
    e = s->pool.get((uint64_t)groups * kThreads * 4, buf, s->diag.scratch_hits, s->diag.scratch_misses);
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    hipStream_t st = s->streams[0];
    if (e == hipSuccess) e = launch_valu_chains((uint32_t *)buf.ptr, groups, 64, st);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(a, st);
    if (e == hipSuccess) e = launch_valu_chains((uint32_t *)buf.ptr, groups, iterations, st);
    if (e == hipSuccess) e = hipEventRecord(b, st);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (buf.ptr) s->pool.put(buf);
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (e != hipSuccess) return hip_fail(e, "VALU calibration");
  out->ms = ms;
  out->wave_instructions = (double)groups * (kThreads / 64) * iterations * valu_per_iteration();
  out->simds = s->cus * 4;
  out->clock_khz = (uint32_t)prop.clockRate;
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_session_synchronize(jxrhip_session *s) {
  if (!s) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null session");
  make_current(s);
  for (uint32_t i = 0; i < kStreams; i++) HIP_TRY(hipStreamSynchronize(s->streams[i]), "hipStreamSynchronize");
  return JXRHIP_OK;
}

extern "C" void *jxrhip_image_device_ptr(const jxrhip_image *im) { return im ? im->buf.ptr : nullptr; }
extern "C" uint64_t jxrhip_image_bytes(const jxrhip_image *im) { return im ? im->bytes : 0; }
extern "C" void jxrhip_image_destroy(jxrhip_image *im) {
  if (!im) return;
  jxrhip_session *s = im->session;
  if (im->held) im->held->fetch_sub(im->buf.bytes);
  {
    std::lock_guard<std::mutex> g(s->lock);
    make_current(s);
    s->pool.put(im->buf);
  }
  delete im;
  session_unref(s);
}

extern "C" jxrhip_status jxrhip_readback(jxrhip_session *s, const jxrhip_image *im, void *host, uint64_t bytes) {
  if (!s || !im || !host) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  if (bytes < im->bytes) return fail(JXRHIP_ERR_BUFFER_TOO_SMALL, "host buffer smaller than the image");
  make_current(s);
  HIP_TRY(hipMemcpy(host, im->buf.ptr, im->bytes, hipMemcpyDeviceToHost), "readback");
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_decode_to_host(jxrhip_session *s, const jxrhip_plan *p, void *host, uint64_t bytes,
                                               uint32_t *stage) {
  if (stage) *stage = 0;
  if (!p || bytes < p->output_bytes) return fail(JXRHIP_ERR_BUFFER_TOO_SMALL, "host buffer smaller than the output");
  jxrhip_submission *sub;
  jxrhip_status st = jxrhip_submit(s, p, &sub);
  if (st) return st;
  jxrhip_image *im = nullptr;
  st = jxrhip_submission_wait(sub, &im, stage);
  if (st) return st;
  st = jxrhip_readback(s, im, host, bytes);
  jxrhip_image_destroy(im);
  return st;
}

// ---- persistent batches --------------------------------------------------
// One persistent launch per shape bucket. A prepared batch owns what those
// launches read besides the images' descriptors and coefficients: the slot
// records, the owning slot of every work item, the share counters and the
// per-image status words.
struct BucketLaunch {
  size_t first;  // first slot of the bucket
  uint32_t n, items, tx, ty, words, storage, grid;
  uint32_t reduced = 0;  // the bucket's images are reduced decodes
};

struct PreparedBatch {
  std::vector<BucketLaunch> launches;
  std::vector<const jxrhip_plan *> plans;
  std::vector<void *> dsts;
  DeviceBuffer slots, item_slots, counters, statuses;
  std::vector<DeviceBuffer> transient;  // descriptors and coefficients of arenas without identity
  std::vector<uint64_t> coef_pins, desc_pins;  // cache entries held resident until release
};

static void pin_uploads(jxrhip_session *s, const jxrhip_plan *p, PreparedBatch &pb) {
  for (uint32_t a = 0; a < p->arena_count; a++) {
    auto it = s->coef_cache.find(p->arenas[a]->identity);
    if (p->arenas[a]->identity && it != s->coef_cache.end()) {
      it->second.pins++;
      pb.coef_pins.push_back(it->first);
    }
  }
  auto it = s->desc_cache.find(p->id);
  if (it != s->desc_cache.end()) {
    it->second.pins++;
    pb.desc_pins.push_back(it->first);
  }
}

static void release_prepared(jxrhip_session *s, PreparedBatch &pb) {
  for (uint64_t k : pb.coef_pins) {
    auto it = s->coef_cache.find(k);
    if (it != s->coef_cache.end() && it->second.pins) it->second.pins--;
  }
  for (uint64_t k : pb.desc_pins) {
    auto it = s->desc_cache.find(k);
    if (it != s->desc_cache.end() && it->second.pins) it->second.pins--;
  }
  pb.coef_pins.clear();
  pb.desc_pins.clear();
  for (auto &t : pb.transient) s->pool.put(t);
  pb.transient.clear();
  s->pool.put(pb.slots);
  s->pool.put(pb.item_slots);
  s->pool.put(pb.counters);
  s->pool.put(pb.statuses);
}

// Groups the plans into shape buckets, uploads what is missing and builds the
// slot records. `force` (tile_x, tile_y) overrides every other tile choice.
// Called with the session lock held; on failure the caller releases `pb`.
static jxrhip_status prepare_persistent(jxrhip_session *s, const jxrhip_plan *const *plans, uint32_t count,
                                        void *const *dsts, const uint32_t *force, PreparedBatch &pb,
                                        hipStream_t upload_stream = nullptr, Slot *pinned_slots = nullptr,
                                        size_t pinned_capacity = 0, uint32_t force_groups = 0) {
  hipStream_t stream = upload_stream ? upload_stream : s->streams[0];
  pb.plans.assign(plans, plans + count);
  pb.dsts.assign(dsts, dsts + count);
  HIP_TRY(s->pool.get(sizeof(uint32_t) * std::max<uint32_t>(count, 1), pb.statuses, s->diag.scratch_hits,
                      s->diag.scratch_misses),
          "status words");
  // shape buckets: plans with the same plane geometry and storage share a
  // kernel instantiation and a tile shape
  std::map<uint64_t, std::vector<uint32_t>> buckets;
  for (uint32_t i = 0; i < count; i++) buckets[plans[i]->geometry_key].push_back(i);
  std::vector<const DevImage *> dimgs(count);
  for (uint32_t i = 0; i < count; i++) {
    jxrhip_submission tmp;
    tmp.session = s;
    jxrhip_status st = device_image(s, plans[i], stream, &dimgs[i], &tmp);
    for (auto &t : tmp.transient) pb.transient.push_back(t);
    if (st) return st;
    // later uploads of this batch must not evict what it already placed
    pin_uploads(s, plans[i], pb);
  }
  std::vector<Slot> slots;
  for (auto &kv : buckets) {
    uint32_t tx, ty, words;
    choose_tile(s, plans[kv.second[0]], tx, ty, words, force);
    // the bucket shares a tile shape; its images' halos (and so their LDS)
    // can differ with the overlap mode
    for (uint32_t i : kv.second) words = std::max(words, tile_words(plans[i], tx, ty));
    if ((size_t)words * 4 > s->max_lds) return fail(JXRHIP_ERR_INVALID_PLAN, "tile does not fit in LDS");
    BucketLaunch L{slots.size(), 0, 0, tx, ty, words, plans[kv.second[0]]->dev.storage, 0};
    L.reduced = plans[kv.second[0]]->scale != JXRHIP_SCALE_FULL ? 1 : 0;
    for (uint32_t i : kv.second) {
      // one slot per plane group; the groups of an image share its status word
      for (size_t g = 0; g < plans[i]->groups.size(); g++) {
        Slot sl{};
        sl.img = dimgs[i] + g;
        sl.out = (uint8_t *)dsts[i];
        sl.status = (uint32_t *)pb.statuses.ptr + i;
        sl.tiles_x = (plans[i]->mbx + tx - 1) / tx;
        sl.tile_start = L.items;
        const uint64_t items = (uint64_t)L.items + (uint64_t)sl.tiles_x * ((plans[i]->mby + ty - 1) / ty);
        if (items > UINT32_MAX / 2) return fail(JXRHIP_ERR_BUDGET, "too many work items in one shape bucket");
        L.items = (uint32_t)items;
        slots.push_back(sl);
      }
    }
    L.n = (uint32_t)(slots.size() - L.first);
    // workgroups a compute unit holds at once (LDS, registers), unless the
    // session fixes it; a multiple of 8 workgroups for the XCD split
    int resident = 0;
    if (!s->groups_per_cu)
      HIP_TRY(kernel_occupancy(L.storage, L.reduced, (size_t)L.words * 4, &resident), "occupancy");
    uint32_t per_cu = s->groups_per_cu ? s->groups_per_cu : (uint32_t)std::max(resident, 1);
    if (!s->groups_per_cu) {
      // fewer than fit when the autotuner found that faster for the bucket
      uint32_t want = force_groups;
      if (!want && !force) {
        auto tg = s->tuned_groups.find(kv.first);
        if (tg != s->tuned_groups.end()) want = tg->second;
      }
      if (want) per_cu = std::min(per_cu, want);
    }
    uint32_t grid = std::min<uint32_t>(L.items, s->cus * per_cu);
    L.grid = std::max<uint32_t>((grid + 7u) & ~7u, 8u);
    pb.launches.push_back(L);
  }
  HIP_TRY(s->pool.get(std::max<size_t>(slots.size(), 1) * sizeof(Slot), pb.slots, s->diag.scratch_hits,
                      s->diag.scratch_misses),
          "slot buffer");
  if (pinned_slots && slots.size() <= pinned_capacity) {
    // pinned staging of the caller's, left alone until the stream passes
    // this copy: the copy is a plain DMA and the host does not wait for it
    std::copy(slots.begin(), slots.end(), pinned_slots);
    HIP_TRY(hipMemcpyAsync(pb.slots.ptr, pinned_slots, slots.size() * sizeof(Slot), hipMemcpyHostToDevice, stream),
            "slot upload");
  } else {
    HIP_TRY(hipMemcpyAsync(pb.slots.ptr, slots.data(), slots.size() * sizeof(Slot), hipMemcpyHostToDevice, stream),
            "slot upload");
    // the host records are staged before this returns
    HIP_TRY(hipStreamSynchronize(stream), "slot upload");
  }
  uint64_t total_items = 0;
  for (const BucketLaunch &L : pb.launches) total_items += L.items;
  HIP_TRY(s->pool.get(std::max<uint64_t>(total_items, 1) * sizeof(uint32_t), pb.item_slots, s->diag.scratch_hits,
                      s->diag.scratch_misses),
          "item table");
  // eight share counters per launch, each on its own 128-byte line
  HIP_TRY(s->pool.get(std::max<size_t>(pb.launches.size(), 1) * 8 * 128, pb.counters, s->diag.scratch_hits,
                      s->diag.scratch_misses),
          "share counters");
  return JXRHIP_OK;
}

// Enqueues a prepared batch behind the work already on streams[0] and joins
// it back there: one-bit rows are cleared, then each bucket runs as an
// item-table kernel (which also zeroes the share counters and the bucket's
// status words) and a persistent launch. When every bucket holds at least two
// items per workgroup, each launch fills the device and the buckets run one
// after another on streams[0]; otherwise they go round-robin over the
// session's streams, so that small buckets run side by side. Only kernel
// launches and event dependencies, so that the sequence can be captured into
// a graph. With `only`, everything goes on that one stream instead.
static jxrhip_status enqueue_persistent(jxrhip_session *s, const PreparedBatch &pb, hipStream_t only = nullptr) {
  hipStream_t s0 = only ? only : s->streams[0];
  for (size_t i = 0; i < pb.plans.size(); i++) HIP_TRY(clear_bit_rows(pb.plans[i], pb.dsts[i], s0), "one-bit rows");
  bool spread = false;
  for (const BucketLaunch &L : pb.launches) spread |= L.items < 2 * L.grid;
  const uint32_t used = spread && !only ? (uint32_t)std::min<size_t>(pb.launches.size(), kStreams) : 1;
  if (used > 1) HIP_TRY(hipEventRecord(s->fork, s0), "hipEventRecord");
  for (uint32_t i = 1; i < used; i++) HIP_TRY(hipStreamWaitEvent(s->streams[i], s->fork, 0), "hipStreamWaitEvent");
  uint64_t item_base = 0;
  for (size_t li = 0; li < pb.launches.size(); li++) {
    const BucketLaunch &L = pb.launches[li];
    hipStream_t ls = used == 1 ? s0 : s->streams[li % used];
    KernelArgs args{};
    args.slots = (const Slot *)pb.slots.ptr + L.first;
    args.item_slot = (const uint32_t *)pb.item_slots.ptr + item_base;
    args.counters = (uint32_t *)((uint8_t *)pb.counters.ptr + li * 8 * 128);
    item_base += L.items;
    args.slot_count = L.n;
    args.item_count = L.items;
    args.persistent = 1;
    args.tile_x = L.tx;
    args.tile_y = L.ty;
    args.lds_words = L.words;
    args.storage = L.storage;
    args.reduced = L.reduced;
    HIP_TRY(launch_expand(args.slots, L.n, L.items, (uint32_t *)args.item_slot, args.counters, ls), "item table");
    HIP_TRY(launch_tiles(args, L.grid, ls), "persistent launch");
    s->diag.kernel_launches++;
  }
  for (uint32_t i = 1; i < used; i++) {
    HIP_TRY(hipEventRecord(s->join[i], s->streams[i]), "hipEventRecord");
    HIP_TRY(hipStreamWaitEvent(s0, s->join[i], 0), "hipStreamWaitEvent");
  }
  return JXRHIP_OK;
}

struct jxrhip_batch_graph {
  jxrhip_session *session = nullptr;
  PreparedBatch pb;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t *host_statuses = nullptr;
  hipEvent_t done = nullptr;
  bool in_flight = false;
};

// ---- batches ------------------------------------------------------------
struct jxrhip_batch {
  jxrhip_session *session = nullptr;
  uint32_t count = 0;
  std::vector<jxrhip_submission *> subs;  // per-image mode
  PreparedBatch pb;                       // persistent mode
  uint32_t *host_statuses = nullptr;
  hipEvent_t done = nullptr;
  jxrhip_batch_graph *graph = nullptr;  // replay of a captured batch: owns nothing
  bool timed = false;
};

extern "C" jxrhip_status jxrhip_submit_batch(jxrhip_session *s, const jxrhip_plan *const *plans, uint32_t count,
                                             void *const *dsts, uint32_t mode, jxrhip_batch **out) {
  if (!s || !plans || !dsts || !out) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  for (uint32_t i = 0; i < count; i++)
    if (!plans[i] || !dsts[i]) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null plan or destination");
  std::unique_ptr<jxrhip_batch> b(new (std::nothrow) jxrhip_batch());
  if (!b) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "out of host memory");
  b->session = s;
  b->count = count;
  make_current(s);
  if (mode == JXRHIP_BATCH_PER_IMAGE) {
    // Upload everything first, so that the timed window holds launches only.
    {
      std::lock_guard<std::mutex> g(s->lock);
      HIP_TRY(hipEventRecord(s->t0, s->streams[0]), "hipEventRecord");
      for (uint32_t i = 1; i < kStreams; i++) HIP_TRY(hipStreamWaitEvent(s->streams[i], s->t0, 0), "hipStreamWaitEvent");
      s->next_stream = 0;
    }
    for (uint32_t i = 0; i < count; i++) {
      jxrhip_submission *sub = nullptr;
      jxrhip_status st = submit_impl(s, plans[i], dsts[i], plans[i]->output_bytes, &sub);
      if (st) {
        for (auto *x : b->subs) jxrhip_submission_destroy(x);
        return st;
      }
      b->subs.push_back(sub);
    }
    // the end event on streams[0] waits for the other streams' images
    hipError_t e = hipSuccess;
    {
      std::lock_guard<std::mutex> g(s->lock);
      for (uint32_t i = 1; i < kStreams && e == hipSuccess; i++) {
        e = hipEventRecord(s->join[i], s->streams[i]);
        if (e == hipSuccess) e = hipStreamWaitEvent(s->streams[0], s->join[i], 0);
      }
      if (e == hipSuccess) e = hipEventRecord(s->t1, s->streams[0]);
      if (e == hipSuccess) s->last_launches = count;
    }
    if (e != hipSuccess) {
      const jxrhip_status st = hip_fail(e, "batch end event");
      for (auto *x : b->subs) jxrhip_submission_destroy(x);
      return st;
    }
    b->timed = true;
    session_ref(s);
    *out = b.release();
    return JXRHIP_OK;
  }
  if (mode != JXRHIP_BATCH_PERSISTENT) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "batch mode");
  std::lock_guard<std::mutex> g(s->lock);
  hipStream_t stream = s->streams[0];
  // every failure once the pinned block exists: wait for the stream, then
  // release the prepared batch (pool buffers and cache pins), the pinned
  // block and the completion event
  auto undo = [&](jxrhip_status st) {
    (void)hipStreamSynchronize(stream);
    release_prepared(s, b->pb);
    if (b->host_statuses) (void)hipHostFree(b->host_statuses);
    b->host_statuses = nullptr;
    if (b->done) (void)hipEventDestroy(b->done);
    b->done = nullptr;
    return st;
  };
  // one pinned block: the status words, then the staging of the slot table,
  // which the batch keeps (as it keeps the statuses) until its wait
  const size_t status_bytes = (sizeof(uint32_t) * std::max<uint32_t>(count, 1) + 255) & ~size_t(255);
  size_t slot_count = 0;  // one slot per plane group
  for (uint32_t i = 0; i < count; i++) slot_count += plans[i]->groups.size();
  slot_count = std::max<size_t>(slot_count, 1);
  if (hipHostMalloc((void **)&b->host_statuses, status_bytes + sizeof(Slot) * slot_count, hipHostMallocDefault) !=
      hipSuccess)
    return undo(fail(JXRHIP_ERR_DEVICE, "hipHostMalloc"));
  Slot *slot_staging = reinterpret_cast<Slot *>(reinterpret_cast<uint8_t *>(b->host_statuses) + status_bytes);
  jxrhip_status st = prepare_persistent(s, plans, count, dsts, nullptr, b->pb, nullptr, slot_staging, slot_count);
  if (st) return undo(st);
  if (hipEventRecord(s->t0, stream) != hipSuccess) return undo(fail(JXRHIP_ERR_DEVICE, "hipEventRecord"));
  st = enqueue_persistent(s, b->pb);
  if (st) return undo(st);
  hipError_t e = hipEventRecord(s->t1, stream);
  if (e != hipSuccess) return undo(hip_fail(e, "hipEventRecord"));
  e = hipMemcpyAsync(b->host_statuses, b->pb.statuses.ptr, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, stream);
  if (e != hipSuccess) return undo(hip_fail(e, "status readback"));
  e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
  if (e != hipSuccess) return undo(hip_fail(e, "hipEventCreate"));
  e = hipEventRecord(b->done, stream);
  if (e != hipSuccess) return undo(hip_fail(e, "hipEventRecord"));
  b->timed = true;
  s->last_launches = (uint32_t)b->pb.launches.size();
  s->diag.submissions += count;
  session_ref(s);
  *out = b.release();
  return JXRHIP_OK;
}

extern "C" int32_t jxrhip_batch_is_complete(const jxrhip_batch *b) {
  if (!b) return 1;
  if (b->done) return hipEventQuery(b->done) == hipSuccess ? 1 : 0;
  for (auto *sub : b->subs)
    if (!jxrhip_submission_is_complete(sub)) return 0;
  return 1;
}

// A device-side wait on `stream` for every image of the batch.
extern "C" jxrhip_status jxrhip_batch_enqueue_wait(const jxrhip_batch *b, void *stream) {
  if (!b) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null batch");
  make_current(b->session);
  hipStream_t cs = static_cast<hipStream_t>(stream);
  if (b->done) {
    HIP_TRY(hipStreamWaitEvent(cs, b->done, 0), "hipStreamWaitEvent");
    return JXRHIP_OK;
  }
  for (const jxrhip_submission *sub : b->subs) HIP_TRY(hipStreamWaitEvent(cs, sub->done, 0), "hipStreamWaitEvent");
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_batch_wait(jxrhip_batch *b, uint32_t *stages) {
  if (!b) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null batch");
  jxrhip_session *s = b->session;
  make_current(s);
  jxrhip_status result = JXRHIP_OK;
  std::string first_error;
  if (b->done) {
    hipError_t e = hipEventSynchronize(b->done);
    if (e != hipSuccess) result = hip_fail(e, "hipEventSynchronize");
    for (uint32_t i = 0; i < b->count; i++) {
      const uint32_t st = result == JXRHIP_OK ? b->host_statuses[i] : 0;
      if (stages) stages[i] = st;
      if (st && result == JXRHIP_OK) {
        result = JXRHIP_ERR_KERNEL_ARITHMETIC;
        first_error = "image " + std::to_string(i) + ": reconstruction arithmetic overflow (stage " + std::to_string(st) + ")";
      }
    }
  } else {
    for (uint32_t i = 0; i < b->subs.size(); i++) {
      uint32_t st = 0;
      jxrhip_status r = jxrhip_submission_wait(b->subs[i], nullptr, &st);
      if (stages) stages[i] = st;
      if (r && result == JXRHIP_OK) {
        result = r;
        first_error = "image " + std::to_string(i) + ": " + g_error;
      }
    }
    b->subs.clear();
  }
  if (b->timed) {
    float ms = 0;
    (void)hipEventSynchronize(s->t1);
    if (hipEventElapsedTime(&ms, s->t0, s->t1) == hipSuccess) s->last_ms = ms;
  }
  if (b->graph) {
    std::lock_guard<std::mutex> g(s->lock);
    b->graph->in_flight = false;
  } else {
    {
      std::lock_guard<std::mutex> g(s->lock);
      release_prepared(s, b->pb);
    }
    if (b->host_statuses) (void)hipHostFree(b->host_statuses);
    if (b->done) (void)hipEventDestroy(b->done);
  }
  delete b;
  session_unref(s);
  if (result) g_error = first_error.empty() ? g_error : first_error;
  return result;
}

extern "C" jxrhip_status jxrhip_last_batch_timing(jxrhip_session *s, double *ms, uint32_t *launches) {
  if (!s) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null session");
  if (ms) *ms = s->last_ms;
  if (launches) *launches = s->last_launches;
  return JXRHIP_OK;
}

// ---- allocation-free batches -------------------------------------------------
// The device descriptor of a plan whose coefficients and descriptors are in
// the session's caches already; allocates nothing, and fails when any of
// them is missing.
static jxrhip_status resident_image(jxrhip_session *s, const jxrhip_plan *p, hipStream_t stream,
                                    const DevImage **out) {
  const void *coef[2] = {nullptr, nullptr};
  for (uint32_t a = 0; a < p->arena_count; a++) {
    const uint64_t id = p->arenas[a]->identity;
    auto it = id ? s->coef_cache.find(id) : s->coef_cache.end();
    if (it == s->coef_cache.end())
      return fail(JXRHIP_ERR_INVALID_ARGUMENT, "coefficients are not resident (jxrhip_session_upload first)");
    HIP_TRY(await_upload(s, it->second, stream), "upload dependency");
    it->second.last_use = ++s->tick;
    coef[a] = it->second.buf.ptr;
  }
  auto it = s->desc_cache.find(p->id);
  if (it == s->desc_cache.end() || it->second.deps[0] != coef[0] || it->second.deps[1] != coef[1])
    return fail(JXRHIP_ERR_INVALID_ARGUMENT, "descriptors are not resident (jxrhip_session_upload first)");
  HIP_TRY(await_upload(s, it->second, stream), "descriptor dependency");
  it->second.last_use = ++s->tick;
  *out = (const DevImage *)it->second.buf.ptr;
  return JXRHIP_OK;
}

// Host staging of one allocation-free call: fixed arrays, so the call
// allocates nothing on the host either. One call at a time in the process.
namespace {
struct StaticStaging {
  std::mutex m;
  const DevImage *dimgs[kStaticMaxImages];
  uint32_t bucket_of[kStaticMaxImages];
  uint64_t keys[kStaticMaxLaunches];
  Slot slots[kStaticMaxSlots];
  uint32_t statuses[kStaticMaxImages];
};
StaticStaging g_static;
}  // namespace

extern "C" jxrhip_status jxrhip_decode_batch_static(jxrhip_session *s, const jxrhip_plan *const *plans,
                                                    uint32_t count, void *const *dsts, uint32_t *stages) {
  if (!s || (count && (!plans || !dsts))) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  if (count > kStaticMaxImages)
    return fail(JXRHIP_ERR_BUDGET, "more than " + std::to_string(kStaticMaxImages) +
                                       " images in one allocation-free batch");
  for (uint32_t i = 0; i < count; i++)
    if (!plans[i] || !dsts[i]) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null plan or destination");
  if (count == 0) return JXRHIP_OK;
  std::lock_guard<std::mutex> host(g_static.m);
  std::lock_guard<std::mutex> g(s->lock);
  make_current(s);
  if (!s->scratch) return fail(JXRHIP_ERR_DEVICE, "no static scratch on this device");
  hipStream_t s0 = s->streams[0];
  StaticStaging &st = g_static;
  // shape buckets, in order of first appearance
  uint32_t nb = 0;
  size_t nslots = 0;
  for (uint32_t i = 0; i < count; i++) {
    jxrhip_status r = resident_image(s, plans[i], s0, &st.dimgs[i]);
    if (r) return r;
    uint32_t b = 0;
    while (b < nb && st.keys[b] != plans[i]->geometry_key) b++;
    if (b == nb) {
      if (nb == kStaticMaxLaunches)
        return fail(JXRHIP_ERR_BUDGET, "more than " + std::to_string(kStaticMaxLaunches) +
                                           " shape buckets in one allocation-free batch");
      st.keys[nb++] = plans[i]->geometry_key;
    }
    st.bucket_of[i] = b;
    nslots += plans[i]->groups.size();
  }
  if (nslots > kStaticMaxSlots)
    return fail(JXRHIP_ERR_BUDGET, "more than " + std::to_string(kStaticMaxSlots) +
                                       " plane groups in one allocation-free batch");
  uint32_t *status_dev = reinterpret_cast<uint32_t *>(s->scratch + kStaticStatusOffset);
  Slot *slot_dev = reinterpret_cast<Slot *>(s->scratch + kStaticSlotOffset);
  uint32_t *item_dev = reinterpret_cast<uint32_t *>(s->scratch + kStaticItemOffset);
  BucketLaunch launches[kStaticMaxLaunches];
  size_t next_slot = 0;
  uint64_t total_items = 0;
  for (uint32_t b = 0; b < nb; b++) {
    const jxrhip_plan *first = nullptr;
    for (uint32_t i = 0; i < count && !first; i++)
      if (st.bucket_of[i] == b) first = plans[i];
    uint32_t tx, ty, words;
    choose_tile(s, first, tx, ty, words);
    for (uint32_t i = 0; i < count; i++)
      if (st.bucket_of[i] == b) words = std::max(words, tile_words(plans[i], tx, ty));
    if ((size_t)words * 4 > s->max_lds) return fail(JXRHIP_ERR_INVALID_PLAN, "tile does not fit in LDS");
    BucketLaunch L{next_slot, 0, 0, tx, ty, words, first->dev.storage, 0};
    L.reduced = first->scale != JXRHIP_SCALE_FULL ? 1 : 0;
    for (uint32_t i = 0; i < count; i++) {
      if (st.bucket_of[i] != b) continue;
      for (size_t gi = 0; gi < plans[i]->groups.size(); gi++) {
        Slot &sl = st.slots[next_slot++];
        sl = Slot{};
        sl.img = st.dimgs[i] + gi;
        sl.out = (uint8_t *)dsts[i];
        sl.status = status_dev + i;
        sl.tiles_x = (plans[i]->mbx + tx - 1) / tx;
        sl.tile_start = L.items;
        L.items += sl.tiles_x * ((plans[i]->mby + ty - 1) / ty);
        if (total_items + L.items > kStaticMaxItems)
          return fail(JXRHIP_ERR_BUDGET, "more than " + std::to_string(kStaticMaxItems) +
                                             " work items in one allocation-free batch");
      }
    }
    L.n = (uint32_t)(next_slot - L.first);
    int resident = 0;
    if (!s->groups_per_cu)
      HIP_TRY(kernel_occupancy(L.storage, L.reduced, (size_t)L.words * 4, &resident), "occupancy");
    const uint32_t per_cu = s->groups_per_cu ? s->groups_per_cu : (uint32_t)std::max(resident, 1);
    L.grid = std::max<uint32_t>((std::min<uint32_t>(L.items, s->cus * per_cu) + 7u) & ~7u, 8u);
    total_items += L.items;
    launches[b] = L;
  }
  // the slot table goes up from the host staging; everything after it is
  // kernels on the same stream, and the call waits for all of it
  HIP_TRY(hipMemcpyAsync(slot_dev, st.slots, next_slot * sizeof(Slot), hipMemcpyHostToDevice, s0), "slot upload");
  for (uint32_t i = 0; i < count; i++) HIP_TRY(clear_bit_rows(plans[i], dsts[i], s0), "one-bit rows");
  HIP_TRY(hipEventRecord(s->t0, s0), "hipEventRecord");
  uint64_t item_base = 0;
  for (uint32_t b = 0; b < nb; b++) {
    const BucketLaunch &L = launches[b];
    KernelArgs args{};
    args.slots = slot_dev + L.first;
    args.item_slot = item_dev + item_base;
    args.counters = reinterpret_cast<uint32_t *>(s->scratch + kStaticCounterOffset + (size_t)b * 8 * 128);
    item_base += L.items;
    args.slot_count = L.n;
    args.item_count = L.items;
    args.persistent = 1;
    args.tile_x = L.tx;
    args.tile_y = L.ty;
    args.lds_words = L.words;
    args.storage = L.storage;
    args.reduced = L.reduced;
    // the item-table kernel also zeroes the share counters and the status words
    HIP_TRY(launch_expand(args.slots, L.n, L.items, (uint32_t *)args.item_slot, args.counters, s0), "item table");
    HIP_TRY(launch_tiles(args, L.grid, s0), "persistent launch");
    s->diag.kernel_launches++;
  }
  HIP_TRY(hipEventRecord(s->t1, s0), "hipEventRecord");
  HIP_TRY(hipStreamSynchronize(s0), "hipStreamSynchronize");
  HIP_TRY(hipMemcpy(st.statuses, status_dev, sizeof(uint32_t) * count, hipMemcpyDeviceToHost), "status readback");
  float ms = 0;
  if (hipEventElapsedTime(&ms, s->t0, s->t1) == hipSuccess) s->last_ms = ms;
  s->last_launches = nb;
  s->diag.submissions += count;
  jxrhip_status result = JXRHIP_OK;
  for (uint32_t i = 0; i < count; i++) {
    if (stages) stages[i] = st.statuses[i];
    if (st.statuses[i] && result == JXRHIP_OK) {
      result = fail(JXRHIP_ERR_KERNEL_ARITHMETIC, "image " + std::to_string(i) +
                                                      ": reconstruction arithmetic overflow (stage " +
                                                      std::to_string(st.statuses[i]) + ")");
    }
  }
  return result;
}

// ---- captured batches -----------------------------------------------------
extern "C" jxrhip_status jxrhip_batch_graph_create(jxrhip_session *s, const jxrhip_plan *const *plans, uint32_t count,
                                                   void *const *dsts, jxrhip_batch_graph **out) {
  if (!s || !plans || !dsts || !out) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (count == 0) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "empty batch");
  for (uint32_t i = 0; i < count; i++) {
    if (!plans[i] || !dsts[i]) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null plan or destination");
    for (uint32_t a = 0; a < plans[i]->arena_count; a++)
      if (!plans[i]->arenas[a]->identity)
        return fail(JXRHIP_ERR_INVALID_ARGUMENT, "a captured batch needs arenas with identity (cached uploads)");
  }
  std::unique_ptr<jxrhip_batch_graph> gr(new (std::nothrow) jxrhip_batch_graph());
  if (!gr) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "out of host memory");
  gr->session = s;
  std::lock_guard<std::mutex> g(s->lock);
  make_current(s);
  hipStream_t s0 = s->streams[0];
  auto undo = [&](jxrhip_status st) {
    (void)hipStreamSynchronize(s0);
    release_prepared(s, gr->pb);
    if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
    if (gr->graph) (void)hipGraphDestroy(gr->graph);
    if (gr->host_statuses) (void)hipHostFree(gr->host_statuses);
    if (gr->done) (void)hipEventDestroy(gr->done);
    return st;
  };
  if (hipHostMalloc((void **)&gr->host_statuses, sizeof(uint32_t) * count, hipHostMallocDefault) != hipSuccess)
    return undo(fail(JXRHIP_ERR_DEVICE, "hipHostMalloc"));
  if (hipEventCreateWithFlags(&gr->done, hipEventDisableTiming) != hipSuccess)
    return undo(fail(JXRHIP_ERR_DEVICE, "hipEventCreate"));
  jxrhip_status st = prepare_persistent(s, plans, count, dsts, nullptr, gr->pb);
  if (st) return undo(st);
  const uint64_t launches_before = s->diag.kernel_launches;
  hipError_t e = hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return undo(hip_fail(e, "hipStreamBeginCapture"));
  st = enqueue_persistent(s, gr->pb);
  std::string reason = st ? g_error : std::string();
  e = hipStreamEndCapture(s0, &gr->graph);
  s->diag.kernel_launches = launches_before;
  if (st) return undo(fail(st, reason));
  if (e != hipSuccess) return undo(hip_fail(e, "hipStreamEndCapture"));
  e = hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) return undo(hip_fail(e, "hipGraphInstantiate"));
  s->live_graphs++;
  *out = gr.release();
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_batch_graph_launch(jxrhip_batch_graph *gr, jxrhip_batch **out) {
  if (!gr || !out) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  jxrhip_session *s = gr->session;
  std::lock_guard<std::mutex> g(s->lock);
  if (gr->in_flight) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "the previous replay of this graph has not been waited for");
  make_current(s);
  std::unique_ptr<jxrhip_batch> b(new (std::nothrow) jxrhip_batch());
  if (!b) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "out of host memory");
  const uint32_t count = (uint32_t)gr->pb.plans.size();
  b->session = s;
  b->count = count;
  b->graph = gr;
  b->host_statuses = gr->host_statuses;
  b->done = gr->done;
  hipStream_t s0 = s->streams[0];
  HIP_TRY(hipEventRecord(s->t0, s0), "hipEventRecord");
  HIP_TRY(hipGraphLaunch(gr->exec, s0), "hipGraphLaunch");
  HIP_TRY(hipEventRecord(s->t1, s0), "hipEventRecord");
  HIP_TRY(hipMemcpyAsync(gr->host_statuses, gr->pb.statuses.ptr, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, s0),
          "status readback");
  HIP_TRY(hipEventRecord(gr->done, s0), "hipEventRecord");
  gr->in_flight = true;
  b->timed = true;
  s->last_launches = (uint32_t)gr->pb.launches.size();
  s->diag.kernel_launches += gr->pb.launches.size();
  s->diag.submissions += count;
  session_ref(s);
  *out = b.release();
  return JXRHIP_OK;
}

extern "C" void jxrhip_batch_graph_destroy(jxrhip_batch_graph *gr) {
  if (!gr) return;
  jxrhip_session *s = gr->session;
  make_current(s);
  (void)hipEventSynchronize(gr->done);
  std::lock_guard<std::mutex> g(s->lock);
  (void)hipStreamSynchronize(s->streams[0]);
  release_prepared(s, gr->pb);
  if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
  if (gr->graph) (void)hipGraphDestroy(gr->graph);
  if (gr->host_statuses) (void)hipHostFree(gr->host_statuses);
  if (gr->done) (void)hipEventDestroy(gr->done);
  s->live_graphs--;
  delete gr;
}

// ---- tile autotuner ---------------------------------------------------------
extern "C" uint64_t jxrhip_plan_bucket_key(const jxrhip_plan *p) { return p ? p->geometry_key : 0; }

extern "C" jxrhip_status jxrhip_session_set_bucket_tile(jxrhip_session *s, uint64_t key, uint32_t tx, uint32_t ty) {
  if (!s) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null session");
  if ((tx == 0) != (ty == 0) || tx > 16 || ty > 16) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "tile shape");
  std::lock_guard<std::mutex> g(s->lock);
  if (tx == 0) s->tuned.erase(key);
  else s->tuned[key] = {tx, ty};
  s->tuned_groups.erase(key);
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_session_bucket_tile(jxrhip_session *s, uint64_t key, uint32_t *tx, uint32_t *ty) {
  if (!s || !tx || !ty) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(s->lock);
  auto it = s->tuned.find(key);
  *tx = it == s->tuned.end() ? 0 : it->second.first;
  *ty = it == s->tuned.end() ? 0 : it->second.second;
  return JXRHIP_OK;
}

// Times every candidate tile shape on each shape bucket of the batch (a
// persistent launch of that bucket alone, `repeats` times after one untimed
// run) and keeps the fastest for the bucket key. The destinations receive
// the images, as a batch would.
extern "C" jxrhip_status jxrhip_session_autotune(jxrhip_session *s, const jxrhip_plan *const *plans, uint32_t count,
                                                 void *const *dsts, uint32_t repeats, jxrhip_tile_choice *choices,
                                                 uint32_t capacity, uint32_t *bucket_count) {
  if (!s || !plans || !dsts) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  if (bucket_count) *bucket_count = 0;
  for (uint32_t i = 0; i < count; i++)
    if (!plans[i] || !dsts[i]) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null plan or destination");
  repeats = std::min<uint32_t>(std::max<uint32_t>(repeats, 1), 64);
  std::map<uint64_t, std::vector<uint32_t>> buckets;
  for (uint32_t i = 0; i < count; i++) buckets[plans[i]->geometry_key].push_back(i);
  if (bucket_count) *bucket_count = (uint32_t)buckets.size();
  // widths and heights that are not powers of two pack a stage's items
  // into fewer rounds of 256 threads for some plane layouts
  static const uint32_t cand[][2] = {{4, 4}, {8, 4}, {4, 8}, {8, 2}, {2, 8}, {4, 2}, {2, 4}, {8, 8},
                                     {16, 2}, {16, 4}, {2, 2}, {4, 1}, {6, 4}, {4, 6}, {6, 2}, {6, 3},
                                     {5, 4}, {3, 4}, {4, 3}, {5, 3}, {6, 6}, {12, 2}, {10, 2}, {3, 3}};
  std::lock_guard<std::mutex> g(s->lock);
  make_current(s);
  hipStream_t s0 = s->streams[0];
  uint32_t k = 0;
  for (auto &kv : buckets) {
    std::vector<const jxrhip_plan *> bp;
    std::vector<void *> bd;
    for (uint32_t i : kv.second) {
      bp.push_back(plans[i]);
      bd.push_back(dsts[i]);
    }
    const jxrhip_plan *p0 = bp[0];
    // the shape the bucket gets without tuning
    const auto saved = s->tuned.find(kv.first) != s->tuned.end() ? s->tuned[kv.first] : std::make_pair(0u, 0u);
    s->tuned.erase(kv.first);
    uint32_t dx, dy, dw;
    choose_tile(s, p0, dx, dy, dw);
    if (saved.first) s->tuned[kv.first] = saved;
    std::vector<std::pair<uint32_t, uint32_t>> shapes{{dx, dy}};
    for (auto &c : cand) {
      if (c[0] == dx && c[1] == dy) continue;
      uint32_t w = 0;
      for (const jxrhip_plan *p : bp) w = std::max(w, tile_words(p, c[0], c[1]));
      if ((size_t)w * 4 > std::min<size_t>(s->max_lds, 64 * 1024)) continue;
      shapes.push_back({c[0], c[1]});
    }
    // median device time of the bucket alone at a shape and a number of
    // workgroups per CU (0: as many as fit)
    auto time_shape = [&](uint32_t tx, uint32_t ty, uint32_t groups, double &med) -> jxrhip_status {
      const uint32_t force[2] = {tx, ty};
      PreparedBatch pb;
      jxrhip_status st = prepare_persistent(s, bp.data(), (uint32_t)bp.size(), bd.data(), force, pb, nullptr,
                                            nullptr, 0, groups);
      std::vector<double> ms;
      for (uint32_t r = 0; st == JXRHIP_OK && r <= repeats; r++) {
        if (hipEventRecord(s->t0, s0) != hipSuccess) st = fail(JXRHIP_ERR_DEVICE, "hipEventRecord");
        if (!st) st = enqueue_persistent(s, pb);
        if (!st && hipEventRecord(s->t1, s0) != hipSuccess) st = fail(JXRHIP_ERR_DEVICE, "hipEventRecord");
        if (!st && hipEventSynchronize(s->t1) != hipSuccess) st = fail(JXRHIP_ERR_DEVICE, "hipEventSynchronize");
        float t = 0;
        if (!st && hipEventElapsedTime(&t, s->t0, s->t1) != hipSuccess) st = fail(JXRHIP_ERR_DEVICE, "hipEventElapsedTime");
        if (!st && r > 0) ms.push_back(t);
      }
      (void)hipStreamSynchronize(s0);
      release_prepared(s, pb);
      if (st) return st;
      std::sort(ms.begin(), ms.end());
      med = ms[ms.size() / 2];
      return JXRHIP_OK;
    };
    double best_ms = 1e30, default_ms = 0;
    std::pair<uint32_t, uint32_t> best{dx, dy};
    for (auto &sh : shapes) {
      double med = 0;
      const jxrhip_status st = time_shape(sh.first, sh.second, 0, med);
      if (st) return st;
      if (sh.first == dx && sh.second == dy) default_ms = med;
      if (med < best_ms) {
        best_ms = med;
        best = sh;
      }
    }
    s->tuned[kv.first] = best;
    // then the workgroups per CU at that shape: as many as fit, or one or
    // two fewer (less L2 and LDS pressure per compute unit) when that is
    // faster by more than 1%
    s->tuned_groups.erase(kv.first);
    uint32_t groups = 0, fit = 0;
    if (!s->groups_per_cu) {
      uint32_t words = 0;
      for (const jxrhip_plan *p : bp) words = std::max(words, tile_words(p, best.first, best.second));
      int resident = 0;
      HIP_TRY(kernel_occupancy(p0->dev.storage, p0->scale != JXRHIP_SCALE_FULL ? 1 : 0, (size_t)words * 4,
                               &resident),
              "occupancy");
      fit = (uint32_t)std::max(resident, 1);
      for (uint32_t g2 = fit > 1 ? fit - 1 : 0; g2 >= 1 && g2 + 2 >= fit; g2--) {
        double med = 0;
        const jxrhip_status st = time_shape(best.first, best.second, g2, med);
        if (st) return st;
        if (med < best_ms * 0.99) {
          best_ms = med;
          groups = g2;
        }
      }
      if (groups) s->tuned_groups[kv.first] = groups;
    }
    if (choices && k < capacity) {
      jxrhip_tile_choice &c = choices[k];
      c.bucket_key = kv.first;
      c.tile_mbs_x = best.first;
      c.tile_mbs_y = best.second;
      c.images = (uint32_t)bp.size();
      c.candidates = (uint32_t)shapes.size();
      c.best_ms = best_ms;
      c.default_ms = default_ms;
      c.groups_per_cu = groups ? groups : (s->groups_per_cu ? s->groups_per_cu : fit);
      c.reserved = 0;
    }
    k++;
  }
  return JXRHIP_OK;
}

// ---- serving loop -----------------------------------------------------------
// A worker thread admits requests first come, first served while their device
// footprint fits the budget, and launches the admitted ones as waves: one
// prepared persistent batch per wave, on one of four compute streams, with
// uploads on a stream of their own. A wave starts whenever a compute stream is
// free, so requests that arrive while others run are admitted as memory frees
// up and their wave overlaps the running ones on the device.
static constexpr uint32_t kServerStreams = 4;

struct ServerRequest {
  uint64_t ticket = 0, tag = 0;
  const jxrhip_plan *plan = nullptr;
  void *dst = nullptr;
  uint64_t footprint = 0;
  uint64_t arrival_ns = 0, admitted_ns = 0;
  DeviceBuffer output;  // when the request brought no destination
};

struct ServerWave {
  std::vector<ServerRequest *> reqs;
  PreparedBatch pb;
  uint32_t slot = 0;  // compute stream, completion event and status mirror
};

struct jxrhip_server {
  jxrhip_session *s = nullptr;
  jxrhip_server_config cfg{};
  hipStream_t compute[kServerStreams] = {};
  hipEvent_t done[kServerStreams] = {};
  uint32_t *host_status[kServerStreams] = {};
  // per-stream pinned staging for the wave's slot table; a stream's wave
  // (and so its upload) has finished before the stream takes the next one
  Slot *host_slots[kServerStreams] = {};
  bool slot_busy[kServerStreams] = {};
  std::thread worker;
  std::mutex m;                   // arrivals, results, outstanding, stop, statistics
  std::condition_variable wake;   // the worker: arrivals or stop
  std::condition_variable ready;  // pollers and drain: results
  std::deque<ServerRequest *> arrivals;
  std::deque<jxrhip_server_result> results;
  uint64_t next_ticket = 1;
  uint64_t outstanding = 0;
  bool stop = false;
  jxrhip_server_stats stats{};
  std::vector<double> latencies, queue_waits;
  uint64_t wave_images = 0;
  // the worker's own
  std::deque<ServerRequest *> pending;
  std::vector<ServerRequest *> admitted;
  std::deque<ServerWave *> waves;
  uint64_t inflight_bytes = 0;
  uint32_t inflight_count = 0;
  // device bytes of result images the caller has not destroyed yet: they
  // count against the budget too, so a caller that falls behind slows
  // admission instead of exhausting device memory
  std::shared_ptr<std::atomic<uint64_t>> held = std::make_shared<std::atomic<uint64_t>>(0);
};

extern "C" uint64_t jxrhip_server_clock_ns(void) {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

static void server_finish(jxrhip_server *sv, ServerRequest *r, jxrhip_status status, uint32_t stage, uint64_t t) {
  jxrhip_session *s = sv->s;
  jxrhip_server_result res{};
  res.ticket = r->ticket;
  res.tag = r->tag;
  res.status = status;
  res.stage = stage;
  res.latency_ms = (double)(t - r->arrival_ns) / 1e6;
  res.queue_ms = r->admitted_ns ? (double)(r->admitted_ns - r->arrival_ns) / 1e6 : 0.0;
  if (status == JXRHIP_OK && r->output.ptr) {
    jxrhip_image *im = new (std::nothrow) jxrhip_image();
    if (im) {
      im->session = s;
      im->buf = r->output;
      im->bytes = r->plan->output_bytes;
      im->held = sv->held;
      sv->held->fetch_add(im->buf.bytes);
      r->output = DeviceBuffer{};
      session_ref(s);
      res.image = im;
    } else {
      res.status = JXRHIP_ERR_INVALID_ARGUMENT;
    }
  }
  if (r->output.ptr) {
    std::lock_guard<std::mutex> g(s->lock);
    s->pool.put(r->output);
  }
  {
    std::lock_guard<std::mutex> g(sv->m);
    if (res.status == JXRHIP_OK) {
      sv->stats.completed++;
      sv->stats.pixels_completed += r->plan->pixels;
      sv->latencies.push_back(res.latency_ms);
      sv->queue_waits.push_back(res.queue_ms);
    } else if (res.status == JXRHIP_ERR_BUDGET) {
      sv->stats.rejected++;
    } else {
      sv->stats.failed++;
    }
    sv->results.push_back(res);
    sv->outstanding--;
  }
  sv->ready.notify_all();
  delete r;
}

static void server_launch_wave(jxrhip_server *sv, uint32_t slot) {
  jxrhip_session *s = sv->s;
  const size_t n = std::min<size_t>(sv->admitted.size(), sv->cfg.max_wave_images);
  ServerWave *w = new ServerWave();
  w->slot = slot;
  w->reqs.assign(sv->admitted.begin(), sv->admitted.begin() + (long)n);
  sv->admitted.erase(sv->admitted.begin(), sv->admitted.begin() + (long)n);
  std::vector<const jxrhip_plan *> plans;
  std::vector<void *> dsts;
  jxrhip_status st = JXRHIP_OK;
  {
    std::lock_guard<std::mutex> g(s->lock);
    make_current(s);
    for (ServerRequest *r : w->reqs) {
      if (!r->dst && !r->output.ptr &&
          s->pool.get(r->plan->output_bytes, r->output, s->diag.scratch_hits, s->diag.scratch_misses) != hipSuccess) {
        st = fail(JXRHIP_ERR_DEVICE, "output buffer");
        break;
      }
      plans.push_back(r->plan);
      dsts.push_back(r->dst ? r->dst : r->output.ptr);
    }
    if (!st) st = prepare_persistent(s, plans.data(), (uint32_t)n, dsts.data(), nullptr, w->pb, sv->compute[slot],
                                     sv->host_slots[slot], sv->cfg.max_wave_images);
    if (!st) st = enqueue_persistent(s, w->pb, sv->compute[slot]);
    if (!st && hipMemcpyAsync(sv->host_status[slot], w->pb.statuses.ptr, sizeof(uint32_t) * n,
                              hipMemcpyDeviceToHost, sv->compute[slot]) != hipSuccess)
      st = fail(JXRHIP_ERR_DEVICE, "status readback");
    if (!st && hipEventRecord(sv->done[slot], sv->compute[slot]) != hipSuccess)
      st = fail(JXRHIP_ERR_DEVICE, "hipEventRecord");
    if (st) {
      (void)hipStreamSynchronize(sv->compute[slot]);
      release_prepared(s, w->pb);
    }
  }
  if (st) {
    const uint64_t t = jxrhip_server_clock_ns();
    for (ServerRequest *r : w->reqs) {
      sv->inflight_bytes -= r->footprint;
      sv->inflight_count--;
      server_finish(sv, r, st, 0, t);
    }
    delete w;
    return;
  }
  {
    std::lock_guard<std::mutex> g(sv->m);
    sv->stats.waves++;
    sv->wave_images += n;
    sv->stats.max_wave_images = std::max<uint64_t>(sv->stats.max_wave_images, n);
  }
  sv->slot_busy[slot] = true;
  sv->waves.push_back(w);
}

static void server_loop(jxrhip_server *sv) {
  jxrhip_session *s = sv->s;
  (void)hipSetDevice(s->device);
  for (;;) {
    bool progressed = false;
    {
      std::lock_guard<std::mutex> g(sv->m);
      while (!sv->arrivals.empty()) {
        sv->pending.push_back(sv->arrivals.front());
        sv->arrivals.pop_front();
        progressed = true;
      }
      if (sv->stop && sv->pending.empty() && sv->admitted.empty() && sv->waves.empty()) break;
    }
    // completed waves release their requests' footprint
    for (auto it = sv->waves.begin(); it != sv->waves.end();) {
      ServerWave *w = *it;
      const hipError_t q = hipEventQuery(sv->done[w->slot]);
      if (q == hipErrorNotReady) {
        ++it;
        continue;
      }
      const uint64_t t = jxrhip_server_clock_ns();
      for (size_t i = 0; i < w->reqs.size(); i++) {
        const uint32_t stage = q == hipSuccess ? sv->host_status[w->slot][i] : 0;
        const jxrhip_status st =
            q != hipSuccess ? JXRHIP_ERR_DEVICE : stage ? JXRHIP_ERR_KERNEL_ARITHMETIC : JXRHIP_OK;
        sv->inflight_bytes -= w->reqs[i]->footprint;
        sv->inflight_count--;
        server_finish(sv, w->reqs[i], st, stage, t);
      }
      {
        std::lock_guard<std::mutex> g(s->lock);
        release_prepared(s, w->pb);
      }
      sv->slot_busy[w->slot] = false;
      delete w;
      it = sv->waves.erase(it);
      progressed = true;
    }
    // admission, first come first served: a request waits for the ones
    // before it, and one larger than the whole budget is refused
    while (!sv->pending.empty()) {
      ServerRequest *r = sv->pending.front();
      if (r->footprint > sv->cfg.hbm_budget_bytes) {
        sv->pending.pop_front();
        server_finish(sv, r, JXRHIP_ERR_BUDGET, 0, jxrhip_server_clock_ns());
        progressed = true;
        continue;
      }
      // results not yet destroyed by the caller count against the budget;
      // with nothing in flight one request is admitted regardless, so a
      // caller that drains before it polls still sees every request finish
      const uint64_t used = sv->inflight_bytes + sv->held->load();
      if (sv->inflight_count >= sv->cfg.max_inflight ||
          (sv->inflight_count > 0 && used + r->footprint > sv->cfg.hbm_budget_bytes))
        break;
      sv->pending.pop_front();
      r->admitted_ns = jxrhip_server_clock_ns();
      sv->inflight_bytes += r->footprint;
      sv->inflight_count++;
      {
        std::lock_guard<std::mutex> g(sv->m);
        sv->stats.admitted++;
        sv->stats.inflight_high_water_bytes = std::max(sv->stats.inflight_high_water_bytes, sv->inflight_bytes);
      }
      sv->admitted.push_back(r);
      progressed = true;
    }
    if (!sv->admitted.empty()) {
      for (uint32_t i = 0; i < sv->cfg.max_waves; i++)
        if (!sv->slot_busy[i]) {
          server_launch_wave(sv, i);
          progressed = true;
          break;
        }
    }
    if (!progressed) {
      std::unique_lock<std::mutex> g(sv->m);
      if (sv->arrivals.empty() && !sv->stop)
        sv->wake.wait_for(g, std::chrono::microseconds(sv->waves.empty() ? 1000 : 20));
    }
  }
}

extern "C" jxrhip_status jxrhip_server_create(jxrhip_session *s, const jxrhip_server_config *cfg,
                                              jxrhip_server **out) {
  if (!s || !out) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  std::unique_ptr<jxrhip_server> sv(new (std::nothrow) jxrhip_server());
  if (!sv) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "out of host memory");
  sv->s = s;
  if (cfg) sv->cfg = *cfg;
  make_current(s);
  if (!sv->cfg.hbm_budget_bytes) {
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    // 8 GiB (or 3/4 of free memory, if less): enough work in flight to fill
    // the device, and within what the buffer pool keeps for reuse, so that
    // admitted outputs recycle pooled buffers instead of allocating and
    // freeing device memory under load (which serialises the device)
    sv->cfg.hbm_budget_bytes = std::min<uint64_t>(free_b / 4 * 3, uint64_t(8) << 30);
  }
  if (!sv->cfg.max_inflight) sv->cfg.max_inflight = 4096;
  if (!sv->cfg.max_wave_images) sv->cfg.max_wave_images = 256;
  if (!sv->cfg.max_waves) sv->cfg.max_waves = kServerStreams;
  if (sv->cfg.max_waves > kServerStreams) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "max_waves above 4");
  if (sv->cfg.max_wave_images > 65536) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "max_wave_images above 65536");
  // idempotent: every handle is cleared once released
  auto undo = [&](jxrhip_status st) {
    for (uint32_t i = 0; i < kServerStreams; i++) {
      if (sv->compute[i]) (void)hipStreamDestroy(sv->compute[i]);
      if (sv->done[i]) (void)hipEventDestroy(sv->done[i]);
      if (sv->host_status[i]) (void)hipHostFree(sv->host_status[i]);
      if (sv->host_slots[i]) (void)hipHostFree(sv->host_slots[i]);
      sv->compute[i] = nullptr;
      sv->done[i] = nullptr;
      sv->host_status[i] = nullptr;
      sv->host_slots[i] = nullptr;
    }
    return st;
  };
  for (uint32_t i = 0; i < kServerStreams; i++) {
    if (hipStreamCreateWithFlags(&sv->compute[i], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sv->done[i], hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void **)&sv->host_status[i], sizeof(uint32_t) * sv->cfg.max_wave_images,
                      hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&sv->host_slots[i], sizeof(Slot) * sv->cfg.max_wave_images,
                      hipHostMallocDefault) != hipSuccess)
      return undo(fail(JXRHIP_ERR_DEVICE, "server streams"));
  }
  jxrhip_server *raw = sv.get();
  try {
    sv->worker = std::thread(server_loop, raw);
  } catch (...) {
    return undo(fail(JXRHIP_ERR_INVALID_ARGUMENT, "cannot start the server thread"));
  }
  {
    std::lock_guard<std::mutex> g(s->lock);
    s->live_servers++;
  }
  *out = sv.release();
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_server_submit(jxrhip_server *sv, const jxrhip_plan *p, void *dst, uint64_t dst_bytes,
                                              uint64_t tag, uint64_t arrival_ns, uint64_t *ticket) {
  if (!sv || !p) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  if (dst && dst_bytes < p->output_bytes) return fail(JXRHIP_ERR_BUFFER_TOO_SMALL, "destination smaller than the plan's output");
  ServerRequest *r = new (std::nothrow) ServerRequest();
  if (!r) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "out of host memory");
  r->plan = p;
  r->dst = dst;
  r->tag = tag;
  r->footprint = p->coefficient_bytes + p->descriptor_bytes + (dst ? 0 : p->output_bytes);
  const uint64_t now = jxrhip_server_clock_ns();
  r->arrival_ns = arrival_ns && arrival_ns <= now ? arrival_ns : now;
  {
    std::lock_guard<std::mutex> g(sv->m);
    if (sv->stop) {
      delete r;
      return fail(JXRHIP_ERR_INVALID_ARGUMENT, "server is stopping");
    }
    r->ticket = sv->next_ticket++;
    sv->outstanding++;
    sv->stats.submitted++;
    sv->arrivals.push_back(r);
    if (ticket) *ticket = r->ticket;
  }
  sv->wake.notify_one();
  return JXRHIP_OK;
}

extern "C" uint32_t jxrhip_server_poll(jxrhip_server *sv, jxrhip_server_result *out, uint32_t capacity,
                                       uint32_t timeout_us) {
  if (!sv || !out || !capacity) return 0;
  std::unique_lock<std::mutex> g(sv->m);
  if (sv->results.empty() && timeout_us)
    sv->ready.wait_for(g, std::chrono::microseconds(timeout_us), [&] { return !sv->results.empty(); });
  uint32_t n = 0;
  while (n < capacity && !sv->results.empty()) {
    out[n++] = sv->results.front();
    sv->results.pop_front();
  }
  return n;
}

extern "C" jxrhip_status jxrhip_server_drain(jxrhip_server *sv) {
  if (!sv) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null server");
  std::unique_lock<std::mutex> g(sv->m);
  sv->ready.wait(g, [&] { return sv->outstanding == 0; });
  return JXRHIP_OK;
}

static double nearest_rank(std::vector<double> v, double p) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  size_t k = (size_t)std::ceil(p / 100.0 * (double)v.size());
  k = std::min(std::max<size_t>(k, 1), v.size());
  return v[k - 1];
}

extern "C" jxrhip_status jxrhip_server_get_stats(jxrhip_server *sv, jxrhip_server_stats *out) {
  if (!sv || !out) return fail(JXRHIP_ERR_INVALID_ARGUMENT, "null argument");
  std::vector<double> lat, queue;
  {
    std::lock_guard<std::mutex> g(sv->m);
    *out = sv->stats;
    lat = sv->latencies;
    queue = sv->queue_waits;
    out->mean_wave_images = sv->stats.waves ? (double)sv->wave_images / (double)sv->stats.waves : 0.0;
  }
  {
    std::lock_guard<std::mutex> g(sv->s->lock);
    out->hbm_high_water_bytes = sv->s->pool.high_water();
  }
  out->p50_ms = nearest_rank(lat, 50);
  out->p99_ms = nearest_rank(lat, 99);
  out->queue_p50_ms = nearest_rank(queue, 50);
  out->queue_p99_ms = nearest_rank(queue, 99);
  return JXRHIP_OK;
}

extern "C" void jxrhip_server_destroy(jxrhip_server *sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> g(sv->m);
    sv->stop = true;
  }
  sv->wake.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  for (auto &r : sv->results)
    if (r.image) jxrhip_image_destroy(r.image);
  make_current(sv->s);
  for (uint32_t i = 0; i < kServerStreams; i++) {
    if (sv->compute[i]) (void)hipStreamDestroy(sv->compute[i]);
    if (sv->done[i]) (void)hipEventDestroy(sv->done[i]);
    if (sv->host_status[i]) (void)hipHostFree(sv->host_status[i]);
    if (sv->host_slots[i]) (void)hipHostFree(sv->host_slots[i]);
  }
  {
    std::lock_guard<std::mutex> g(sv->s->lock);
    sv->s->live_servers--;
  }
  delete sv;
}
