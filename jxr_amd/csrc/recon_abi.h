// recon_abi.h — device-side descriptors shared by the host runtime and the
// gfx950 reconstruction kernels. The host validates a plan and builds these
// once; the kernels only read them.
#pragma once
#include <stdint.h>

namespace jxr {

constexpr uint32_t kMaxPlanes = 5;     // planes of one work item: Y, U, V (or CMYK/K), alpha
// Colour planes of an N-component image (DecodeLimits::max_components,
// crates/jxr-core/src/request.rs:85); more than four run as plane groups.
constexpr uint32_t kMaxColors = 64;
constexpr uint32_t kThreads = 256;     // threads per workgroup (4 wave64)
// Waves each SIMD must hold: four workgroups per compute unit (LDS allows
// four 4:2:0 tiles of 4x4 macroblocks), so at most 128 VGPRs per lane.
constexpr uint32_t kMinWavesPerSimd = 4;

// Raster-ordered macroblock record (the reference's JxrMacroblockAbi words,
// crates/jxr-math/data/reconstruction.abi:3, without coded_x/y because the
// record's index is its position).
struct alignas(32) MbRec {
  uint32_t coef_off;  // i32 index into the plane's arena coefficients
  uint32_t qdc, qlp, qhp;
  uint32_t bands;     // JXRHIP_BANDS_*
  uint32_t pred;      // JXRHIP_PRED_* (NONE when the high-pass band is absent)
  uint32_t pad0, pad1;
};

struct DevPlane {
  const int32_t *coef;     // arena coefficients
  const MbRec *mbs;        // [mby][mbx]
  const uint32_t *colmask; // bit m: hard tile boundary at MB column m (or null)
  const uint32_t *rowmask;
  uint32_t bc, br;         // blocks across / down one macroblock
  uint32_t scale_first;    // x2 after the first-level transform
  uint32_t overlap;        // 0, 1, 2
  uint32_t alpha;
  uint32_t pad[3];
};

// How a reduced decode turns the coefficients it stops at into samples
// (crates/jxr-native/src/reconstruct/sample_scale.rs:57-70, 255-316), one per
// plane: a shift by the gain of the transform stages left out, or, under
// overlap filtering, the line through the reduced pictures of flat probe
// images, looked up for each sample at the probe sample placed like it.
enum ReduceMode : uint32_t {
  kReduceShift = 0,  // sample = coefficient >> shift
  kReduceFlat = 1,   // sample = floor((c * gain - bias) / divisor), 0 for 0
  kReduceNone = 2,   // probe pictures: the coefficient itself
};
struct DevReduce {
  const int64_t *terms;   // flat: {divisor, bias} of each probe sample
  const uint32_t *cols;   // flat: probe sample column of each reduced column
  const uint32_t *rows;   // flat: probe sample row of each reduced row
  int64_t gain;           // flat: probe pictures times the level span
  uint32_t mode;          // ReduceMode
  uint32_t shift;         // shift: gain bits of the stages left out
  uint32_t stride;        // flat: probe samples across
  uint32_t pad[3];        // 64 bytes
};

// Bias, rounding, shift, postscale of one channel (scaling.rs:42-124).
struct ChanScale {
  int32_t shifted_bias;
  int32_t rounding;
  uint32_t shift;
  uint32_t post_shift;
  uint32_t flags;  // 1 invalid, 2 padding, 4 post-shift present
  uint32_t pad[3];
};

// Output paths specialised for common requests; applied to tiles whose
// output arithmetic cannot overflow (integer recipes: every channel shares
// one scale with no post-shift, no alpha; the float recipe: one sample
// format for the colour channels and one for alpha).
enum OutRecipe : uint32_t {
  kRecipeGeneric = 0,
  kRecipeRgb3 = 1,   // YUV -> RGB or BGR, three stored channels
  kRecipeLuma1 = 2,  // Luma -> Luma, one stored channel
  // F32 storage: YUV -> RGB/BGR, luma -> RGB or luma -> luma, optionally
  // alpha last, no padding channel, no premultiply, |exponent bias| <= 2^23
  kRecipeF32 = 3,
};

// What the output stage stores (output_format/{packing,packed_color,planar}.rs).
enum OutKind : uint32_t {
  kOutTyped = 0,      // one typed sample per channel (interleaved or planar)
  kOutBits = 1,       // one-bit luma, eight pixels a byte, first pixel highest
  kOutPacked16 = 2,   // RGB555 / RGB565 word
  kOutRgb101010 = 3,  // RGB101010 word
  kOutRgbe = 4,       // shared-exponent RGBE word
  kOutYuv = 5,        // native-resolution planar YUV (+ alpha)
};

enum Conversion : uint32_t {
  kConvLuma = 0,      // Luma -> Luma
  kConvLumaRgb = 1,   // Luma -> RGB (replicated)
  kConvYuvRgb = 2,    // YUV -> RGB/RGBE
  kConvYuvkCmyk = 3,  // YUVK -> CMYK
  kConvYuvkDirect = 4,// YUVK -> CMYK direct
  kConvDirect = 5,    // N-component direct
};

struct alignas(64) DevImage {
  DevPlane planes[kMaxPlanes];
  uint32_t plane_count;
  uint32_t colors;         // colour planes
  int32_t alpha_plane;     // plane index of alpha, -1 none
  uint32_t mbx, mby;       // macroblock window
  uint32_t upsample;       // 0 none, 1 = 4:2:0, 2 = 4:2:2
  uint32_t crop_x, crop_y; // crop, relative to the window's sample origin
  uint32_t crop_w, crop_h;
  uint32_t conv;           // Conversion
  uint32_t primary_count;  // primaries of the layout
  uint32_t channels;       // stored samples per pixel
  uint32_t has_alpha;      // layout stores alpha as its last channel
  uint32_t reverse;        // BGR order
  uint32_t storage;        // JXRHIP_STORE_*
  uint32_t bytes;          // bytes per stored sample
  uint32_t planar;         // channel c -> surface plane c
  int32_t u16_max;
  uint32_t premultiply;
  uint32_t centering_x, centering_y;
  uint32_t f_scaled, f_alpha_scaled;       // F16/F32 scaled flags
  uint32_t f_mant, f_alpha_mant;           // F32 mantissa lengths
  int32_t f_bias, f_alpha_bias;            // F32 exponent biases
  uint32_t out_fast;       // every channel's |bias| + rounding < 2^29: plain output arithmetic allowed
  uint32_t recipe;         // OutRecipe: a specialised output path for plain-arithmetic tiles
  int32_t recipe_add;      // recipes: bias + rounding shared by every channel
  uint32_t recipe_shift;   // recipes: shift shared by every channel
  int32_t recipe_lo, recipe_hi;  // recipes: clamp of the stored value
  uint32_t out_kind;       // OutKind
  uint32_t out_depth;      // JXRHIP_DEPTH_* (packed field widths, one-bit polarity)
  uint32_t chroma_sub_y;   // planar YUV: 2 for 4:2:0 chroma rows, 1 otherwise
  // plane groups of an N-component image: this descriptor's channels are
  // channel_base .. channel_base + store_channels - 1 of a pixel of
  // pixel_channels samples (0, channels and channels for every other image)
  uint32_t pixel_channels;
  uint32_t channel_base;
  uint32_t store_channels;
  // reduced decode (JXRHIP_SCALE_*): luma samples across a macroblock
  // (16, 4 or 1), whether the tile stores raw reduced pictures (the probe
  // images of the sample scale) instead of output pixels, and each plane's
  // sample scale
  uint32_t scale;
  uint32_t mb_px;
  uint32_t raw;
  uint32_t pad_reduce;
  const DevReduce *reduce;
  uint64_t plane_offset[kMaxPlanes];       // output surface planes
  uint64_t row_stride[kMaxPlanes];
  ChanScale color_scale[8];
  ChanScale alpha_scale;
};

// One image of a launch: descriptor, output and status word.
struct alignas(32) Slot {
  const DevImage *img;
  uint8_t *out;
  uint32_t *status;
  uint32_t tiles_x;     // tiles across the window for this launch's tile shape
  uint32_t tile_start;  // first work item of this image in a persistent launch
};

struct KernelArgs {
  Slot inline_slot;      // per-image launch: the one image
  const Slot *slots;     // persistent: images, ordered by tile_start
  const uint32_t *item_slot;  // persistent: owning image of each work item
  uint32_t *counters;         // persistent: eight share counters, 128 bytes apart
  uint32_t slot_count;
  uint32_t item_count;   // work items (tiles) of the launch
  uint32_t persistent;
  uint32_t tile_x, tile_y;  // macroblocks per tile
  uint32_t lds_words;    // dynamic LDS in 32-bit words
  uint32_t storage;      // JXRHIP_STORE_* of every image of the launch
  uint32_t reduced;      // every image of the launch is a reduced decode
};

// Scratch of the allocation-free batch entry point, a static device array of
// kStaticScratchBytes per device: status words, the share counters of each
// launch (eight 128-byte lines), slot records, then the item table.
constexpr uint32_t kStaticMaxImages = 4096;
constexpr uint32_t kStaticMaxSlots = 8192;    // one per plane group of an image
constexpr uint32_t kStaticMaxLaunches = 64;   // shape buckets of one call
constexpr size_t kStaticScratchBytes = size_t(4) << 20;
constexpr size_t kStaticStatusOffset = 0;
constexpr size_t kStaticCounterOffset = kStaticStatusOffset + kStaticMaxImages * 4;
constexpr size_t kStaticSlotOffset = kStaticCounterOffset + kStaticMaxLaunches * 8 * 128;
constexpr size_t kStaticItemOffset = kStaticSlotOffset + kStaticMaxSlots * sizeof(Slot);
constexpr uint32_t kStaticMaxItems = (uint32_t)((kStaticScratchBytes - kStaticItemOffset) / 4);
static_assert(kStaticSlotOffset % 256 == 0 && kStaticItemOffset % 256 == 0, "scratch regions stay aligned");

// Halo of the low-pass region in macroblocks: one macroblock holds the DC of
// every block in the one-block halo of the second level; the subsampled
// first-level overlap (mode 2) reads a corner's neighbour one pair further.
__host__ __device__ inline uint32_t low_halo(uint32_t bc, uint32_t overlap) {
  return bc == 2 && overlap == 2 ? 2u : 1u;
}

// Rows of second-level output kept above and below a tile: the sample-level
// overlap reaches two rows past the tile edge and chroma upsampling one.
// Columns keep the whole halo block (four samples), so that rows start on
// 16-byte boundaries.
constexpr int kSampleHaloRows = 2;

// Dynamic LDS words one workgroup needs for a tile shape and plane set.
__host__ __device__ inline uint32_t sample_region_words(uint32_t bc, uint32_t br, uint32_t tx,
                                                        uint32_t ty) {
  return (tx * bc * 4 + 8) * (ty * br * 4 + 2 * kSampleHaloRows);
}
__host__ __device__ inline uint32_t low_region_words(uint32_t bc, uint32_t br, uint32_t overlap, uint32_t tx,
                                                     uint32_t ty) {
  uint32_t h = low_halo(bc, overlap);
  return (tx + 2 * h) * bc * (ty + 2 * h) * br;
}
// A reduced decode keeps a sample per block (a quarter) or per macroblock (a
// sixteenth): samples across one macroblock of a plane of `blocks` blocks.
__host__ __device__ inline uint32_t reduced_units(uint32_t blocks, uint32_t scale) {
  return scale == JXRHIP_SCALE_QUARTER ? blocks : 1u;
}
// Reduced sample region of a tile: rows start on a 16-byte boundary four
// samples left of the tile's first, and hold the tile, a one-sample halo and
// the overhang of the output stage's four-sample groups; two rows of halo
// above and below.
__host__ __device__ inline uint32_t reduced_region_width(uint32_t ux, uint32_t tx) { return (tx * ux + 15u) & ~3u; }
__host__ __device__ inline uint32_t reduced_sample_words(uint32_t ux, uint32_t uy, uint32_t tx, uint32_t ty) {
  return reduced_region_width(ux, tx) * (ty * uy + 2 * kSampleHaloRows);
}
// Per plane, the second level's view of the macroblock records of the tile
// and its one-macroblock halo (coefficient offset, HP step, bands and HP
// direction: four words each), written by the first level.
__host__ __device__ inline uint32_t mb_table_words(uint32_t tx, uint32_t ty) { return (tx + 2) * (ty + 2) * 4; }

}  // namespace jxr
