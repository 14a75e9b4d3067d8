/*
 * jxr_hip_codestream.h — host-side codestream reader and writer.
 *
 * The CPU half of the decode that sits before the reconstruction boundary:
 * parsing a JPEG XR codestream (raw T.832, or inside an Annex-A container)
 * and entropy-decoding its tile packets into the coefficient arenas
 * jxrhip_plan_create takes, as the reference's native decoder does on its
 * CPU workers before handing arenas to its device backend; and the mirror
 * image for the encoder: coding the arenas of jxrhip_encode into tile
 * packets, headers, an index table and, optionally, an Annex-A container.
 *
 * Reference interfaces these entry points replace (paths relative to the
 * reference repository):
 *   parse_codestream / ParsedCodestream / ParsedHeaders / CodestreamDirectory
 *       crates/jxr-native/src/codestream.rs:10-81,
 *       crates/jxr-native/src/header/{image,plane,types}.rs,
 *       crates/jxr-native/src/tail.rs:1-260           -> jxrhip_codestream_parse
 *   parse_annex_a / AnnexAImage
 *       crates/jxr-native/src/annex_a.rs:48-120       -> jxrhip_codestream_parse
 *   image_info (ImageInfo, PlaneInfo, TileGrid)
 *       crates/jxr-native/src/metadata.rs:14-140      -> jxrhip_codestream_get_info
 *   decode_tiles / decode_tiles_into (spatial and frequency packets,
 *       adaptive-VLC entropy decoding, DC/LP prediction)
 *       crates/jxr-native/src/tile_decode/mod.rs:31-146,
 *       crates/jxr-native/src/entropy/ (all files)     -> jxrhip_codestream_decode
 *   encode_codestream / write_codestream_headers / write_directory /
 *       encode_tile / write_annex_a
 *       crates/jxr-native/src/encode.rs, header/write.rs:19-258,
 *       tail.rs:110-140, tile_encode/ (all files), annex_a/writer.rs
 *                                                     -> jxrhip_encoded_codestream
 */
#ifndef JXR_HIP_CODESTREAM_H
#define JXR_HIP_CODESTREAM_H

#include "jxr_hip.h"
#include "jxr_hip_encode.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Codestream failures (NativeError / TileDecodeError / EntropyError of the
 * reference, crates/jxr-native/src/error.rs, tile_decode/error.rs,
 * entropy/error.rs): malformed or truncated syntax, a valid feature this
 * reader does not take, and an arithmetic overflow while decoding. */
#define JXRHIP_ERR_SYNTAX 9
#define JXRHIP_ERR_UNSUPPORTED 10
#define JXRHIP_ERR_OVERFLOW 11

/* What the headers, index and container say of one image
 * (ImageInfo / PlaneInfo / CodestreamHeader / ImagePlaneHeader). Colour,
 * depth and band values use this library's codes (JXRHIP_COLOR_*,
 * JXRHIP_DEPTH_*, JXRHIP_BANDS_*), mapped from the T.832 codes as
 * metadata.rs:198-300 maps them. */
typedef struct jxrhip_codestream_info {
  uint32_t width;            /* output samples across, before orientation   */
  uint32_t height;
  uint32_t margin_top;       /* coded margins (windowing, or inferred)       */
  uint32_t margin_left;
  uint32_t margin_bottom;
  uint32_t margin_right;
  uint32_t macroblocks_x;    /* coded plane, in macroblocks                 */
  uint32_t macroblocks_y;
  uint32_t tile_column_count;
  uint32_t tile_row_count;
  uint32_t overlap;          /* 0, 1, 2                                     */
  uint32_t hard_tiles;
  uint32_t frequency_mode;   /* packets per band rather than per tile       */
  uint32_t orientation;      /* SPATIAL_XFRM_SUBORDINATE (or the Annex-A tag) */
  uint32_t output_color;     /* JXRHIP_COLOR_* of OUTPUT_CLR_FMT            */
  uint32_t bit_depth;        /* JXRHIP_DEPTH_* of OUTPUT_BITDEPTH           */
  uint32_t color_format;     /* JXRHIP_COLOR_* of INTERNAL_CLR_FMT (primary) */
  uint32_t component_count;  /* primary components                          */
  uint32_t bands;            /* JXRHIP_BANDS_* of the primary plane         */
  uint32_t scaled;
  uint32_t chroma_centering_x;
  uint32_t chroma_centering_y;
  uint32_t shift_bits;
  uint32_t mantissa_length;
  int32_t exponent_bias;
  uint32_t alpha_mode;       /* JXRHIP_ALPHA_*                              */
  uint32_t alpha_bands;
  uint32_t alpha_scaled;
  uint32_t alpha_shift_bits;
  uint32_t alpha_mantissa_length;
  int32_t alpha_exponent_bias;
  uint32_t alpha_bit_depth;  /* JXRHIP_DEPTH_* of the separate alpha codestream */
  uint32_t red_blue_not_swapped;
  uint32_t premultiplied_alpha;
  uint32_t short_header;
  uint32_t long_word;
  uint32_t trim_flexbits_present;
  uint32_t index_table_present;
  uint32_t profile_idc;      /* first PROFILE_LEVEL_INFO, or 111/255 when absent */
  uint32_t level_idc;
  uint32_t annex_a;          /* nonzero: the input is an Annex-A container  */
  uint8_t pixel_format_guid[16]; /* Annex-A PIXEL_FORMAT (zero otherwise)   */
  uint64_t codestream_offset; /* primary codestream within the input        */
  uint64_t codestream_bytes;
  uint64_t alpha_offset;      /* separate alpha codestream, or 0 / 0        */
  uint64_t alpha_bytes;
} jxrhip_codestream_info;

typedef struct jxrhip_codestream jxrhip_codestream;
typedef struct jxrhip_decoded jxrhip_decoded;

/* Parses the container (if any), image and plane headers, index table and
 * profile area of `bytes` (copied; the caller's buffer may go afterwards).
 * Refuses malformed syntax with JXRHIP_ERR_SYNTAX before any tile is read. */
jxrhip_status jxrhip_codestream_parse(const uint8_t *bytes, uint64_t length, jxrhip_codestream **out);
void jxrhip_codestream_destroy(jxrhip_codestream *codestream);
jxrhip_status jxrhip_codestream_get_info(const jxrhip_codestream *codestream, jxrhip_codestream_info *out);
/* Tile column widths and row heights in macroblocks (the complete grid,
 * the last column and row inferred); arrays of tile_column_count and
 * tile_row_count entries. */
jxrhip_status jxrhip_codestream_tile_grid(const jxrhip_codestream *codestream, uint32_t *column_widths,
                                          uint32_t *row_heights);

/* Entropy-decodes the tile packets into coefficient arenas
 * (decode_tiles, tile_decode/mod.rs:31-146): the primary arena holds the
 * primary components and, after them, an integrated alpha plane; a separate
 * alpha codestream becomes a second arena. `scale` (JXRHIP_SCALE_*) keeps
 * the bands a reduced decode reads (DecodeScale::retained_bands,
 * crates/jxr-core/src/request.rs:33-39): a reduced decode needs
 * frequency-mode packets, a quarter one also the LP band (prepare.rs
 * validate_native_scale). `tile_mask`, when non-NULL, holds one byte per
 * tile in raster order; tiles whose byte is zero are not read and not
 * stored (TilePlan::required_for_reconstruction). Tiles are decoded on up to
 * `threads` host threads (0 = one per host core, at most 16). */
jxrhip_status jxrhip_codestream_decode(const jxrhip_codestream *codestream, uint32_t scale,
                                       const uint8_t *tile_mask, uint32_t threads, jxrhip_decoded **out);
/* The coefficient counts of the arenas a decode with these arguments
 * produces (tile_decode/mod.rs:121-127 coefficient_count): the primary
 * arena's and the separate alpha arena's (0 without one). */
jxrhip_status jxrhip_codestream_coefficient_count(const jxrhip_codestream *codestream, uint32_t scale,
                                                  const uint8_t *tile_mask, uint64_t *primary, uint64_t *alpha);
/* jxrhip_codestream_decode with the coefficients written into caller-owned
 * storage (tile_decode/mod.rs:99-119 decode_tiles_into), for example pinned
 * host memory from jxrhip_host_alloc, from which a session's upload is an
 * asynchronous DMA. The arenas point into that storage, which must outlive
 * them and every plan made from them; a submission reads it on its stream,
 * so pinned storage stays unchanged until that submission has completed.
 * Storage below the counts
 * jxrhip_codestream_coefficient_count gives is JXRHIP_ERR_BUFFER_TOO_SMALL. */
jxrhip_status jxrhip_codestream_decode_into(const jxrhip_codestream *codestream, uint32_t scale,
                                            const uint8_t *tile_mask, uint32_t threads, int32_t *primary,
                                            uint64_t primary_capacity, int32_t *alpha, uint64_t alpha_capacity,
                                            jxrhip_decoded **out);
void jxrhip_decoded_destroy(jxrhip_decoded *decoded);
/* which = 0: primary arena; 1: separate alpha arena (NULL if none). Valid as
 * long as `decoded`. */
const jxrhip_arena *jxrhip_decoded_arena(const jxrhip_decoded *decoded, uint32_t which);

/* ---- writer ------------------------------------------------------------- */
#define JXRHIP_TRIM_NONE 0xffffffffu

typedef struct jxrhip_codestream_options {
  uint32_t frequency_mode; /* one packet per band of each tile               */
  uint32_t trim_flexbits;  /* JXRHIP_TRIM_NONE, or 0..15 (TRIM_FLEXBITS)     */
  uint32_t annex_a;        /* wrap in an Annex-A container (required for a
                              separate alpha plane)                          */
  uint32_t short_header;   /* 0: choose (short when the sizes fit), 1: long  */
  uint32_t profile_idc;    /* PROFILE_LEVEL_INFO written after the index     */
  uint32_t level_idc;
} jxrhip_codestream_options;

/* Codes the arenas of an encoder output as a codestream: headers from the
 * encode parameters, the macroblocks' quantization parameters as image- or
 * tile-level sets with per-macroblock indices, and every coefficient through
 * the adaptive entropy coder. *bytes (freed with jxrhip_free_bytes) receives
 * the file. A decode of the result reproduces the arenas exactly when every
 * band is coded and no flexbits are trimmed. */
jxrhip_status jxrhip_encoded_codestream(const jxrhip_encoded *encoded, const jxrhip_codestream_options *options,
                                        uint8_t **bytes, uint64_t *length);
void jxrhip_free_bytes(uint8_t *bytes);

#ifdef __cplusplus
}
#endif
#endif
