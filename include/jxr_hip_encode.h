/*
 * jxr_hip_encode.h — forward analysis into a coefficient arena.
 *
 * The encoder's transform and quantization stages without entropy coding:
 * colour transform, chroma decimation, overlap prefilters, forward core
 * transforms, quantization and high-pass prediction residuals, laid out as
 * the reference's entropy decoder lays out an arena (component-major,
 * tile-major, crates/jxr-native/src/tile_decode/output.rs:232-299). It
 * stands in for the reference encoder (crates/jxr-native/src/encode,
 * src/analyze.rs) so that seeded synthetic images become the arenas the
 * reconstruction consumes; with all steps 1 and no chroma decimation the
 * reconstruction returns the input exactly.
 */
#ifndef JXR_HIP_ENCODE_H
#define JXR_HIP_ENCODE_H

#include "jxr_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define JXRHIP_INPUT_U8 0u
#define JXRHIP_INPUT_U16 1u
#define JXRHIP_INPUT_F32 2u

typedef struct jxrhip_encode_params {
  uint32_t width;          /* pixels                                          */
  uint32_t height;
  uint32_t input_kind;     /* JXRHIP_INPUT_*                                  */
  uint32_t input_channels; /* 1 gray, 2 gray+alpha, 3 RGB, 4 RGBA (interleaved) */
  uint32_t color_format;   /* JXRHIP_COLOR_LUMA or JXRHIP_COLOR_YUV420/422/444 */
  uint32_t alpha_mode;     /* JXRHIP_ALPHA_* for inputs with alpha            */
  uint32_t overlap;        /* 0, 1, 2                                         */
  uint32_t hard_tiles;
  uint32_t tile_column_count; /* 0 = one tile                                 */
  const uint32_t *tile_column_widths; /* macroblocks                         */
  uint32_t tile_row_count;
  const uint32_t *tile_row_heights;
  uint32_t scaled;         /* SCALED_FLAG                                     */
  uint32_t qp_dc;          /* coded QP indices, 0 = step 1 (lossless)         */
  uint32_t qp_lp;
  uint32_t qp_hp;
  uint32_t bands;          /* JXRHIP_BANDS_*                                  */
  uint32_t shift_bits;     /* fixed-point inputs                              */
  uint32_t mantissa_length; /* f32 inputs                                     */
  int32_t exponent_bias;
  uint32_t alpha_qp;       /* QP of the alpha plane (all bands)               */
  uint32_t seed;           /* varies quantizer sets between macroblocks       */
  uint32_t vary_quantizers; /* nonzero: per-tile and per-macroblock steps vary */
  uint32_t chroma_qp_set;  /* nonzero: components 1.. take the qp_chroma_* QPs */
  uint32_t qp_chroma_dc;
  uint32_t qp_chroma_lp;
  uint32_t qp_chroma_hp;
} jxrhip_encode_params;

typedef struct jxrhip_encoded jxrhip_encoded;

jxrhip_status jxrhip_encode(const void *pixels, uint64_t pixel_bytes,
                            const jxrhip_encode_params *params, jxrhip_encoded **out);
void jxrhip_encoded_destroy(jxrhip_encoded *encoded);
/* which = 0: primary arena; 1: separate alpha arena (NULL if none). The
 * returned struct and everything it points to live as long as `encoded`. */
const jxrhip_arena *jxrhip_encoded_arena(const jxrhip_encoded *encoded, uint32_t which);
/* Quantization step of a coded QP (T.832 QP mapping,
 * crates/jxr-native/src/tile_decode/quantizer.rs:309-331). */
uint32_t jxrhip_quant_step(uint32_t qp, uint32_t scaled, uint32_t scaled_shift);
/* Quantization parameters (coded QPs, 0 = lossless) and overlap mode that a
 * quality 0..100 resolves to for an integer source whose samples span
 * `range` values (255 for 8-bit, 65535 for 16-bit, 31 and 63 for RGB555 and
 * RGB565 fields), as the reference encoder resolves a quality request
 * (crates/jxr-native/src/encode/quality.rs:101-480): `luma` receives DC, LP
 * and HP parameters of component 0, `chroma` those of components 1 and 2,
 * which are colour differences at color_format's chroma sampling when
 * `rgb_source` is nonzero and take the luma parameters otherwise; *overlap
 * receives the mode the quality picks when a request leaves it open. */
jxrhip_status jxrhip_quality_parameters(uint32_t quality, uint32_t color_format, uint32_t rgb_source,
                                        uint32_t range, uint32_t luma[3], uint32_t chroma[3],
                                        uint32_t *overlap);
/* As jxrhip_quality_parameters for a source of any sample scale: `range`,
 * `is_float` and `lowest_quality` as jxrhip_sample_scale gives them. Float
 * sources keep their differences at the primary's whole levels, and every
 * quality below `lowest_quality` codes as it does (quality.rs:295-329). */
jxrhip_status jxrhip_quality_parameters_scaled(uint32_t quality, uint32_t color_format, uint32_t rgb_source,
                                               uint32_t range, uint32_t is_float, uint32_t lowest_quality,
                                               uint32_t luma[3], uint32_t chroma[3], uint32_t *overlap);
/* SampleScale (quality.rs:194-292): how the coded samples of a source
 * stored at `depth` (JXRHIP_DEPTH_*; RGBE is JXRHIP_DEPTH_U8 with
 * output_color JXRHIP_COLOR_RGBE) compare with 8-bit samples, for its colour
 * planes (alpha = 0) or its alpha plane (alpha = 1). *range is the coded
 * distance standing for the 8-bit range: an integer's maximum less the bits
 * shift_bits drops, a fixed-point sample's one, and for floats eight
 * octaves of bit patterns (two for alpha, at most the patterns up to one),
 * fewer for a 32-bit float whose exponent bias is below six; *is_float says
 * the samples are float patterns; *lowest_quality is the quality every
 * lower one codes as (10 for RGBE, else 0). */
jxrhip_status jxrhip_sample_scale(uint32_t depth, uint32_t output_color, uint32_t shift_bits,
                                  uint32_t mantissa_length, int32_t exponent_bias, uint32_t alpha,
                                  uint32_t *range, uint32_t *is_float, uint32_t *lowest_quality);
/* The reconstructed sample whose f32 output is nearest `bits`
 * (crates/jxr-native/src/output_format/scaling.rs:177-229). */
jxrhip_status jxrhip_float32_sample(uint32_t bits, uint32_t scaled, uint32_t mantissa_length,
                                    int32_t exponent_bias, int32_t *sample);

#ifdef __cplusplus
}
#endif
#endif
