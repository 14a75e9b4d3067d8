/*
 * jxr_oracle.h — CPU restatement of the reference decoder's reconstruction
 * and output stages. TEST INFRASTRUCTURE ONLY: the parity tests, the smoke
 * check and the CPU-baseline leg of the benchmark call it as the checker.
 * Nothing in the shipped library links or calls it.
 *
 * Every operator uses checked i32 arithmetic (i64 intermediates with a range
 * test): overflow is an error, never a wrap, as in
 * crates/jxr-math/src/arithmetic.rs:58-84. Right shifts of negative values
 * are arithmetic (gcc semantics), as Rust's i32 >> is.
 *
 * Return codes: 0 = ok, JXRHIP_ERR_* otherwise; *stage receives the
 * JXRHIP_STAGE_* of an arithmetic failure.
 */
#ifndef JXR_ORACLE_H
#define JXR_ORACLE_H

#include "../include/jxr_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Whole image: arena(s) -> output bytes laid out as desc->surface says. */
int jxro_decode(const jxrhip_image_desc *desc, uint8_t *out, uint64_t out_len,
                uint32_t *stage, char *message, size_t message_len);

/* Reconstructed signed samples of one plane (before colour and output):
 * plane_index < primary plane count selects a primary plane, == it selects
 * the separate alpha plane. Writes width*height samples. */
int jxro_reconstruct_plane(const jxrhip_image_desc *desc, uint32_t plane_index,
                           int32_t *samples, uint64_t capacity, uint32_t *width,
                           uint32_t *height, uint32_t *origin_x, uint32_t *origin_y,
                           uint32_t *stage);

/* Decodes `count` images on `threads` host threads (per-image split, as the
 * reference's batch decoder does). outs[i] has out_lens[i] bytes. Returns the
 * first failure. */
int jxro_decode_many(const jxrhip_image_desc *const *descs, uint8_t *const *outs,
                     const uint64_t *out_lens, uint32_t count, uint32_t threads);

/* Operators, exported for the known-answer and round-trip tests.
 * Each returns 0, or 1 on overflow (the values are then unspecified). */
int jxro_inverse_core_transform(int32_t v[16]);
int jxro_forward_core_transform(int32_t v[16]);
int jxro_inverse_chroma_420(int32_t v[4]);
int jxro_forward_chroma_420(int32_t v[4]);
int jxro_inverse_chroma_422(int32_t v[8]);
int jxro_forward_chroma_422(int32_t v[8]);
int jxro_inverse_overlap_4(int32_t v[4]);
int jxro_forward_overlap_4(int32_t v[4]);
int jxro_inverse_overlap_4x4(int32_t v[16]);
int jxro_forward_overlap_4x4(int32_t v[16]);
int jxro_inverse_overlap_2x2(int32_t v[4]);
int jxro_forward_overlap_2x2(int32_t v[4]);
int jxro_inverse_overlap_2(int32_t v[2]);
int jxro_forward_overlap_2(int32_t v[2]);
int jxro_upsample_pair(int32_t previous, int32_t current, int32_t next, uint32_t centering,
                       int32_t out[2]);
/* HP prediction of one macroblock's 256 coefficients (prediction.rs:30-97). */
int jxro_predict_high_pass(int32_t coefficients[256], uint32_t direction,
                           uint32_t block_columns, uint32_t block_rows);
int jxro_premultiply_unsigned(uint32_t value, uint32_t alpha, uint32_t maximum, uint32_t *out);
int jxro_premultiply_signed(int32_t value, int32_t alpha, int32_t maximum, int32_t *out);
/* RGBE word of three scaled channels (rgbe.rs:6-100): bytes R, G, B, E. */
uint32_t jxro_rgbe_word(const int32_t rgb[3]);
int jxro_float32_bits(int32_t sample, uint32_t scaled, uint32_t mantissa_length,
                      int32_t exponent_bias, uint32_t *bits);
int jxro_float16_bits(int32_t sample, uint32_t scaled, uint16_t *bits);

/* Internal-to-output colour of one pixel (output_format/color.rs:9-47):
 * JXRHIP_COLOR_* codes; 0 ok, 1 overflow, 2 unsupported combination. */
int jxro_convert(uint32_t internal, uint32_t output, const int32_t in[4], int32_t packed_rgb,
                 int32_t red_blue_not_swapped, int32_t out[4]);

/* ProbeAxis (sample_scale.rs:336-381) of an axis whose regions have the
 * macroblock boundaries `regions` (count entries, from 0): the probe's
 * region boundaries and, for each macroblock, its probe macroblock.
 * 0 ok, 1 a partition that does not advance or more positions than
 * `capacity`. */
int jxro_probe_axis(const uint32_t *regions, uint32_t count, uint32_t *bounds, uint32_t *bound_count,
                    uint32_t *positions, uint32_t capacity);

/* Forward analysis at unit steps of one component (mbx x mby macroblocks
 * of bc x br blocks; samples row-major, mbx*bc*4 across) with hard regions
 * at the macroblock boundaries column_bounds x row_bounds (from 0 to the
 * extent): the exact inverse of its full reconstruction, as
 * analyze_component (analyze.rs:56-86) runs it for the sample-scale probe.
 * `doubled` halves the DC and low-pass coefficients, rounding the magnitude
 * half up. Writes 256 coefficients per macroblock (block k at 16k). */
int jxro_analyze_plane(const int32_t *samples, uint32_t mbx, uint32_t mby, uint32_t bc, uint32_t br,
                       int32_t doubled, uint32_t overlap, const uint32_t *column_bounds, uint32_t column_count,
                       const uint32_t *row_bounds, uint32_t row_count, int32_t *coefficients);

/* Postfilter over a whole plane with the given tile boundaries (sample
 * units, including 0 and the extent), as reconstruct/overlap.rs:124-161. */
int jxro_overlap_plane(int32_t *samples, uint32_t width, uint32_t height,
                       const uint32_t *column_boundaries, uint32_t column_count,
                       const uint32_t *row_boundaries, uint32_t row_count, int32_t hard);
/* Subsampled first-level postfilter over a plane
 * (reconstruct/subsampled_overlap.rs:110-283). */
int jxro_subsampled_overlap_plane(int32_t *samples, uint32_t width, uint32_t height,
                                  const uint32_t *column_boundaries, uint32_t column_count,
                                  const uint32_t *row_boundaries, uint32_t row_count,
                                  int32_t hard);

#ifdef __cplusplus
}
#endif
#endif
