/*
 * jxr_hip.h — C ABI of the MI355X (gfx950) JPEG XR reconstruction backend.
 *
 * The boundary sits where the reference's CUDA backend sits: after CPU
 * entropy decoding has filled a coefficient arena, and before device-resident
 * output pixels. Everything crossing it is a plain C struct, pointer or size;
 * no HIP, torch or C++ types appear in any signature, and no C++ exception
 * crosses it. Every call returns a jxrhip_status; on failure the
 * thread-local jxrhip_last_error() names the reason.
 *
 * Reference interfaces this ABI replaces (paths relative to the reference
 * repository):
 *   CoefficientArena / MacroblockMetadata / CoefficientPlane
 *       crates/jxr-core/src/plan.rs:64-162            -> jxrhip_arena
 *   OutputFormatRequest / OutputBitDepth / PixelFormat
 *       crates/jxr-core/src/output_policy.rs:20-186    -> jxrhip_output_format
 *   SurfaceLayout / SurfacePlaneLayout
 *       crates/jxr-core/src/surface.rs:10-27           -> jxrhip_surface_layout
 *   CudaDecodePlan::from_prepared
 *       crates/jxr-cuda/src/plan.rs:91-148             -> jxrhip_plan_create
 *   CudaDecoderSession::{new,submit,submit_into,submit_batch,
 *       submit_dense_batch,decode_to_host,readback}
 *       crates/jxr-cuda/src/session.rs:47-268          -> jxrhip_session_*, jxrhip_submit*
 *   CudaSubmission::{is_complete,wait}
 *       crates/jxr-cuda/src/submission.rs:41-118       -> jxrhip_submission_*
 *   plan_cuda_route / CUDA_AUTO_THRESHOLD
 *       crates/jxr-cuda/src/route.rs:8-50              -> jxrhip_route
 *   CudaBufferPoolDiagnostics / CudaUploadCacheDiagnostics
 *       crates/jxr-cuda/src/buffer_pool.rs:12-27,
 *       crates/jxr-cuda/src/upload_cache.rs:18-33      -> jxrhip_diagnostics
 *
 * Sizes and offsets are 64-bit throughout, which lifts the reference's u32
 * device-ABI limits (crates/jxr-cuda/src/plan.rs:723-730,
 * crates/jxr-cuda/src/destination.rs:191-205).
 */
#ifndef JXR_HIP_H
#define JXR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JXRHIP_ABI_VERSION 1u

/* ---- status codes (crates/jxr-cuda/src/error.rs:8-92) ------------------ */
typedef int32_t jxrhip_status;
#define JXRHIP_OK 0
#define JXRHIP_ERR_INVALID_PLAN 1        /* CudaError::InvalidPlan             */
#define JXRHIP_ERR_UNSUPPORTED_OUTPUT 2  /* CudaError::UnsupportedOutputFormat */
#define JXRHIP_ERR_UNAVAILABLE 3         /* no usable device                   */
#define JXRHIP_ERR_DEVICE 4              /* HIP runtime failure                */
#define JXRHIP_ERR_KERNEL_ARITHMETIC 5   /* CudaError::KernelArithmetic{status}*/
#define JXRHIP_ERR_INVALID_ARGUMENT 6
#define JXRHIP_ERR_BUFFER_TOO_SMALL 7
#define JXRHIP_ERR_BUDGET 8              /* admission or scratch budget        */

/* First failing stage recorded in a submission's device status word
 * (crates/jxr-cuda/src/kernels/main.rs:52-71,106-207). */
#define JXRHIP_STAGE_FIRST_TRANSFORM 1u
#define JXRHIP_STAGE_OVERLAP 2u
#define JXRHIP_STAGE_PREDICTION 3u
#define JXRHIP_STAGE_SECOND_TRANSFORM 4u
#define JXRHIP_STAGE_SAMPLE_SCALE 5u /* reduced decode: the flat-probe sample scale */
#define JXRHIP_STAGE_OUTPUT 16u

/* ---- coefficient arena (crates/jxr-core/src/plan.rs:11-162) ------------ */
/* BandPresence */
#define JXRHIP_BANDS_DC_ONLY 0u
#define JXRHIP_BANDS_NO_HIGH_PASS 1u
#define JXRHIP_BANDS_NO_FLEXBITS 2u
#define JXRHIP_BANDS_ALL 3u
/* PredictionMode (high-pass direction) */
#define JXRHIP_PRED_NONE 0u
#define JXRHIP_PRED_FROM_LEFT 1u
#define JXRHIP_PRED_FROM_TOP 2u
#define JXRHIP_PRED_FROM_TOP_LEFT 3u /* rejected, as plan.rs:700-705 does */

/* CoefficientPlane: one component's macroblocks inside an arena. */
typedef struct jxrhip_coefficient_plane {
  uint64_t coefficient_offset;
  uint64_t coefficient_count;
  uint64_t macroblock_offset;
  uint64_t macroblock_count;
  uint32_t block_columns; /* 4x4 blocks across one component macroblock: 4 or 2 */
  uint32_t block_rows;    /* 4x4 blocks down one component macroblock: 4 or 2   */
} jxrhip_coefficient_plane;

/* CoefficientArena plus the per-arena parts of PreparedPlan the
 * reconstruction reads (overlap mode and tile grid, plan.rs:198-247).
 * Coefficients are i32, component-major and tile-major within a component;
 * a macroblock holds 1 (DC only), blocks (no high pass) or blocks*16 values,
 * block k's DC/low-pass value at offset + 16k in the last case. The arena is
 * borrowed: it must outlive every plan built from it. */
typedef struct jxrhip_arena {
  const int32_t *coefficients;
  uint64_t coefficient_count;
  uint64_t macroblock_count;
  const uint32_t *coefficient_offsets; /* [macroblock_count]                    */
  const uint32_t *quantizers;          /* [macroblock_count][3]: dc, lp, hp step */
  const uint8_t *bands;                /* [macroblock_count] JXRHIP_BANDS_*      */
  const uint8_t *hp_predictions;       /* [macroblock_count] JXRHIP_PRED_*       */
  const uint32_t *coded_x;             /* [macroblock_count] column in coded plane */
  const uint32_t *coded_y;             /* [macroblock_count] row in coded plane    */
  const jxrhip_coefficient_plane *planes;
  uint32_t plane_count;
  uint32_t overlap;    /* 0, 1 or 2 (OverlapMode)                              */
  uint32_t hard_tiles; /* nonzero: each tile is filtered on its own           */
  uint32_t tile_column_count;
  const uint32_t *tile_column_widths; /* macroblocks, whole coded image        */
  uint32_t tile_row_count;
  const uint32_t *tile_row_heights;   /* macroblocks, whole coded image        */
  uint32_t reserved0;
  /* Caller-chosen identity of immutable coefficients: the session's upload
   * cache keys device copies on it, as the reference keys on Arc identity
   * (crates/jxr-cuda/src/upload_cache.rs:15-153). 0 disables caching. */
  uint64_t identity;
} jxrhip_arena;

/* ---- output policy (crates/jxr-core/src/output_policy.rs) -------------- */
/* ColorFormat */
#define JXRHIP_COLOR_LUMA 0u
#define JXRHIP_COLOR_YUV420 1u
#define JXRHIP_COLOR_YUV422 2u
#define JXRHIP_COLOR_YUV444 3u
#define JXRHIP_COLOR_RGB 4u
#define JXRHIP_COLOR_CMYK 5u
#define JXRHIP_COLOR_CMYK_DIRECT 6u
#define JXRHIP_COLOR_YUVK 7u
#define JXRHIP_COLOR_RGBE 8u
#define JXRHIP_COLOR_NCOMPONENT 9u
/* AlphaMode */
#define JXRHIP_ALPHA_NONE 0u
#define JXRHIP_ALPHA_INTEGRATED 1u
#define JXRHIP_ALPHA_SEPARATE 2u
/* OutputBitDepth */
#define JXRHIP_DEPTH_BIT1_WHITE 0u
#define JXRHIP_DEPTH_BIT1_BLACK 1u
#define JXRHIP_DEPTH_U8 2u
#define JXRHIP_DEPTH_U10 3u
#define JXRHIP_DEPTH_U16 4u  /* shift_bits */
#define JXRHIP_DEPTH_I16 5u  /* shift_bits */
#define JXRHIP_DEPTH_F16 6u
#define JXRHIP_DEPTH_I32 7u  /* shift_bits */
#define JXRHIP_DEPTH_F32 8u  /* mantissa_length, exponent_bias */
#define JXRHIP_DEPTH_RGB555 9u
#define JXRHIP_DEPTH_RGB101010 10u
#define JXRHIP_DEPTH_RGB565 11u
/* Storage kind of one output sample (PixelFormat::storage_kind) */
#define JXRHIP_STORE_BITS 0u
#define JXRHIP_STORE_U8 1u
#define JXRHIP_STORE_U16 2u
#define JXRHIP_STORE_I16 3u
#define JXRHIP_STORE_I32 4u
#define JXRHIP_STORE_F16 5u
#define JXRHIP_STORE_F32 6u
#define JXRHIP_STORE_PACKED16 7u
#define JXRHIP_STORE_PACKED32 8u
/* ChannelLayout (codes of crates/jxr-core/src/device_plan/output.rs:269-295) */
#define JXRHIP_LAYOUT_LUMA 0u
#define JXRHIP_LAYOUT_LUMA_ALPHA 1u
#define JXRHIP_LAYOUT_YUV 2u
#define JXRHIP_LAYOUT_YUVA 3u
#define JXRHIP_LAYOUT_RGB 4u
#define JXRHIP_LAYOUT_RGBA 5u
#define JXRHIP_LAYOUT_BGR 6u
#define JXRHIP_LAYOUT_BGRA 7u
#define JXRHIP_LAYOUT_CMYK 8u
#define JXRHIP_LAYOUT_CMYKA 9u
#define JXRHIP_LAYOUT_NCOMPONENT 10u
#define JXRHIP_LAYOUT_NCOMPONENT_ALPHA 11u
#define JXRHIP_LAYOUT_RGBX 12u
#define JXRHIP_LAYOUT_BGRX 13u

/* OutputFormatRequest resolved for one image, plus the image facts the
 * output stage reads (primary colour format, chroma centering). */
typedef struct jxrhip_output_format {
  uint32_t internal_color; /* JXRHIP_COLOR_* reconstructed components       */
  uint32_t output_color;   /* JXRHIP_COLOR_* semantic output colour          */
  uint32_t bit_depth;      /* JXRHIP_DEPTH_*                                  */
  uint32_t shift_bits;
  uint32_t mantissa_length;
  int32_t exponent_bias;
  uint32_t has_alpha_format; /* nonzero: the alpha_* fields below apply      */
  uint32_t alpha_bit_depth;
  uint32_t alpha_shift_bits;
  uint32_t alpha_mantissa_length;
  int32_t alpha_exponent_bias;
  uint32_t alpha_scaled;
  uint32_t storage;        /* JXRHIP_STORE_*                                  */
  uint32_t channel_layout; /* JXRHIP_LAYOUT_*                                 */
  uint32_t channel_count;  /* samples per pixel of the layout                */
  uint32_t scaled;         /* SCALED_FLAG of the primary plane               */
  uint32_t red_blue_not_swapped;
  uint32_t premultiply_alpha;
  uint32_t chroma_centering_x; /* 0..4 (T.832 Table 181)                     */
  uint32_t chroma_centering_y;
  uint32_t crop_x;      /* in the coded plane: coded margin + region origin  */
  uint32_t crop_y;
  uint32_t crop_width;
  uint32_t crop_height;
} jxrhip_output_format;

/* One plane of an output surface. */
typedef struct jxrhip_surface_plane {
  uint64_t byte_offset;
  uint64_t row_stride_bytes;
  uint32_t width;
  uint32_t height;
  uint32_t channels;
  uint32_t reserved;
} jxrhip_surface_plane;

#define JXRHIP_MAX_SURFACE_PLANES 5u
/* SurfaceLayout. planar == 0: one interleaved plane, as the reference's
 * CUDA store writes (crates/jxr-cuda/src/kernels/output.rs:350-495).
 * planar == 1: channel c of every pixel goes to planes[c] (NCHW), which the
 * reference offers only on the CPU (crates/jxr/src/batch/cpu/output.rs:94-160). */
typedef struct jxrhip_surface_layout {
  uint32_t width;
  uint32_t height;
  uint32_t plane_count;
  uint32_t planar;
  jxrhip_surface_plane planes[JXRHIP_MAX_SURFACE_PLANES];
  uint64_t byte_len;
} jxrhip_surface_layout;

/* DecodeScale (crates/jxr-core/src/request.rs:10-40). A reduced decode
 * stops after the first-level transform (quarter: one sample per block, from
 * the DC and low-pass bands) or at the macroblock DC (sixteenth: one sample
 * per macroblock) and scales those coefficients to the samples they stand
 * for (crates/jxr-native/src/reconstruct/pipeline.rs:175-260,
 * sample_scale.rs). The reference runs it on the CPU only; its CUDA plan
 * rejects it (crates/jxr-cuda/src/plan.rs:634). With a reduced scale the
 * output crop is in the coded plane at that scale:
 * (coded margin + region origin) / 4 or / 16, as output_crop does
 * (crates/jxr-native/src/decode/output.rs:484-501). The caller's prepared
 * plan has already required frequency-mode packets and, for a quarter,
 * the low-pass band (crates/jxr-native/src/prepare.rs:154-194). */
#define JXRHIP_SCALE_FULL 0u
#define JXRHIP_SCALE_QUARTER 1u
#define JXRHIP_SCALE_SIXTEENTH 2u

/* Everything CudaDecodePlan::from_prepared takes (plan.rs:91-148), plus the
 * decode scale. */
typedef struct jxrhip_image_desc {
  uint32_t abi_version;     /* JXRHIP_ABI_VERSION                              */
  uint32_t color_format;    /* primary ColorFormat, JXRHIP_COLOR_*             */
  uint32_t component_count; /* primary components, alpha excluded              */
  uint32_t alpha_mode;      /* JXRHIP_ALPHA_*                                  */
  uint32_t scaled;          /* primary SCALED_FLAG: chroma doubled after the
                               first-level transform (jxr-core image.rs:83-85) */
  uint32_t scale;           /* JXRHIP_SCALE_*: native decode resolution       */
  const jxrhip_arena *primary;
  const jxrhip_arena *separate_alpha; /* one luma plane, or NULL             */
  jxrhip_output_format output;
  jxrhip_surface_layout surface;
} jxrhip_image_desc;

/* ---- opaque handles ---------------------------------------------------- */
typedef struct jxrhip_session jxrhip_session;
typedef struct jxrhip_plan jxrhip_plan;
typedef struct jxrhip_submission jxrhip_submission;
typedef struct jxrhip_image jxrhip_image;
typedef struct jxrhip_batch jxrhip_batch;
typedef struct jxrhip_batch_graph jxrhip_batch_graph;

/* Counters of the session's pools and caches. */
typedef struct jxrhip_diagnostics {
  uint64_t upload_hits;
  uint64_t upload_misses;
  uint64_t upload_evictions;
  uint64_t upload_resident_bytes;
  uint64_t scratch_hits;
  uint64_t scratch_misses;
  uint64_t scratch_high_water_bytes;
  uint64_t submissions;
  uint64_t kernel_launches;
  uint32_t stream_count;
  int32_t device_ordinal;
  uint32_t compute_units;
  uint32_t reserved0;
} jxrhip_diagnostics;

/* ---- library ----------------------------------------------------------- */
uint32_t jxrhip_abi_version(void);
/* Message of the last failure on this thread ("" if none). */
const char *jxrhip_last_error(void);
/* 1 when a gfx950 device is usable (CudaDecoderSession::is_available,
 * session.rs:35-39), 0 otherwise. */
int32_t jxrhip_is_available(void);

/* Route (route.rs:23-50): 1 = device, 0 = CPU. strict != 0 makes the
 * device mandatory; otherwise reconstructed_coefficients below
 * jxrhip_route_threshold() stays on the CPU. */
uint64_t jxrhip_route_threshold(void);
jxrhip_status jxrhip_route(uint64_t reconstructed_coefficients, int32_t strict,
                           int32_t device_available, int32_t *use_device);

/* ---- plans (plan.rs:91-354) -------------------------------------------- */
/* Validates the arenas, geometry, tile grids and output policy before any
 * device resource exists, and builds the device descriptors on the host. */
jxrhip_status jxrhip_plan_create(const jxrhip_image_desc *desc, jxrhip_plan **out);
void jxrhip_plan_destroy(jxrhip_plan *plan);
uint64_t jxrhip_plan_output_bytes(const jxrhip_plan *plan);
uint64_t jxrhip_plan_reconstructed_coefficients(const jxrhip_plan *plan);
uint64_t jxrhip_plan_output_pixels(const jxrhip_plan *plan);
/* Device bytes a submission of this plan holds: uploaded coefficients and
 * descriptors plus its output (the serving loop admits by this). */
uint64_t jxrhip_plan_device_footprint(const jxrhip_plan *plan);
/* Algorithmic HBM bytes of one reconstruction: 4 B per stored coefficient,
 * 32 B per macroblock per plane, plus the output bytes. */
uint64_t jxrhip_plan_algorithmic_bytes(const jxrhip_plan *plan);

/* ---- sessions (session.rs:16-268) --------------------------------------- */
jxrhip_status jxrhip_session_create(int32_t device_ordinal, jxrhip_session **out);
/* Releases the caller's handle. Refused (the session stays, and
 * jxrhip_last_error says why) while captured batches or servers of the
 * session exist: destroy those first. Submissions, batches, result images
 * and destinations of the session that are still alive keep it (its
 * streams, pool and cache) until the last of them is released, as a
 * resident image keeps its device context (resident.rs:8-12); the session
 * then waits for its streams and frees itself. The handle must not be used
 * for new work after this call. */
void jxrhip_session_destroy(jxrhip_session *session);
jxrhip_status jxrhip_session_diagnostics(jxrhip_session *session, jxrhip_diagnostics *out);
/* Upload-cache byte budget (default 64 GiB of the device's HBM). */
jxrhip_status jxrhip_session_set_upload_budget(jxrhip_session *session, uint64_t bytes);
/* Makes the coefficients and descriptors of `plan` device-resident now. */
jxrhip_status jxrhip_session_upload(jxrhip_session *session, const jxrhip_plan *plan);
/* Drops every cached upload (device memory returns to the pool). */
jxrhip_status jxrhip_session_clear_cache(jxrhip_session *session);

/* Plain per-image path: one reconstruction launch per image on one of the
 * session's streams, into a device buffer the session allocates. */
jxrhip_status jxrhip_submit(jxrhip_session *session, const jxrhip_plan *plan,
                            jxrhip_submission **out);
/* Same, into caller device memory of at least the plan's output bytes. */
jxrhip_status jxrhip_submit_into(jxrhip_session *session, const jxrhip_plan *plan,
                                 void *device_destination, uint64_t destination_bytes,
                                 jxrhip_submission **out);
int32_t jxrhip_submission_is_complete(const jxrhip_submission *submission);
/* Waits, then frees the submission. On success *image (may be NULL for
 * submit_into) receives the resident output. A nonzero device status word
 * returns JXRHIP_ERR_KERNEL_ARITHMETIC and *stage the failing stage. */
jxrhip_status jxrhip_submission_wait(jxrhip_submission *submission, jxrhip_image **image,
                                     uint32_t *stage);
/* Frees a submission, waiting first (Drop, submission.rs:132-139). */
void jxrhip_submission_destroy(jxrhip_submission *submission);

/* Consumers of a pending submission, with no host wait
 * (CudaSubmission::{enqueue_consumer_wait, output_layout,
 * pending_device_buffer}, submission.rs:50-93).
 * enqueue_wait: work queued next on `hip_stream` (a hipStream_t of the
 *   session's device and of the HIP runtime this library is bound to; NULL
 *   is the null stream) runs after the submission's reconstruction. A
 *   process that loads PyTorch-ROCm before this library binds both to the
 *   HIP runtime torch loaded (one soname), so torch's streams qualify.
 * output_layout: the validated surface layout of the output, before
 *   completion.
 * pending_device_ptr: the output buffer of a jxrhip_submit submission (not
 *   of submit_into, whose destination the caller holds). A consumer stream
 *   may read it after enqueue_wait, and must be done with it before the
 *   submission is waited for or destroyed (a wait hands the buffer to the
 *   result image, which keeps it). */
jxrhip_status jxrhip_submission_enqueue_wait(const jxrhip_submission *submission, void *hip_stream);
jxrhip_status jxrhip_submission_output_layout(const jxrhip_submission *submission, jxrhip_surface_layout *out);
jxrhip_status jxrhip_submission_pending_device_ptr(const jxrhip_submission *submission, void **device_ptr,
                                                   uint64_t *bytes);

/* ---- decode reports (crates/jxr-core/src/report.rs:9-70) --------------- */
/* DecodeStage */
#define JXRHIP_DECODE_STAGE_PARSE 0u
#define JXRHIP_DECODE_STAGE_ENTROPY_DECODE 1u
#define JXRHIP_DECODE_STAGE_INVERSE_SCAN 2u
#define JXRHIP_DECODE_STAGE_COEFFICIENT_REMAP 3u
#define JXRHIP_DECODE_STAGE_DC_LOW_PASS_PREDICTION 4u
#define JXRHIP_DECODE_STAGE_DEQUANTIZE_AND_FIRST_INVERSE_TRANSFORM 5u
#define JXRHIP_DECODE_STAGE_FIRST_OVERLAP 6u
#define JXRHIP_DECODE_STAGE_HIGH_PASS_PREDICTION 7u
#define JXRHIP_DECODE_STAGE_SECOND_INVERSE_TRANSFORM 8u
#define JXRHIP_DECODE_STAGE_SECOND_OVERLAP 9u
#define JXRHIP_DECODE_STAGE_CHROMA_RECONSTRUCTION 10u
#define JXRHIP_DECODE_STAGE_COLOR_AND_ALPHA_CONVERSION 11u
#define JXRHIP_DECODE_STAGE_CROP_CLIP_AND_PACK 12u
#define JXRHIP_DECODE_STAGE_HOST_READBACK 13u
/* StageExecutor; HIP takes the place of the reference's Cuda */
#define JXRHIP_EXECUTOR_CPU_SCALAR 0u
#define JXRHIP_EXECUTOR_CPU_SIMD 1u
#define JXRHIP_EXECUTOR_METAL 2u
#define JXRHIP_EXECUTOR_HIP 3u
/* BackendRequest / BackendKind (the caller's request, passed through) */
#define JXRHIP_BACKEND_AUTO 0u
#define JXRHIP_BACKEND_CPU 1u
#define JXRHIP_BACKEND_HIP 2u
/* FallbackReason (none for a device decode) */
#define JXRHIP_FALLBACK_NONE 0u
typedef struct jxrhip_stage_report {
  uint32_t stage;    /* JXRHIP_DECODE_STAGE_* */
  uint32_t executor; /* JXRHIP_EXECUTOR_*     */
} jxrhip_stage_report;
typedef struct jxrhip_decode_report {
  uint32_t requested; /* JXRHIP_BACKEND_* */
  uint32_t selected;  /* JXRHIP_BACKEND_HIP */
  uint32_t fallback;  /* JXRHIP_FALLBACK_NONE */
  uint32_t stage_count;
  jxrhip_stage_report stages[14];
} jxrhip_decode_report;
/* The route and per-stage ownership of a device decode of `plan`
 * (CudaDecodePlan::decode_report, plan.rs:214-253); `host_readback`
 * appends the readback stage, as decode_to_host reports. */
jxrhip_status jxrhip_plan_decode_report(const jxrhip_plan *plan, uint32_t requested, int32_t host_readback,
                                        jxrhip_decode_report *out);

/* Device destinations (CudaDecoderSession::allocate_destination /
 * allocate_batch_destination, session.rs:85-101): plain device memory of the
 * session's device, usable as submit_into / submit_batch targets. A dense
 * batch is one allocation with per-image offsets. jxrhip_free_destination
 * takes only pointers this session allocated, and works after
 * jxrhip_session_destroy (each live destination keeps the session). */
jxrhip_status jxrhip_allocate_destination(jxrhip_session *session, uint64_t bytes, void **device_ptr);
jxrhip_status jxrhip_free_destination(jxrhip_session *session, void *device_ptr);
jxrhip_status jxrhip_fill_destination(jxrhip_session *session, void *device_ptr, uint8_t value,
                                      uint64_t bytes);
jxrhip_status jxrhip_copy_to_host(jxrhip_session *session, const void *device_ptr, void *host,
                                  uint64_t bytes);
/* Page-locked host memory (hipHostMalloc) of the session's device, for
 * staging that uploads or readbacks should reach by DMA without a copy:
 * for example the coefficient storage of jxrhip_codestream_decode_into.
 * No counterpart in the reference (its uploads go through cudarc from
 * pageable memory). A block keeps its session until it is freed.
 * jxrhip_host_free takes only pointers jxrhip_host_alloc returned; the
 * caller frees a block once every submission that reads it has completed
 * (the call also waits for the session's streams). */
jxrhip_status jxrhip_host_alloc(jxrhip_session *session, uint64_t bytes, void **host_ptr);
jxrhip_status jxrhip_host_free(jxrhip_session *session, void *host_ptr);
/* Device-to-device copy of `bytes` between destinations, on the session's
 * first stream (waits for completion). */
jxrhip_status jxrhip_copy_device(jxrhip_session *session, void *dst, const void *src, uint64_t bytes);
/* Waits for all work queued on the session's streams. */
jxrhip_status jxrhip_session_synchronize(jxrhip_session *session);

/* Issue rate of the vector ALU, the reference the reconstruction kernel's
 * VALU instruction rate is read against (no counterpart in the reference:
 * measurement support for the benchmark). Runs `iterations` rounds of
 * sixteen independent integer VALU instructions per wave at four 256-thread
 * workgroups per compute unit on the session's first stream and reports the
 * timed launch. wave_instructions / (simds * ms * clock) is the peak rate in
 * wave instructions per SIMD cycle. */
typedef struct jxrhip_valu_rate {
  double ms;                 /* device time of the timed launch           */
  double wave_instructions;  /* VALU instructions it issued, per wave      */
  uint32_t simds;            /* SIMDs of the device (four per compute unit) */
  uint32_t clock_khz;        /* the device's reported peak shader clock    */
} jxrhip_valu_rate;
jxrhip_status jxrhip_calibrate_valu(jxrhip_session *session, uint32_t iterations, jxrhip_valu_rate *out);

void *jxrhip_image_device_ptr(const jxrhip_image *image);
uint64_t jxrhip_image_bytes(const jxrhip_image *image);
void jxrhip_image_destroy(jxrhip_image *image);
jxrhip_status jxrhip_readback(jxrhip_session *session, const jxrhip_image *image,
                              void *host, uint64_t host_bytes);
/* submit + wait + readback (session.rs:214-218). */
jxrhip_status jxrhip_decode_to_host(jxrhip_session *session, const jxrhip_plan *plan,
                                    void *host, uint64_t host_bytes, uint32_t *stage);

/* ---- batches (session.rs:112-211; crates/jxr/src/batch/cuda.rs) --------- */
#define JXRHIP_BATCH_PER_IMAGE 0u  /* one launch per image, streams round-robin */
#define JXRHIP_BATCH_PERSISTENT 1u /* one persistent launch per shape bucket     */
/* Submits `count` plans. Image i goes to device_destinations[i] (each at
 * least its plan's output bytes). In persistent mode the images are grouped
 * into shape buckets and each bucket runs as one persistent kernel that
 * pulls macroblock-tile work items from a device queue. */
jxrhip_status jxrhip_submit_batch(jxrhip_session *session, const jxrhip_plan *const *plans,
                                  uint32_t count, void *const *device_destinations,
                                  uint32_t mode, jxrhip_batch **out);
int32_t jxrhip_batch_is_complete(const jxrhip_batch *batch);
/* A device-side wait on `hip_stream` for every image of the batch (as
 * jxrhip_submission_enqueue_wait); the destinations may then be read on
 * that stream, with no host wait. */
jxrhip_status jxrhip_batch_enqueue_wait(const jxrhip_batch *batch, void *hip_stream);
/* Waits, frees the batch, and writes each image's stage code (0 = ok) to
 * stages[count] when non-NULL. Returns JXRHIP_ERR_KERNEL_ARITHMETIC if any
 * image failed; siblings are unaffected. */
jxrhip_status jxrhip_batch_wait(jxrhip_batch *batch, uint32_t *stages);
/* Device time of the batch's reconstruction launches, from HIP events on
 * the streams they ran on; valid after jxrhip_batch_wait. */
jxrhip_status jxrhip_last_batch_timing(jxrhip_session *session, double *milliseconds,
                                       uint32_t *launches);

/* Allocation-free persistent batch. No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture
 * set-up, tight serving loops, harnesses that forbid hipMalloc). Decodes
 * like jxrhip_submit_batch(..., JXRHIP_BATCH_PERSISTENT) + jxrhip_batch_wait
 * and returns when the images are done, writing each image's stage code
 * (0 = ok) to stages[count] when non-NULL. Its scratch (status words, share
 * counters, slot records, item table) is a static device array of 4 MiB
 * that every device the library runs on holds until the process ends; the
 * call allocates no device or pinned memory. So every plan's coefficients
 * and descriptors must be resident already (jxrhip_session_upload, or an
 * earlier submission), else JXRHIP_ERR_INVALID_ARGUMENT. Limits: 4096
 * images, 8192 plane groups, 64 shape buckets and 962,560 work items per
 * call; beyond them JXRHIP_ERR_BUDGET. The array is shared by every call on
 * a device and keeps the last call's contents (each call clears what must
 * start at zero). Calls that overlapped on one device would race for it, so
 * the library runs these calls one at a time in a process: a call made while
 * another runs waits for it. */
jxrhip_status jxrhip_decode_batch_static(jxrhip_session *session, const jxrhip_plan *const *plans,
                                         uint32_t count, void *const *device_destinations, uint32_t *stages);

/* ---- tuning ------------------------------------------------------------ */
/* Preferred macroblock tile of one work item (0 = automatic). A shape whose
 * LDS does not fit a workgroup is halved along its longer side until it
 * does. */
jxrhip_status jxrhip_session_set_tile(jxrhip_session *session, uint32_t tile_mbs_x,
                                      uint32_t tile_mbs_y);
/* Persistent grid: workgroups per compute unit (0 = automatic). */
jxrhip_status jxrhip_session_set_occupancy(jxrhip_session *session, uint32_t groups_per_cu);

/* Shape buckets. Plans with the same plane geometry (component count and
 * block layout per plane) and storage kind share a kernel instantiation and
 * a tile shape; jxrhip_plan_bucket_key names the bucket. The reference has
 * no buckets (one launch sequence per image, crates/jxr-cuda/src/encode.rs:66-159). */
uint64_t jxrhip_plan_bucket_key(const jxrhip_plan *plan);

typedef struct jxrhip_tile_choice {
  uint64_t bucket_key;
  uint32_t tile_mbs_x, tile_mbs_y; /* the fastest shape, now in use          */
  uint32_t images;                 /* images of the batch in this bucket     */
  uint32_t candidates;             /* shapes timed                           */
  double best_ms;                  /* median device time at that shape       */
  double default_ms;               /* median device time at the untuned shape*/
  uint32_t groups_per_cu;          /* workgroups per CU the persistent launch uses */
  uint32_t reserved;
} jxrhip_tile_choice;

/* Autotuner: times every candidate tile shape on each shape bucket of the
 * batch (one persistent launch of the bucket alone, `repeats` timed runs
 * after an untimed one), then, at the fastest shape, one and two fewer
 * workgroups per CU than fit; keeps the fastest per bucket key (a smaller
 * workgroup count only when more than 1% faster) for later persistent
 * submissions of this session, and the shape also for per-image ones. The
 * destinations receive the images. Writes up to `capacity` choices and the
 * bucket count. A fixed shape from jxrhip_session_set_tile, or a count from
 * jxrhip_session_set_occupancy, still takes precedence. No counterpart in
 * the reference. */
jxrhip_status jxrhip_session_autotune(jxrhip_session *session, const jxrhip_plan *const *plans, uint32_t count,
                                      void *const *device_dsts, uint32_t repeats, jxrhip_tile_choice *choices,
                                      uint32_t capacity, uint32_t *bucket_count);
/* Installs (or, with 0 x 0, removes) a bucket's tile shape, e.g. one saved
 * from an earlier autotune, with as many workgroups per CU as fit; a shape
 * whose LDS does not fit is ignored. */
jxrhip_status jxrhip_session_set_bucket_tile(jxrhip_session *session, uint64_t bucket_key, uint32_t tile_mbs_x,
                                             uint32_t tile_mbs_y);
/* The bucket's installed shape, 0 x 0 when none. */
jxrhip_status jxrhip_session_bucket_tile(jxrhip_session *session, uint64_t bucket_key, uint32_t *tile_mbs_x,
                                         uint32_t *tile_mbs_y);

/* ---- captured batches ---------------------------------------------------- */
/* A persistent batch (JXRHIP_BATCH_PERSISTENT) captured once into a HIP graph:
 * per shape bucket an item-table kernel and the persistent launch, buckets
 * spread over the session's streams. jxrhip_batch_graph_launch replays it and
 * returns a batch to wait on like jxrhip_submit_batch's. Every arena needs an
 * identity: the graph keeps those uploads and the plans' descriptors resident
 * until it is destroyed. Plans and destinations must outlive the graph. One
 * replay at a time: launching again before the previous batch was waited for
 * fails with JXRHIP_ERR_INVALID_ARGUMENT. Wait for the last replay before
 * destroying the graph, and destroy graphs before their session. Replaces
 * the reference's per-submission launch sequencing
 * (crates/jxr-cuda/src/encode.rs:66-159, session.rs:103-211) for a batch
 * decoded repeatedly into the same destinations. */
jxrhip_status jxrhip_batch_graph_create(jxrhip_session *session, const jxrhip_plan *const *plans, uint32_t count,
                                        void *const *device_dsts, jxrhip_batch_graph **out);
jxrhip_status jxrhip_batch_graph_launch(jxrhip_batch_graph *graph, jxrhip_batch **out);
void jxrhip_batch_graph_destroy(jxrhip_batch_graph *graph);

/* ---- serving loop -------------------------------------------------------- */
/* A request loop over one session: requests are admitted first come, first
 * served while the device footprint of the admitted ones (uploaded
 * coefficients and descriptors, plus the output when the server allocates
 * it) fits the budget, and run as waves, each one persistent batch on one of
 * four compute streams, with uploads on a stream of their own. A wave starts
 * whenever a compute stream is free, so requests arriving while others run
 * are admitted as memory frees up and their waves overlap on the device
 * (continuous batching). Results carry each request's status, its
 * arrival-to-resident latency and, when the request gave no destination, the
 * output image. Such an image counts against the budget until the caller
 * destroys it, so a caller that falls behind slows admission rather than
 * exhausting device memory; when nothing is in flight one request is
 * admitted regardless, so a caller that drains before polling still sees
 * every request finish. The nearest callers in the reference are the batch decoder's
 * grouped submissions (crates/jxr/src/batch/cuda.rs:20-156, indexed
 * per-image errors); admission and continuous batching have no counterpart
 * there. Plans and destinations must outlive their requests; submit and poll
 * may be called from any thread, destroy from one thread after the last
 * submit. */
typedef struct jxrhip_server jxrhip_server;

typedef struct jxrhip_server_config {
  uint64_t hbm_budget_bytes; /* admitted footprint limit (0: 8 GiB, or 3/4 of
                                free HBM if less)                              */
  uint32_t max_inflight;     /* admitted requests at once (0: 4096)           */
  uint32_t max_wave_images;  /* images per wave, up to 65536 (0: 256)         */
  uint32_t max_waves;        /* waves on the device at once, 1..4 (0: 4)      */
  uint32_t reserved0;
} jxrhip_server_config;

typedef struct jxrhip_server_result {
  uint64_t ticket;
  uint64_t tag;
  jxrhip_status status; /* OK; KERNEL_ARITHMETIC with `stage`; BUDGET when the
                           request alone exceeds the budget; DEVICE          */
  uint32_t stage;
  jxrhip_image *image;  /* output when the request had no destination; the
                           caller destroys it                                */
  double latency_ms;    /* arrival -> output resident on the device          */
  double queue_ms;      /* arrival -> admission                              */
} jxrhip_server_result;

typedef struct jxrhip_server_stats {
  uint64_t submitted, admitted, completed, failed, rejected;
  uint64_t waves, max_wave_images;
  uint64_t pixels_completed;
  uint64_t inflight_high_water_bytes; /* admitted footprint                 */
  uint64_t hbm_high_water_bytes;      /* the session's device pool          */
  double p50_ms, p99_ms;              /* latency of completed requests,
                                         nearest rank                        */
  double queue_p50_ms, queue_p99_ms;
  double mean_wave_images;
} jxrhip_server_stats;

jxrhip_status jxrhip_server_create(jxrhip_session *session, const jxrhip_server_config *config,
                                   jxrhip_server **out);
/* The steady clock that arrival times are measured on, in nanoseconds. */
uint64_t jxrhip_server_clock_ns(void);
/* Queues a request. `arrival_ns` (0: now; later than now counts as now) is
 * when the request arrived, for its latency. */
jxrhip_status jxrhip_server_submit(jxrhip_server *server, const jxrhip_plan *plan, void *device_dst,
                                   uint64_t dst_bytes, uint64_t tag, uint64_t arrival_ns, uint64_t *ticket);
/* Moves up to `capacity` finished results to `out`, waiting up to
 * `timeout_us` for the first; returns how many. */
uint32_t jxrhip_server_poll(jxrhip_server *server, jxrhip_server_result *out, uint32_t capacity,
                            uint32_t timeout_us);
/* Waits until every submitted request has a result. */
jxrhip_status jxrhip_server_drain(jxrhip_server *server);
jxrhip_status jxrhip_server_get_stats(jxrhip_server *server, jxrhip_server_stats *out);
/* Finishes every queued request, then stops; images of results never
 * polled are released. */
void jxrhip_server_destroy(jxrhip_server *server);

#ifdef __cplusplus
}
#endif
#endif /* JXR_HIP_H */
