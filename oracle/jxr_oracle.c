/*
 * jxr_oracle.c — CPU restatement of the reference CPU decoder's
 * reconstruction and output stages, used only as the checker of the HIP path
 * (parity tests, smoke check, CPU-baseline timing). Not part of the shipped
 * library.
 *
 * Followed reference files (paths relative to the reference repository):
 *   operators    crates/jxr-math/src/{transform,overlap,prediction,sampling,
 *                alpha,quantization,arithmetic}.rs
 *   stages       crates/jxr-native/src/reconstruct/{pipeline,overlap,
 *                subsampled_overlap,chroma}.rs,
 *                crates/jxr-native/src/decode/reconstruction.rs:432-613
 *   output       crates/jxr-native/src/output_format/{scaling,color,packing,
 *                premultiply}.rs
 *   geometry     crates/jxr-cuda/src/plan.rs:405-721 (windows, validation),
 *                crates/jxr-core/src/device_plan/overlap.rs:65-253 (regions)
 *
 * Parity pins (tests/test_oracle.py): every known-answer vector the
 * reference's unit tests hold for these operators and stages, forward/inverse
 * round trips of every operator, lossless decode(analysis(x)) == x, brute
 * force against i64 formulas, and interval bounds. The output stage's
 * formats are pinned by the reference's scaling/alpha vectors plus closed
 * forms; whole-image composition on real codestreams has no reference vector
 * offline (restatement vs restatement).
 */
#include "jxr_oracle.h"

#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* checked i32 arithmetic (crates/jxr-math/src/arithmetic.rs:58-84)        */
/* ---------------------------------------------------------------------- */
static inline int32_t fit(int64_t r, int *e) {
  if (r > INT32_MAX || r < INT32_MIN) {
    *e = 1;
    return 0;
  }
  return (int32_t)r;
}
#define ADD(a, b) fit((int64_t)(a) + (int64_t)(b), e)
#define SUB(a, b) fit((int64_t)(a) - (int64_t)(b), e)
#define MUL(a, b) fit((int64_t)(a) * (int64_t)(b), e)
#define NEG(a) fit(-(int64_t)(a), e)
#define SWAP(a, b)      \
  do {                  \
    int32_t t_ = (a);   \
    (a) = (b);          \
    (b) = t_;           \
  } while (0)

/* ---------------------------------------------------------------------- */
/* core transform (crates/jxr-math/src/transform.rs:8-347)                */
/* ---------------------------------------------------------------------- */
/* crates/jxr-math/data/reconstruction.abi:2 */
static const int PERM[16] = {0, 8, 4, 13, 2, 15, 3, 14, 1, 12, 5, 9, 7, 11, 6, 10};

typedef void (*op4)(int32_t *v, int *e);

static void group4(int32_t *c, int a, int b, int d, int f, op4 op, int *e) {
  int32_t v[4] = {c[a], c[b], c[d], c[f]};
  op(v, e);
  c[a] = v[0];
  c[b] = v[1];
  c[d] = v[2];
  c[f] = v[3];
}

/* transform.rs:274-285 */
static void t2x2h(int32_t *v, int32_t r, int *e) {
  v[0] = ADD(v[0], v[3]);
  v[1] = SUB(v[1], v[2]);
  int32_t d = SUB(v[0], v[1]);
  int32_t first = ADD(d, r) >> 1;
  int32_t second = v[2];
  v[2] = SUB(first, v[3]);
  v[3] = SUB(first, second);
  v[0] = SUB(v[0], v[3]);
  v[1] = ADD(v[1], v[2]);
}
static void t2x2h_round(int32_t *v, int *e) { t2x2h(v, 1, e); }
static void t2x2h_zero(int32_t *v, int *e) { t2x2h(v, 0, e); }

/* transform.rs:287-321 */
static void inverse_todd(int32_t *v, int *e) {
  v[1] = ADD(v[1], v[3]);
  v[0] = SUB(v[0], v[2]);
  v[3] = SUB(v[3], v[1] >> 1);
  v[2] = ADD(v[2], ADD(v[0], 1) >> 1);
  v[0] = SUB(v[0], ADD(MUL(v[1], 3), 4) >> 3);
  v[1] = ADD(v[1], ADD(MUL(v[0], 3), 4) >> 3);
  v[2] = SUB(v[2], ADD(MUL(v[3], 3), 4) >> 3);
  v[3] = ADD(v[3], ADD(MUL(v[2], 3), 4) >> 3);
  v[2] = SUB(v[2], ADD(v[1], 1) >> 1);
  v[3] = SUB(ADD(v[0], 1) >> 1, v[3]);
  v[1] = ADD(v[1], v[2]);
  v[0] = SUB(v[0], v[3]);
}

/* transform.rs:323-361 */
static void inverse_todd_odd(int32_t *v, int *e) {
  v[3] = ADD(v[3], v[0]);
  v[2] = SUB(v[2], v[1]);
  int32_t first = v[3] >> 1;
  int32_t second = v[2] >> 1;
  v[0] = SUB(v[0], first);
  v[1] = ADD(v[1], second);
  v[0] = SUB(v[0], ADD(MUL(v[1], 3), 3) >> 3);
  v[1] = ADD(v[1], ADD(MUL(v[0], 3), 3) >> 2);
  v[0] = SUB(v[0], ADD(MUL(v[1], 3), 4) >> 3);
  v[1] = SUB(v[1], second);
  v[0] = ADD(v[0], first);
  v[2] = ADD(v[2], v[1]);
  v[3] = SUB(v[3], v[0]);
  v[1] = NEG(v[1]);
  v[2] = NEG(v[2]);
}

/* transform.rs:363-398 */
static void forward_todd(int32_t *v, int *e) {
  v[0] = ADD(v[0], v[3]);
  v[1] = SUB(v[1], v[2]);
  v[3] = SUB(ADD(v[0], 1) >> 1, v[3]);
  v[2] = ADD(v[2], ADD(v[1], 1) >> 1);
  v[3] = SUB(v[3], ADD(MUL(v[2], 3), 4) >> 3);
  v[2] = ADD(v[2], ADD(MUL(v[3], 3), 4) >> 3);
  v[1] = SUB(v[1], ADD(MUL(v[0], 3), 4) >> 3);
  v[0] = ADD(v[0], ADD(MUL(v[1], 3), 4) >> 3);
  v[2] = SUB(v[2], ADD(v[0], 1) >> 1);
  v[3] = ADD(v[3], v[1] >> 1);
  v[0] = ADD(v[0], v[2]);
  v[1] = SUB(v[1], v[3]);
}

/* transform.rs:400-440 */
static void forward_todd_odd(int32_t *v, int *e) {
  v[1] = NEG(v[1]);
  v[2] = NEG(v[2]);
  v[3] = ADD(v[3], v[0]);
  v[2] = SUB(v[2], v[1]);
  int32_t first = v[3] >> 1;
  int32_t second = v[2] >> 1;
  v[0] = SUB(v[0], first);
  v[1] = ADD(v[1], second);
  v[0] = ADD(v[0], ADD(MUL(v[1], 3), 4) >> 3);
  v[1] = SUB(v[1], ADD(MUL(v[0], 3), 3) >> 2);
  v[0] = ADD(v[0], ADD(MUL(v[1], 3), 3) >> 3);
  v[1] = SUB(v[1], second);
  v[0] = ADD(v[0], first);
  v[2] = ADD(v[2], v[1]);
  v[3] = SUB(v[3], v[0]);
}

/* transform.rs:8-22 */
int jxro_inverse_core_transform(int32_t c[16]) {
  int err = 0, *e = &err;
  int32_t in[16];
  memcpy(in, c, sizeof in);
  for (int i = 0; i < 16; i++) c[PERM[i]] = in[i];
  group4(c, 0, 1, 4, 5, t2x2h_round, e);
  group4(c, 2, 3, 6, 7, inverse_todd, e);
  group4(c, 8, 12, 9, 13, inverse_todd, e);
  group4(c, 10, 11, 14, 15, inverse_todd_odd, e);
  group4(c, 0, 3, 12, 15, t2x2h_zero, e);
  group4(c, 5, 6, 9, 10, t2x2h_zero, e);
  group4(c, 1, 2, 13, 14, t2x2h_zero, e);
  group4(c, 4, 7, 8, 11, t2x2h_zero, e);
  return err;
}

/* transform.rs:29-41 */
int jxro_forward_core_transform(int32_t c[16]) {
  int err = 0, *e = &err;
  group4(c, 4, 7, 8, 11, t2x2h_zero, e);
  group4(c, 1, 2, 13, 14, t2x2h_zero, e);
  group4(c, 5, 6, 9, 10, t2x2h_zero, e);
  group4(c, 0, 3, 12, 15, t2x2h_zero, e);
  group4(c, 10, 11, 14, 15, forward_todd_odd, e);
  group4(c, 8, 12, 9, 13, forward_todd, e);
  group4(c, 2, 3, 6, 7, forward_todd, e);
  group4(c, 0, 1, 4, 5, t2x2h_round, e);
  int32_t in[16];
  memcpy(in, c, sizeof in);
  for (int i = 0; i < 16; i++) c[i] = in[PERM[i]];
  return err;
}

/* transform.rs:52-57 */
int jxro_inverse_chroma_420(int32_t v[4]) {
  int err = 0, *e = &err;
  t2x2h(v, 0, e);
  SWAP(v[1], v[2]);
  return err;
}
/* transform.rs:98-102 */
int jxro_forward_chroma_420(int32_t v[4]) {
  int err = 0, *e = &err;
  SWAP(v[1], v[2]);
  t2x2h(v, 0, e);
  return err;
}

/* transform.rs:138-157 */
static void inverse_t2pt(int32_t *p, int *e) {
  p[0] = SUB(p[0], ADD(p[1], 1) >> 1);
  p[1] = ADD(p[1], p[0]);
}
static void forward_t2pt(int32_t *p, int *e) {
  p[1] = SUB(p[1], p[0]);
  p[0] = ADD(p[0], ADD(p[1], 1) >> 1);
}

/* transform.rs:60-89 */
int jxro_inverse_chroma_422(int32_t v[8]) {
  int err = 0, *e = &err;
  int32_t pair[2] = {v[0], v[4]};
  inverse_t2pt(pair, e);
  v[0] = pair[0];
  v[4] = pair[1];
  t2x2h(v, 0, e); /* v[0..4] */
  SWAP(v[1], v[2]);
  int32_t s[4] = {v[4], v[6], v[5], v[7]};
  t2x2h(s, 0, e);
  v[4] = s[0];
  v[6] = s[1];
  v[5] = s[2];
  v[7] = s[3];
  SWAP(v[5], v[6]);
  return err;
}
/* transform.rs:108-131 */
int jxro_forward_chroma_422(int32_t v[8]) {
  int err = 0, *e = &err;
  SWAP(v[5], v[6]);
  int32_t s[4] = {v[4], v[6], v[5], v[7]};
  t2x2h(s, 0, e);
  v[4] = s[0];
  v[6] = s[1];
  v[5] = s[2];
  v[7] = s[3];
  SWAP(v[1], v[2]);
  t2x2h(v, 0, e);
  int32_t pair[2] = {v[0], v[4]};
  forward_t2pt(pair, e);
  v[0] = pair[0];
  v[4] = pair[1];
  return err;
}

/* ---------------------------------------------------------------------- */
/* overlap filters (crates/jxr-math/src/overlap.rs:10-444)                */
/* ---------------------------------------------------------------------- */
/* overlap.rs:266-278 */
static void inverse_rotate(int32_t *a, int32_t *b, int *e) {
  *a = SUB(*a, ADD(*b, 1) >> 1);
  *b = ADD(*b, ADD(*a, 1) >> 1);
}
/* overlap.rs:281-293 */
static void forward_rotate(int32_t *a, int32_t *b, int *e) {
  *b = SUB(*b, ADD(*a, 1) >> 1);
  *a = ADD(*a, ADD(*b, 1) >> 1);
}
/* overlap.rs:379-393 */
static void inverse_scale(int32_t *a, int32_t *b, int *e) {
  *a = ADD(*a, *b);
  *b = SUB(*a >> 1, *b);
  *a = ADD(*a, MUL(*b, 3) >> 3);
  *b = ADD(*b, MUL(*a, 3) >> 4);
  *b = ADD(*b, *a >> 7);
  *b = SUB(*b, *a >> 10);
}
/* overlap.rs:296-310 */
static void forward_scale(int32_t *a, int32_t *b, int *e) {
  *b = ADD(*b, *a >> 10);
  *b = SUB(*b, *a >> 7);
  *b = SUB(*b, MUL(*a, 3) >> 4);
  *a = SUB(*a, MUL(*b, 3) >> 3);
  *b = SUB(*a >> 1, *b);
  *a = SUB(*a, *b);
}

static void inverse_4_impl(int32_t *v, int *e) {
  v[0] = ADD(v[0], v[3]);
  v[1] = ADD(v[1], v[2]);
  v[3] = SUB(v[3], ADD(v[0], 1) >> 1);
  v[2] = SUB(v[2], ADD(v[1], 1) >> 1);
  inverse_scale(&v[0], &v[3], e);
  inverse_scale(&v[1], &v[2], e);
  v[0] = ADD(v[0], ADD(MUL(v[3], 3), 4) >> 3);
  v[1] = ADD(v[1], ADD(MUL(v[2], 3), 4) >> 3);
  v[3] = SUB(v[3], v[0] >> 1);
  v[2] = SUB(v[2], v[1] >> 1);
  v[0] = ADD(v[0], v[3]);
  v[1] = ADD(v[1], v[2]);
  v[3] = NEG(v[3]);
  v[2] = NEG(v[2]);
  inverse_rotate(&v[2], &v[3], e);
  v[3] = ADD(v[3], ADD(v[0], 1) >> 1);
  v[2] = ADD(v[2], ADD(v[1], 1) >> 1);
  v[0] = SUB(v[0], v[3]);
  v[1] = SUB(v[1], v[2]);
}
/* overlap.rs:10-53 */
int jxro_inverse_overlap_4(int32_t v[4]) {
  int err = 0;
  inverse_4_impl(v, &err);
  return err;
}
/* overlap.rs:59-102 */
int jxro_forward_overlap_4(int32_t v[4]) {
  int err = 0, *e = &err;
  v[1] = ADD(v[1], v[2]);
  v[0] = ADD(v[0], v[3]);
  v[2] = SUB(v[2], ADD(v[1], 1) >> 1);
  v[3] = SUB(v[3], ADD(v[0], 1) >> 1);
  forward_rotate(&v[2], &v[3], e);
  v[2] = NEG(v[2]);
  v[3] = NEG(v[3]);
  v[1] = SUB(v[1], v[2]);
  v[0] = SUB(v[0], v[3]);
  v[2] = ADD(v[2], v[1] >> 1);
  v[3] = ADD(v[3], v[0] >> 1);
  v[1] = SUB(v[1], ADD(MUL(v[2], 3), 4) >> 3);
  v[0] = SUB(v[0], ADD(MUL(v[3], 3), 4) >> 3);
  forward_scale(&v[1], &v[2], e);
  forward_scale(&v[0], &v[3], e);
  v[2] = ADD(v[2], ADD(v[1], 1) >> 1);
  v[3] = ADD(v[3], ADD(v[0], 1) >> 1);
  v[1] = SUB(v[1], v[2]);
  v[0] = SUB(v[0], v[3]);
  return err;
}

/* overlap.rs:395-409 */
static void inverse_hadamard_post(int32_t *v, int *e) {
  v[1] = SUB(v[1], v[2]);
  v[0] = ADD(v[0], ADD(MUL(v[3], 3), 4) >> 3);
  v[3] = SUB(v[3], v[1] >> 1);
  v[2] = SUB(SUB(v[0], v[1]) >> 1, v[2]);
  SWAP(v[2], v[3]);
  v[0] = SUB(v[0], v[3]);
  v[1] = ADD(v[1], v[2]);
}
/* overlap.rs:312-328 */
static void forward_hadamard_pre(int32_t *v, int *e) {
  v[1] = SUB(v[1], v[2]);
  v[0] = ADD(v[0], v[3]);
  SWAP(v[2], v[3]);
  v[2] = SUB(SUB(v[0], v[1]) >> 1, v[2]);
  v[3] = ADD(v[3], v[1] >> 1);
  v[0] = SUB(v[0], ADD(MUL(v[3], 3), 4) >> 3);
  v[1] = ADD(v[1], v[2]);
}
/* overlap.rs:411-436 */
static void inverse_todd_odd_post(int32_t *v, int *e) {
  v[3] = ADD(v[3], v[0]);
  v[2] = SUB(v[2], v[1]);
  int32_t first = v[3] >> 1;
  int32_t second = v[2] >> 1;
  v[0] = SUB(v[0], first);
  v[1] = ADD(v[1], second);
  v[0] = SUB(v[0], ADD(MUL(v[1], 3), 6) >> 3);
  v[1] = ADD(v[1], ADD(MUL(v[0], 3), 2) >> 2);
  v[0] = SUB(v[0], ADD(MUL(v[1], 3), 4) >> 3);
  v[1] = SUB(v[1], second);
  v[0] = ADD(v[0], first);
  v[2] = ADD(v[2], v[1]);
  v[3] = SUB(v[3], v[0]);
}
/* overlap.rs:330-356 */
static void forward_todd_odd_pre(int32_t *v, int *e) {
  v[3] = ADD(v[3], v[0]);
  v[2] = SUB(v[2], v[1]);
  int32_t first = v[3] >> 1;
  int32_t second = v[2] >> 1;
  v[0] = SUB(v[0], first);
  v[1] = ADD(v[1], second);
  v[0] = ADD(v[0], ADD(MUL(v[1], 3), 4) >> 3);
  v[1] = SUB(v[1], ADD(MUL(v[0], 3), 2) >> 2);
  v[0] = ADD(v[0], ADD(MUL(v[1], 3), 6) >> 3);
  v[1] = SUB(v[1], second);
  v[0] = ADD(v[0], first);
  v[2] = ADD(v[2], v[1]);
  v[3] = SUB(v[3], v[0]);
}

static const int GROUPS[4][4] = {{0, 3, 12, 15}, {1, 2, 13, 14}, {4, 7, 8, 11}, {5, 6, 9, 10}};
static const int ROTATIONS[4][2] = {{13, 12}, {9, 8}, {7, 3}, {6, 2}};
static const int SCALES[4][2] = {{0, 15}, {1, 14}, {4, 11}, {5, 10}};

/* overlap.rs:139-161 */
int jxro_inverse_overlap_4x4(int32_t v[16]) {
  int err = 0, *e = &err;
  for (int g = 0; g < 4; g++)
    group4(v, GROUPS[g][0], GROUPS[g][1], GROUPS[g][2], GROUPS[g][3], t2x2h_zero, e);
  for (int r = 0; r < 4; r++) inverse_rotate(&v[ROTATIONS[r][0]], &v[ROTATIONS[r][1]], e);
  group4(v, 10, 11, 14, 15, inverse_todd_odd_post, e);
  for (int s = 0; s < 4; s++) inverse_scale(&v[SCALES[s][0]], &v[SCALES[s][1]], e);
  for (int g = 0; g < 4; g++)
    group4(v, GROUPS[g][0], GROUPS[g][1], GROUPS[g][2], GROUPS[g][3], inverse_hadamard_post, e);
  return err;
}
/* overlap.rs:108-136 */
int jxro_forward_overlap_4x4(int32_t v[16]) {
  int err = 0, *e = &err;
  for (int g = 0; g < 4; g++)
    group4(v, GROUPS[g][0], GROUPS[g][1], GROUPS[g][2], GROUPS[g][3], forward_hadamard_pre, e);
  for (int s = 0; s < 4; s++) forward_scale(&v[SCALES[s][0]], &v[SCALES[s][1]], e);
  group4(v, 10, 11, 14, 15, forward_todd_odd_pre, e);
  for (int r = 0; r < 4; r++) forward_rotate(&v[ROTATIONS[r][0]], &v[ROTATIONS[r][1]], e);
  for (int g = 0; g < 4; g++)
    group4(v, GROUPS[g][0], GROUPS[g][1], GROUPS[g][2], GROUPS[g][3], t2x2h_zero, e);
  return err;
}

static void inverse_2_impl(int32_t *v, int *e) {
  v[1] = ADD(v[1], ADD(v[0], 2) >> 2);
  v[0] = ADD(v[0], ADD(v[1], 1) >> 1);
  v[0] = ADD(v[0], v[1] >> 5);
  v[0] = ADD(v[0], v[1] >> 9);
  v[0] = ADD(v[0], v[1] >> 13);
  v[1] = ADD(v[1], ADD(v[0], 2) >> 2);
}
static void forward_2_impl(int32_t *v, int *e) {
  v[1] = SUB(v[1], ADD(v[0], 2) >> 2);
  v[0] = SUB(v[0], v[1] >> 13);
  v[0] = SUB(v[0], v[1] >> 9);
  v[0] = SUB(v[0], v[1] >> 5);
  v[0] = SUB(v[0], ADD(v[1], 1) >> 1);
  v[1] = SUB(v[1], ADD(v[0], 2) >> 2);
}
/* overlap.rs:210-227 */
int jxro_inverse_overlap_2(int32_t v[2]) {
  int err = 0;
  inverse_2_impl(v, &err);
  return err;
}
/* overlap.rs:254-263 */
int jxro_forward_overlap_2(int32_t v[2]) {
  int err = 0;
  forward_2_impl(v, &err);
  return err;
}
/* overlap.rs:167-204 */
int jxro_inverse_overlap_2x2(int32_t v[4]) {
  int err = 0, *e = &err;
  v[0] = ADD(v[0], v[3]);
  v[1] = ADD(v[1], v[2]);
  v[3] = SUB(v[3], ADD(v[0], 1) >> 1);
  v[2] = SUB(v[2], ADD(v[1], 1) >> 1);
  inverse_2_impl(v, e);
  v[3] = ADD(v[3], ADD(v[0], 1) >> 1);
  v[2] = ADD(v[2], ADD(v[1], 1) >> 1);
  v[0] = SUB(v[0], v[3]);
  v[1] = SUB(v[1], v[2]);
  return err;
}
/* overlap.rs:233-250 */
int jxro_forward_overlap_2x2(int32_t v[4]) {
  int err = 0, *e = &err;
  v[1] = ADD(v[1], v[2]);
  v[0] = ADD(v[0], v[3]);
  v[2] = SUB(v[2], ADD(v[1], 1) >> 1);
  v[3] = SUB(v[3], ADD(v[0], 1) >> 1);
  forward_2_impl(v, e);
  v[2] = ADD(v[2], ADD(v[1], 1) >> 1);
  v[3] = ADD(v[3], ADD(v[0], 1) >> 1);
  v[1] = SUB(v[1], v[2]);
  v[0] = SUB(v[0], v[3]);
  return err;
}

/* ---------------------------------------------------------------------- */
/* chroma upsampling (crates/jxr-math/src/sampling.rs:29-58,105-115)      */
/* ---------------------------------------------------------------------- */
static const int32_t CENTERING[5][4] = {
    {4, 4, 0, 8}, {5, 3, 1, 7}, {6, 2, 2, 6}, {7, 1, 3, 5}, {8, 0, 4, 4}};

/* The T.832 Table 180 sum, taken exactly in 64 bits. */
static int32_t weighted(int32_t a, int32_t wa, int32_t b, int32_t wb) {
  int64_t s = (int64_t)a * wa + (int64_t)b * wb + 4;
  return (int32_t)(s >> 3);
}
int jxro_upsample_pair(int32_t previous, int32_t current, int32_t next, uint32_t centering,
                       int32_t out[2]) {
  if (centering > 4) return 1;
  const int32_t *h = CENTERING[centering];
  out[0] = weighted(previous, h[2], current, h[3]);
  out[1] = weighted(current, h[0], next, h[1]);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* high-pass prediction (crates/jxr-math/src/prediction.rs:53-97)         */
/* ---------------------------------------------------------------------- */
int jxro_predict_high_pass(int32_t c[256], uint32_t direction, uint32_t cols, uint32_t rows) {
  int err = 0, *e = &err;
  if (direction == JXRHIP_PRED_FROM_LEFT) {
    for (uint32_t r = 0; r < rows; r++)
      for (uint32_t col = 1; col < cols; col++) {
        uint32_t b = r * cols + col;
        for (int p = 4; p <= 12; p += 4) c[b * 16 + p] = ADD(c[b * 16 + p], c[(b - 1) * 16 + p]);
      }
  } else if (direction == JXRHIP_PRED_FROM_TOP) {
    for (uint32_t r = 1; r < rows; r++)
      for (uint32_t col = 0; col < cols; col++) {
        uint32_t b = r * cols + col;
        for (int p = 1; p <= 3; p++) c[b * 16 + p] = ADD(c[b * 16 + p], c[(b - cols) * 16 + p]);
      }
  } else if (direction != JXRHIP_PRED_NONE) {
    return 2;
  }
  return err;
}

/* ---------------------------------------------------------------------- */
/* alpha premultiplication (crates/jxr-math/src/alpha.rs:6-54)            */
/* ---------------------------------------------------------------------- */
int jxro_premultiply_unsigned(uint32_t value, uint32_t alpha, uint32_t maximum, uint32_t *out) {
  if (maximum == 0) return 1;
  if (value > maximum) value = maximum;
  if (alpha > maximum) alpha = maximum;
  uint64_t p = (uint64_t)value * alpha + maximum / 2;
  *out = (uint32_t)(p / maximum);
  return 0;
}
int jxro_premultiply_signed(int32_t value, int32_t alpha, int32_t maximum, int32_t *out) {
  if (maximum <= 0) return 1;
  if (alpha < 0) alpha = 0;
  if (alpha > maximum) alpha = maximum;
  uint64_t mag = value < 0 ? (uint64_t)(-(int64_t)value) : (uint64_t)value;
  uint64_t m = (mag * (uint64_t)alpha + (uint64_t)maximum / 2) / (uint64_t)maximum;
  int64_t s = value < 0 ? -(int64_t)m : (int64_t)m;
  if (s > INT32_MAX || s < INT32_MIN) return 1;
  *out = (int32_t)s;
  return 0;
}
static int premultiply_sm15(uint16_t value, uint16_t alpha, uint16_t *out) {
  uint32_t a = (alpha & 0x8000) ? 0 : (alpha & 0x7fffu);
  uint32_t m;
  if (jxro_premultiply_unsigned(value & 0x7fffu, a, 0x7fff, &m)) return 1;
  *out = (uint16_t)((value & 0x8000) | m);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* output scaling (crates/jxr-native/src/output_format/scaling.rs)        */
/* ---------------------------------------------------------------------- */
typedef struct {
  int invalid; /* ChannelScale::Invalid: every apply fails */
  int padding;
  int32_t shifted_bias;
  int32_t rounding;
  uint32_t shift;
  int has_post;
  uint32_t post_shift;
} channel_scale;

/* scaling.rs:299-318 */
static int depth_bias(uint32_t depth, uint32_t shift_bits, int32_t *bias) {
  int32_t base = 0;
  uint32_t shift = 0;
  switch (depth) {
    case JXRHIP_DEPTH_RGB555: base = 1 << 4; break;
    case JXRHIP_DEPTH_RGB565: base = 1 << 5; break;
    case JXRHIP_DEPTH_U8: base = 1 << 7; break;
    case JXRHIP_DEPTH_U10:
    case JXRHIP_DEPTH_RGB101010: base = 1 << 9; break;
    case JXRHIP_DEPTH_U16: base = 1 << 15; shift = shift_bits; break;
    case JXRHIP_DEPTH_I16:
    case JXRHIP_DEPTH_I32: base = 0; shift = shift_bits; break;
    default: base = 0; break;
  }
  if (shift >= 32) return 1;
  *bias = base >> shift;
  return 0;
}

/* scaling.rs:42-97 */
static channel_scale resolve_scale(uint32_t component, uint32_t output_color, uint32_t depth,
                                   uint32_t shift_bits, uint32_t scaled) {
  channel_scale s;
  memset(&s, 0, sizeof s);
  int32_t bias;
  if (depth_bias(depth, shift_bits, &bias)) {
    s.invalid = 1;
    return s;
  }
  int32_t cb = bias;
  if (output_color == JXRHIP_COLOR_CMYK)
    cb = component < 3 ? (bias >> 1) : -(bias >> 1);
  else if (output_color == JXRHIP_COLOR_RGBE)
    cb = 0;
  uint32_t scale = scaled ? 3 : 0;
  /* checked_shift_left (scaling.rs:321-331) */
  int64_t sh = (int64_t)cb * ((int64_t)1 << scale);
  if (sh > INT32_MAX || sh < INT32_MIN) {
    s.invalid = 1;
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
    s.has_post = 1;
    s.post_shift = shift_bits;
  }
  return s;
}

/* scaling.rs:319-325: value << bits, held at the i32 limit. */
static int32_t saturating_shl(int32_t v, uint32_t bits) {
  if (bits > 31) bits = 31;
  int64_t r = (int64_t)v * ((int64_t)1 << bits);
  if (r > INT32_MAX) return INT32_MAX;
  if (r < INT32_MIN) return INT32_MIN;
  return (int32_t)r;
}

/* scaling.rs:104-124 */
static int apply_scale(const channel_scale *s, int32_t sample, int32_t *out) {
  if (s->padding) {
    *out = 0;
    return 0;
  }
  if (s->invalid) return 1;
  int err = 0, *e = &err;
  int32_t biased = ADD(sample, s->shifted_bias);
  int32_t r = ADD(biased, s->rounding);
  if (err) return 1;
  int32_t v = r >> s->shift;
  *out = s->has_post ? saturating_shl(v, s->post_shift) : v;
  return 0;
}

int jxro_float16_bits(int32_t sample, uint32_t scaled, uint16_t *bits) {
  channel_scale s = resolve_scale(0, JXRHIP_COLOR_LUMA, JXRHIP_DEPTH_F16, 0, scaled);
  int32_t v;
  if (apply_scale(&s, sample, &v)) return 1;
  uint16_t sign = v < 0 ? 0x8000 : 0;
  uint64_t mag = v < 0 ? (uint64_t)(-(int64_t)v) : (uint64_t)v;
  if (mag > 32767) mag = 32767;
  *bits = (uint16_t)(sign | mag);
  return 0;
}

/* scaling.rs:254-296 */
int jxro_float32_bits(int32_t sample, uint32_t scaled, uint32_t length, int32_t exponent_bias,
                      uint32_t *bits) {
  if (length > 23) return 2;
  channel_scale s = resolve_scale(0, JXRHIP_COLOR_LUMA, JXRHIP_DEPTH_F32, 0, scaled);
  int32_t v;
  if (apply_scale(&s, sample, &v)) return 1;
  uint32_t sign = v < 0 ? 0x80000000u : 0;
  uint64_t mag = v < 0 ? (uint64_t)(-(int64_t)v) : (uint64_t)v;
  uint64_t implicit = (uint64_t)1 << length;
  int64_t exponent = (int64_t)(mag >> length);
  uint64_t mantissa = (mag & (implicit - 1)) | implicit;
  if (exponent == 0) {
    mantissa ^= implicit;
    exponent = 1;
  }
  exponent = exponent - exponent_bias + 127;
  while (mantissa < implicit && exponent > 1 && mantissa > 0) {
    exponent -= 1;
    mantissa <<= 1;
  }
  if (mantissa < implicit)
    exponent = 0;
  else
    mantissa ^= implicit;
  mantissa <<= 23 - length;
  if (exponent < 0 || exponent > 0xff) {
    *bits = sign | 0x7f7fffffu; /* f32::MAX: clipped, scaling.rs:286-291 */
    return 0;
  }
  if (mantissa > 0x7fffff) return 1;
  *bits = sign | ((uint32_t)exponent << 23) | (uint32_t)mantissa;
  return 0;
}

/* ---------------------------------------------------------------------- */
/* reconstruction of one plane                                            */
/* ---------------------------------------------------------------------- */
typedef struct {
  const jxrhip_arena *arena;
  jxrhip_coefficient_plane cp;
  uint32_t bc, br;
  uint32_t mb_ox, mb_oy, mbx, mby;
  uint32_t resolution; /* JXRHIP_SCALE_* */
  uint32_t ux, uy;     /* samples across and down one macroblock at that scale */
  uint32_t *raster; /* [mbx*mby] -> arena macroblock index */
  int scale_after_first;
  int alpha;
  int32_t *low;
  uint32_t lw, lh;
  int32_t *samples;
  uint32_t sw, sh;
  uint32_t *col_bounds; /* MB units, relative to window, incl. 0 and extent */
  uint32_t ncol;
  uint32_t *row_bounds;
  uint32_t nrow;
} plane_t;

static void plane_free(plane_t *p) {
  free(p->raster);
  free(p->low);
  free(p->samples);
  free(p->col_bounds);
  free(p->row_bounds);
  memset(p, 0, sizeof *p);
}

#define FAIL(code, ...)                                         \
  do {                                                          \
    if (msg) snprintf(msg, msglen, __VA_ARGS__);                \
    return (code);                                              \
  } while (0)

/* crates/jxr-core/src/device_plan/overlap.rs:102-123 (clipped_axis) */
static int clipped_axis(const uint32_t *sizes, uint32_t n, uint32_t origin, uint32_t extent,
                        uint32_t **out, uint32_t *count) {
  uint64_t end = (uint64_t)origin + extent;
  uint32_t *b = (uint32_t *)malloc(sizeof(uint32_t) * (n + 2));
  uint32_t k = 0;
  b[k++] = 0;
  uint64_t pos = 0;
  for (uint32_t i = 0; i < n; i++) {
    pos += sizes[i];
    if (pos > origin && pos < end) b[k++] = (uint32_t)(pos - origin);
  }
  b[k++] = extent;
  if (pos < end) {
    free(b);
    return 1;
  }
  *out = b;
  *count = k;
  return 0;
}

/* Window of one coefficient plane and its raster map
 * (crates/jxr-cuda/src/plan.rs:495-574). */
static int plane_setup(plane_t *p, const jxrhip_arena *a, uint32_t index, int scale, int alpha,
                       uint32_t resolution, char *msg, size_t msglen) {
  memset(p, 0, sizeof *p);
  p->resolution = resolution;
  if (index >= a->plane_count) FAIL(JXRHIP_ERR_INVALID_PLAN, "coefficient plane index is absent");
  p->arena = a;
  p->cp = a->planes[index];
  p->bc = p->cp.block_columns;
  p->br = p->cp.block_rows;
  p->scale_after_first = scale;
  p->alpha = alpha;
  if (!((p->bc == 4 && p->br == 4) || (p->bc == 2 && p->br == 4) || (p->bc == 2 && p->br == 2)))
    FAIL(JXRHIP_ERR_INVALID_PLAN, "unsupported component macroblock geometry");
  uint64_t mo = p->cp.macroblock_offset, mc = p->cp.macroblock_count;
  /* checked additions, as check_range (jxr-core/src/plan.rs:193-207) */
  if (mc == 0 || mo + mc < mo || mo + mc > a->macroblock_count)
    FAIL(JXRHIP_ERR_INVALID_PLAN, "macroblock plane range");
  uint64_t co = p->cp.coefficient_offset, cc = p->cp.coefficient_count;
  if (co + cc < co || co + cc > a->coefficient_count)
    FAIL(JXRHIP_ERR_INVALID_PLAN, "coefficient plane range");
  uint32_t x0 = UINT32_MAX, y0 = UINT32_MAX, x1 = 0, y1 = 0;
  for (uint64_t i = mo; i < mo + mc; i++) {
    if (a->coded_x[i] < x0) x0 = a->coded_x[i];
    if (a->coded_y[i] < y0) y0 = a->coded_y[i];
    if (a->coded_x[i] > x1) x1 = a->coded_x[i];
    if (a->coded_y[i] > y1) y1 = a->coded_y[i];
  }
  p->mb_ox = x0;
  p->mb_oy = y0;
  p->mbx = x1 - x0 + 1;
  p->mby = y1 - y0 + 1;
  if ((uint64_t)p->mbx * p->mby != mc)
    FAIL(JXRHIP_ERR_INVALID_PLAN, "coefficient plane reconstruction window is not rectangular");
  p->raster = (uint32_t *)malloc(sizeof(uint32_t) * mc);
  for (uint64_t i = 0; i < mc; i++) p->raster[i] = UINT32_MAX;
  for (uint64_t i = mo; i < mo + mc; i++) {
    uint64_t r = (uint64_t)(a->coded_y[i] - y0) * p->mbx + (a->coded_x[i] - x0);
    if (p->raster[r] != UINT32_MAX) {
      plane_free(p);
      FAIL(JXRHIP_ERR_INVALID_PLAN, "coefficient plane repeats a macroblock coordinate");
    }
    p->raster[r] = (uint32_t)i;
  }
  /* macroblock spans (plan.rs:681-721) and nonzero steps
   * (pipeline.rs:705-707: a zero step has no inverse) */
  uint32_t blocks = p->bc * p->br;
  for (uint64_t i = mo; i < mo + mc; i++) {
    uint32_t bands = a->bands[i];
    if (bands > 3) {
      plane_free(p);
      FAIL(JXRHIP_ERR_INVALID_PLAN, "band presence code");
    }
    uint64_t need = bands == 0 ? 1 : bands == 1 ? blocks : (uint64_t)blocks * 16;
    if ((uint64_t)a->coefficient_offsets[i] + need > a->coefficient_count) {
      plane_free(p);
      FAIL(JXRHIP_ERR_INVALID_PLAN, "macroblock coefficient span exceeds the arena");
    }
    if (a->hp_predictions[i] > JXRHIP_PRED_FROM_TOP) {
      plane_free(p);
      FAIL(JXRHIP_ERR_INVALID_PLAN, "top-left high-pass prediction is invalid");
    }
    if (a->quantizers[3 * i] == 0 || a->quantizers[3 * i + 1] == 0 || a->quantizers[3 * i + 2] == 0) {
      plane_free(p);
      FAIL(JXRHIP_ERR_INVALID_PLAN, "zero quantizer step");
    }
  }
  if (a->hard_tiles) {
    if (clipped_axis(a->tile_column_widths, a->tile_column_count, p->mb_ox, p->mbx, &p->col_bounds,
                     &p->ncol) ||
        clipped_axis(a->tile_row_heights, a->tile_row_count, p->mb_oy, p->mby, &p->row_bounds,
                     &p->nrow)) {
      plane_free(p);
      FAIL(JXRHIP_ERR_INVALID_PLAN, "tile partition does not cover reconstruction window");
    }
  } else {
    p->col_bounds = (uint32_t *)malloc(2 * sizeof(uint32_t));
    p->row_bounds = (uint32_t *)malloc(2 * sizeof(uint32_t));
    p->col_bounds[0] = 0;
    p->col_bounds[1] = p->mbx;
    p->row_bounds[0] = 0;
    p->row_bounds[1] = p->mby;
    p->ncol = p->nrow = 2;
  }
  p->lw = p->mbx * p->bc;
  p->lh = p->mby * p->br;
  /* a quarter keeps one sample per block, a sixteenth one per macroblock
   * (pipeline.rs:198-260) */
  p->ux = resolution == JXRHIP_SCALE_QUARTER ? p->bc : resolution == JXRHIP_SCALE_SIXTEENTH ? 1 : p->bc * 4;
  p->uy = resolution == JXRHIP_SCALE_QUARTER ? p->br : resolution == JXRHIP_SCALE_SIXTEENTH ? 1 : p->br * 4;
  p->sw = p->mbx * p->ux;
  p->sh = p->mby * p->uy;
  p->low = (int32_t *)calloc((size_t)p->lw * p->lh, sizeof(int32_t));
  p->samples = (int32_t *)calloc((size_t)p->sw * p->sh, sizeof(int32_t));
  if (!p->low || !p->samples) {
    plane_free(p);
    FAIL(JXRHIP_ERR_INVALID_PLAN, "out of host memory");
  }
  return 0;
}

static int dequant(int32_t v, uint32_t step, int *e) {
  return fit((int64_t)v * (int64_t)step, e);
}

/* Dequantize DC/LP and run the first-level inverse transform of every
 * macroblock (pipeline.rs:605-680; kernels/reconstruction.rs:6-64). With
 * dc_only the low-pass band is taken as zero, as a sixteenth-size decode
 * reads it (LowBands::Dc, pipeline.rs:363-372, 488-506). */
static int first_level(plane_t *p, int dc_only) {
  const jxrhip_arena *a = p->arena;
  uint32_t blocks = p->bc * p->br;
  for (uint32_t my = 0; my < p->mby; my++)
    for (uint32_t mx = 0; mx < p->mbx; mx++) {
      uint32_t i = p->raster[(uint64_t)my * p->mbx + mx];
      int err = 0, *e = &err;
      const int32_t *c = a->coefficients + a->coefficient_offsets[i];
      uint32_t bands = a->bands[i];
      int32_t low[16] = {0};
      low[0] = dequant(c[0], a->quantizers[3 * i], e);
      if (bands != JXRHIP_BANDS_DC_ONLY && !dc_only)
        for (uint32_t k = 1; k < blocks; k++)
          low[k] = dequant(c[bands >= JXRHIP_BANDS_NO_FLEXBITS ? 16 * k : k], a->quantizers[3 * i + 1], e);
      if (p->bc == 4)
        err |= jxro_inverse_core_transform(low);
      else if (p->br == 4)
        err |= jxro_inverse_chroma_422(low);
      else
        err |= jxro_inverse_chroma_420(low);
      if (p->scale_after_first)
        for (uint32_t k = 0; k < blocks; k++) low[k] = MUL(low[k], 2);
      if (err) return 1;
      for (uint32_t r = 0; r < p->br; r++)
        for (uint32_t col = 0; col < p->bc; col++)
          p->low[(uint64_t)(my * p->br + r) * p->lw + mx * p->bc + col] = low[r * p->bc + col];
    }
  return 0;
}

static int filter_indices4(int32_t *s, uint64_t i0, uint64_t i1, uint64_t i2, uint64_t i3) {
  int32_t v[4] = {s[i0], s[i1], s[i2], s[i3]};
  int err = jxro_inverse_overlap_4(v);
  s[i0] = v[0];
  s[i1] = v[1];
  s[i2] = v[2];
  s[i3] = v[3];
  return err;
}

/* Full-resolution postfilter schedule over one region
 * (reconstruct/overlap.rs:124-323; device_plan/overlap.rs:125-178). */
static int full_region(int32_t *s, uint64_t stride, uint32_t L, uint32_t R, uint32_t T, uint32_t B) {
  int err = 0;
  for (uint32_t y = T + 2; y + 2 < B; y += 4)
    for (uint32_t x = L + 2; x + 2 < R; x += 4) {
      int32_t v[16];
      for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) v[r * 4 + c] = s[(y + r) * stride + x + c];
      err |= jxro_inverse_overlap_4x4(v);
      for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) s[(y + r) * stride + x + c] = v[r * 4 + c];
    }
  for (uint32_t y = T + 2; y + 2 < B; y += 4) {
    uint32_t xs[4] = {L, L + 1, R - 2, R - 1};
    for (int k = 0; k < 4; k++)
      err |= filter_indices4(s, y * stride + xs[k], (y + 1) * stride + xs[k],
                             (y + 2) * stride + xs[k], (y + 3) * stride + xs[k]);
  }
  for (uint32_t x = L + 2; x + 2 < R; x += 4) {
    uint32_t ys[4] = {T, T + 1, B - 2, B - 1};
    for (int k = 0; k < 4; k++) {
      uint64_t b = ys[k] * stride + x;
      err |= filter_indices4(s, b, b + 1, b + 2, b + 3);
    }
  }
  uint32_t cx[4] = {L, R - 2, L, R - 2}, cy[4] = {T, T, B - 2, B - 2};
  for (int k = 0; k < 4; k++) {
    uint64_t b = cy[k] * stride + cx[k];
    err |= filter_indices4(s, b, b + 1, b + stride, b + stride + 1);
  }
  return err;
}

static int pair2(int32_t *s, uint64_t i0, uint64_t i1) {
  int32_t v[2] = {s[i0], s[i1]};
  int err = jxro_inverse_overlap_2(v);
  s[i0] = v[0];
  s[i1] = v[1];
  return err;
}

/* Ordered steps of one boundary row of a subsampled region
 * (subsampled_overlap.rs:234-283). */
static int boundary_row(int32_t *s, uint64_t stride, uint32_t L, uint32_t R, uint32_t row) {
  int err = 0, *e = &err;
  uint64_t c0 = row * stride + L, n0 = row * stride + L + 1;
  uint64_t c1 = row * stride + R - 1, n1 = row * stride + R - 2;
  s[c0] = SUB(s[c0], s[n0]);
  s[c1] = SUB(s[c1], s[n1]);
  for (uint32_t x = L + 2; x < R; x += 2) err |= pair2(s, row * stride + x - 1, row * stride + x);
  s[c0] = ADD(s[c0], s[n0]);
  s[c1] = ADD(s[c1], s[n1]);
  return err;
}

static int subsampled_region(int32_t *s, uint64_t stride, uint32_t L, uint32_t R, uint32_t T,
                             uint32_t B) {
  int err = 0;
  err |= boundary_row(s, stride, L, R, T);
  for (uint32_t y = T + 2; y < B; y += 2) {
    for (uint32_t x = L + 2; x < R; x += 2) {
      int32_t v[4] = {s[(y - 1) * stride + x - 1], s[(y - 1) * stride + x], s[y * stride + x - 1],
                      s[y * stride + x]};
      err |= jxro_inverse_overlap_2x2(v);
      s[(y - 1) * stride + x - 1] = v[0];
      s[(y - 1) * stride + x] = v[1];
      s[y * stride + x - 1] = v[2];
      s[y * stride + x] = v[3];
    }
    err |= pair2(s, (y - 1) * stride + L, y * stride + L);
    err |= pair2(s, (y - 1) * stride + R - 1, y * stride + R - 1);
  }
  err |= boundary_row(s, stride, L, R, B - 1);
  return err;
}

static int overlap_regions(int32_t *s, uint32_t w, uint32_t h, const uint32_t *cb, uint32_t nc,
                           const uint32_t *rb, uint32_t nr, int hard, int subsampled) {
  uint32_t fc[2] = {0, w}, fr[2] = {0, h};
  if (!hard) {
    cb = fc;
    nc = 2;
    rb = fr;
    nr = 2;
  }
  int err = 0;
  for (uint32_t j = 0; j + 1 < nr; j++)
    for (uint32_t i = 0; i + 1 < nc; i++) {
      if (subsampled)
        err |= subsampled_region(s, w, cb[i], cb[i + 1], rb[j], rb[j + 1]);
      else
        err |= full_region(s, w, cb[i], cb[i + 1], rb[j], rb[j + 1]);
    }
  return err;
}

int jxro_overlap_plane(int32_t *samples, uint32_t width, uint32_t height, const uint32_t *cb,
                       uint32_t nc, const uint32_t *rb, uint32_t nr, int32_t hard) {
  return overlap_regions(samples, width, height, cb, nc, rb, nr, hard, 0);
}
int jxro_subsampled_overlap_plane(int32_t *samples, uint32_t width, uint32_t height,
                                  const uint32_t *cb, uint32_t nc, const uint32_t *rb,
                                  uint32_t nr, int32_t hard) {
  return overlap_regions(samples, width, height, cb, nc, rb, nr, hard, 1);
}

static int plane_overlap(plane_t *p, int level_low) {
  uint32_t sx = level_low ? p->bc : p->bc * 4, sy = level_low ? p->br : p->br * 4;
  uint32_t *cb = (uint32_t *)malloc(sizeof(uint32_t) * p->ncol);
  uint32_t *rb = (uint32_t *)malloc(sizeof(uint32_t) * p->nrow);
  for (uint32_t i = 0; i < p->ncol; i++) cb[i] = p->col_bounds[i] * sx;
  for (uint32_t i = 0; i < p->nrow; i++) rb[i] = p->row_bounds[i] * sy;
  int err;
  if (level_low)
    err = overlap_regions(p->low, p->lw, p->lh, cb, p->ncol, rb, p->nrow, 1, p->bc != 4);
  else
    err = overlap_regions(p->samples, p->sw, p->sh, cb, p->ncol, rb, p->nrow, 1, 0);
  free(cb);
  free(rb);
  return err;
}

/* HP unpack, prediction, dequantization and second-level transform
 * (decode/reconstruction.rs:432-613; pipeline.rs:543-586). */
static int second_level(plane_t *p, uint32_t *stage) {
  const jxrhip_arena *a = p->arena;
  uint32_t blocks = p->bc * p->br;
  for (uint32_t my = 0; my < p->mby; my++)
    for (uint32_t mx = 0; mx < p->mbx; mx++) {
      uint32_t i = p->raster[(uint64_t)my * p->mbx + mx];
      const int32_t *c = a->coefficients + a->coefficient_offsets[i];
      int32_t high[256];
      memset(high, 0, sizeof high);
      if (a->bands[i] >= JXRHIP_BANDS_NO_FLEXBITS)
        for (uint32_t k = 0; k < blocks; k++)
          for (int j = 1; j < 16; j++) high[k * 16 + j] = c[k * 16 + j];
      if (jxro_predict_high_pass(high, a->hp_predictions[i], p->bc, p->br)) {
        *stage = JXRHIP_STAGE_PREDICTION;
        return 1;
      }
      for (uint32_t k = 0; k < blocks; k++) {
        int err = 0, *e = &err;
        uint32_t bx = mx * p->bc + k % p->bc, by = my * p->br + k / p->bc;
        int32_t v[16];
        v[0] = p->low[(uint64_t)by * p->lw + bx];
        for (int j = 1; j < 16; j++) v[j] = dequant(high[k * 16 + j], a->quantizers[3 * i + 2], e);
        err |= jxro_inverse_core_transform(v);
        if (err) {
          *stage = JXRHIP_STAGE_SECOND_TRANSFORM;
          return 1;
        }
        for (int r = 0; r < 4; r++)
          for (int col = 0; col < 4; col++)
            p->samples[(uint64_t)(by * 4 + r) * p->sw + bx * 4 + col] = v[r * 4 + col];
      }
    }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* reduced resolution (reconstruct/pipeline.rs:175-260,363-486;            */
/* reconstruct/sample_scale.rs)                                            */
/* ---------------------------------------------------------------------- */
/* Regions of one to five macroblocks have every arrangement an overlap
 * operator can meet (sample_scale.rs:43-52), and the flat probe levels are
 * 0x10001..0x10004 at both signs (sample_scale.rs:33-41). */
#define PROBE_EXTENT 5u
static const int32_t PROBE_LEVELS[4] = {0x10001, 0x10002, 0x10003, 0x10004};

/* probe_offset (sample_scale.rs:406-417) */
static uint32_t probe_offset(uint32_t offset, uint32_t extent) {
  const uint32_t ends = PROBE_EXTENT / 2;
  if (extent <= PROBE_EXTENT || offset < ends) return offset;
  if (extent - offset <= ends) return PROBE_EXTENT - (extent - offset);
  return ends;
}

typedef struct {
  uint32_t bounds[PROBE_EXTENT + 2]; /* probe regions, incl. 0 and the extent */
  uint32_t nb;
  uint32_t mbs;                      /* macroblocks along the probe            */
  uint32_t *pos;                     /* probe macroblock of each macroblock    */
  uint32_t npos;
} probe_axis_t;

/* ProbeAxis::new (sample_scale.rs:336-381): one probe region for each
 * extent a region of the component has, every extent of five or more
 * counting as five; each macroblock maps to the probe macroblock placed in
 * its region as it is placed in its own. */
static int probe_axis(const uint32_t *regions, uint32_t n, probe_axis_t *a) {
  memset(a, 0, sizeof *a);
  if (n < 2 || regions[n - 1] <= regions[0]) return 1;
  a->pos = (uint32_t *)malloc(sizeof(uint32_t) * (regions[n - 1] - regions[0]));
  if (!a->pos) return 1;
  int64_t starts[PROBE_EXTENT + 1];
  for (uint32_t k = 0; k <= PROBE_EXTENT; k++) starts[k] = -1;
  a->bounds[a->nb++] = 0;
  for (uint32_t i = 0; i + 1 < n; i++) {
    if (regions[i + 1] <= regions[i]) {
      free(a->pos);
      a->pos = NULL;
      return 1;
    }
    uint32_t extent = regions[i + 1] - regions[i];
    uint32_t pe = extent < PROBE_EXTENT ? extent : PROBE_EXTENT;
    if (starts[pe] < 0) {
      starts[pe] = a->mbs;
      a->mbs += pe;
      a->bounds[a->nb++] = a->mbs;
    }
    for (uint32_t off = 0; off < extent; off++) a->pos[a->npos++] = (uint32_t)starts[pe] + probe_offset(off, extent);
  }
  return 0;
}

int jxro_probe_axis(const uint32_t *regions, uint32_t count, uint32_t *bounds, uint32_t *bound_count,
                    uint32_t *positions, uint32_t capacity) {
  probe_axis_t a;
  if (probe_axis(regions, count, &a)) return 1;
  int r = a.npos > capacity;
  if (!r) {
    memcpy(bounds, a.bounds, sizeof(uint32_t) * a.nb);
    *bound_count = a.nb;
    memcpy(positions, a.pos, sizeof(uint32_t) * a.npos);
  }
  free(a.pos);
  return r;
}

/* Forward windows of the full-resolution schedule: the postfilter's
 * windows are disjoint, so the prefilter runs them in any order. */
static int forward_indices4(int32_t *s, uint64_t i0, uint64_t i1, uint64_t i2, uint64_t i3) {
  int32_t v[4] = {s[i0], s[i1], s[i2], s[i3]};
  int err = jxro_forward_overlap_4(v);
  s[i0] = v[0];
  s[i1] = v[1];
  s[i2] = v[2];
  s[i3] = v[3];
  return err;
}
static int forward_full_region(int32_t *s, uint64_t stride, uint32_t L, uint32_t R, uint32_t T, uint32_t B) {
  int err = 0;
  for (uint32_t y = T + 2; y + 2 < B; y += 4)
    for (uint32_t x = L + 2; x + 2 < R; x += 4) {
      int32_t v[16];
      for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) v[r * 4 + c] = s[(y + r) * stride + x + c];
      err |= jxro_forward_overlap_4x4(v);
      for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) s[(y + r) * stride + x + c] = v[r * 4 + c];
    }
  for (uint32_t y = T + 2; y + 2 < B; y += 4) {
    uint32_t xs[4] = {L, L + 1, R - 2, R - 1};
    for (int k = 0; k < 4; k++)
      err |= forward_indices4(s, y * stride + xs[k], (y + 1) * stride + xs[k], (y + 2) * stride + xs[k],
                              (y + 3) * stride + xs[k]);
  }
  for (uint32_t x = L + 2; x + 2 < R; x += 4) {
    uint32_t ys[4] = {T, T + 1, B - 2, B - 1};
    for (int k = 0; k < 4; k++) {
      uint64_t b = ys[k] * stride + x;
      err |= forward_indices4(s, b, b + 1, b + 2, b + 3);
    }
  }
  uint32_t cx[4] = {L, R - 2, L, R - 2}, cy[4] = {T, T, B - 2, B - 2};
  for (int k = 0; k < 4; k++) {
    uint64_t b = cy[k] * stride + cx[k];
    err |= forward_indices4(s, b, b + 1, b + stride, b + stride + 1);
  }
  return err;
}

static int forward_pair2(int32_t *s, uint64_t i0, uint64_t i1) {
  int32_t v[2] = {s[i0], s[i1]};
  int err = jxro_forward_overlap_2(v);
  s[i0] = v[0];
  s[i1] = v[1];
  return err;
}
/* The ordered subsampled steps of one boundary row undone last to first. */
static int forward_boundary_row(int32_t *s, uint64_t stride, uint32_t L, uint32_t R, uint32_t row) {
  int err = 0, *e = &err;
  uint64_t c0 = row * stride + L, n0 = row * stride + L + 1;
  uint64_t c1 = row * stride + R - 1, n1 = row * stride + R - 2;
  s[c1] = SUB(s[c1], s[n1]);
  s[c0] = SUB(s[c0], s[n0]);
  for (uint32_t x = L + 2; x < R; x += 2) err |= forward_pair2(s, row * stride + x - 1, row * stride + x);
  s[c1] = ADD(s[c1], s[n1]);
  s[c0] = ADD(s[c0], s[n0]);
  return err;
}
/* subsampled_region run backwards: its boundary rows are the first and last
 * steps, and the steps between touch disjoint samples. */
static int forward_subsampled_region(int32_t *s, uint64_t stride, uint32_t L, uint32_t R, uint32_t T,
                                     uint32_t B) {
  int err = forward_boundary_row(s, stride, L, R, B - 1);
  for (uint32_t y = T + 2; y < B; y += 2) {
    for (uint32_t x = L + 2; x < R; x += 2) {
      int32_t v[4] = {s[(y - 1) * stride + x - 1], s[(y - 1) * stride + x], s[y * stride + x - 1],
                      s[y * stride + x]};
      err |= jxro_forward_overlap_2x2(v);
      s[(y - 1) * stride + x - 1] = v[0];
      s[(y - 1) * stride + x] = v[1];
      s[y * stride + x - 1] = v[2];
      s[y * stride + x] = v[3];
    }
    err |= forward_pair2(s, (y - 1) * stride + L, y * stride + L);
    err |= forward_pair2(s, (y - 1) * stride + R - 1, y * stride + R - 1);
  }
  err |= forward_boundary_row(s, stride, L, R, T);
  return err;
}

/* Forward analysis at unit steps of a component of mbx x mby macroblocks
 * whose samples are `samples` (row-major, mbx*bc*4 across), with hard
 * regions at the macroblock boundaries cb x rb: the stages reconstruction
 * inverts, run backwards (analyze.rs:56-86): sample prefilter, block
 * transforms, low-pass prefilter, first-level transforms. A component whose
 * low bands are doubled after the first level divides them by two,
 * rounding the magnitude half up (analyze.rs:500-530 at unit steps,
 * quantization.rs:27-48). Writes 256 coefficients per macroblock, raster
 * order, block k's at 16k with its DC/low-pass coefficient at 16k. */
int jxro_analyze_plane(const int32_t *samples, uint32_t mbx, uint32_t mby, uint32_t bc, uint32_t br,
                       int32_t doubled, uint32_t overlap, const uint32_t *cb, uint32_t nc, const uint32_t *rb,
                       uint32_t nr, int32_t *out) {
  if (!((bc == 4 && br == 4) || (bc == 2 && br == 4) || (bc == 2 && br == 2)) || nc < 2 || nr < 2) return 1;
  const uint32_t sw = mbx * bc * 4, sh = mby * br * 4, lw = mbx * bc, lh = mby * br;
  int32_t *s = (int32_t *)malloc(sizeof(int32_t) * sw * sh);
  int32_t *low = (int32_t *)malloc(sizeof(int32_t) * lw * lh);
  int32_t *high = (int32_t *)malloc(sizeof(int32_t) * 16 * lw * lh);
  int err = !s || !low || !high;
  if (!err) {
    memcpy(s, samples, sizeof(int32_t) * sw * sh);
    for (uint32_t j = 0; j + 1 < nr && overlap >= 1; j++)
      for (uint32_t i = 0; i + 1 < nc; i++)
        err |= forward_full_region(s, sw, cb[i] * bc * 4, cb[i + 1] * bc * 4, rb[j] * br * 4, rb[j + 1] * br * 4);
    for (uint32_t by = 0; by < lh; by++)
      for (uint32_t bx = 0; bx < lw; bx++) {
        int32_t v[16];
        for (int r = 0; r < 4; r++)
          for (int c = 0; c < 4; c++) v[r * 4 + c] = s[(uint64_t)(by * 4 + r) * sw + bx * 4 + c];
        err |= jxro_forward_core_transform(v);
        low[(uint64_t)by * lw + bx] = v[0];
        memcpy(high + 16 * ((uint64_t)by * lw + bx), v, sizeof v);
      }
    for (uint32_t j = 0; j + 1 < nr && overlap == 2; j++)
      for (uint32_t i = 0; i + 1 < nc; i++) {
        uint32_t L = cb[i] * bc, R = cb[i + 1] * bc, T = rb[j] * br, B = rb[j + 1] * br;
        err |= bc == 4 ? forward_full_region(low, lw, L, R, T, B) : forward_subsampled_region(low, lw, L, R, T, B);
      }
    for (uint32_t my = 0; my < mby; my++)
      for (uint32_t mx = 0; mx < mbx; mx++) {
        int32_t v[16] = {0};
        for (uint32_t r = 0; r < br; r++)
          for (uint32_t c = 0; c < bc; c++) v[r * bc + c] = low[(uint64_t)(my * br + r) * lw + mx * bc + c];
        if (bc == 4)
          err |= jxro_forward_core_transform(v);
        else if (br == 4)
          err |= jxro_forward_chroma_422(v);
        else
          err |= jxro_forward_chroma_420(v);
        int32_t *m = out + ((uint64_t)my * mbx + mx) * 256;
        memset(m, 0, 256 * sizeof(int32_t));
        for (uint32_t k = 0; k < bc * br; k++) {
          uint32_t bx = mx * bc + k % bc, by = my * br + k / bc;
          memcpy(m + 16 * k, high + 16 * ((uint64_t)by * lw + bx), 16 * sizeof(int32_t));
          int32_t q = v[k];
          if (doubled) {
            uint32_t mag = q < 0 ? 0u - (uint32_t)q : (uint32_t)q;
            uint32_t h = mag / 2 + (mag & 1u);
            q = q < 0 ? -(int32_t)h : (int32_t)h;
          }
          m[16 * k] = q;
        }
      }
  }
  free(s);
  free(low);
  free(high);
  return err;
}

/* The DC and low-pass coefficients of the flat probe component of `level`
 * (16 a macroblock: block k's at k). */
static int analyze_flat(int32_t level, uint32_t bc, uint32_t br, int doubled, uint32_t overlap,
                        const probe_axis_t *cols, const probe_axis_t *rows, int32_t *out) {
  const uint32_t mbx = cols->mbs, mby = rows->mbs, n = mbx * bc * 4 * mby * br * 4;
  int32_t *flat = (int32_t *)malloc(sizeof(int32_t) * n);
  int32_t *full = (int32_t *)malloc(sizeof(int32_t) * 256 * mbx * mby);
  int err = !flat || !full;
  if (!err) {
    for (uint32_t i = 0; i < n; i++) flat[i] = level;
    err = jxro_analyze_plane(flat, mbx, mby, bc, br, doubled, overlap, cols->bounds, cols->nb, rows->bounds,
                             rows->nb, full);
    for (uint32_t m = 0; m < mbx * mby && !err; m++)
      for (uint32_t k = 0; k < 16; k++) out[m * 16 + k] = full[m * 256 + 16 * k];
  }
  free(flat);
  free(full);
  return err;
}

/* The unscaled reduced picture of a plane: the block DCs after the
 * first-level transform and its overlap filter (a quarter, block_dc_plane);
 * at a sixteenth the dequantized macroblock DC (doubled where the plane's
 * low bands are), or under two overlap levels the quarter picture of the DC
 * band summed over each macroblock (macroblock_dc_plane, sum_macroblocks). */
static int reduced_picture(plane_t *p, uint32_t *stage) {
  const jxrhip_arena *a = p->arena;
  const uint32_t overlap = a->overlap;
  if (p->resolution == JXRHIP_SCALE_SIXTEENTH && overlap != 2) {
    for (uint32_t m = 0; m < p->mbx * p->mby; m++) {
      uint32_t i = p->raster[m];
      int err = 0, *e = &err;
      int32_t dc = dequant(a->coefficients[a->coefficient_offsets[i]], a->quantizers[3 * i], e);
      if (p->scale_after_first) dc = MUL(dc, 2);
      if (err) {
        *stage = JXRHIP_STAGE_FIRST_TRANSFORM;
        return 1;
      }
      p->samples[m] = dc;
    }
    return 0;
  }
  const int dc_only = p->resolution == JXRHIP_SCALE_SIXTEENTH;
  if (first_level(p, dc_only)) {
    *stage = JXRHIP_STAGE_FIRST_TRANSFORM;
    return 1;
  }
  if (overlap == 2 && plane_overlap(p, 1)) {
    *stage = JXRHIP_STAGE_OVERLAP;
    return 1;
  }
  if (!dc_only) {
    memcpy(p->samples, p->low, sizeof(int32_t) * p->lw * p->lh);
    return 0;
  }
  for (uint32_t my = 0; my < p->mby; my++)
    for (uint32_t mx = 0; mx < p->mbx; mx++) {
      int err = 0, *e = &err;
      int32_t sum = 0;
      for (uint32_t r = 0; r < p->br; r++)
        for (uint32_t c = 0; c < p->bc; c++) sum = ADD(sum, p->low[(uint64_t)(my * p->br + r) * p->lw + mx * p->bc + c]);
      if (err) {
        *stage = JXRHIP_STAGE_FIRST_TRANSFORM;
        return 1;
      }
      p->samples[(uint64_t)my * p->mbx + mx] = sum;
    }
  return 0;
}

/* The unscaled reduced picture of the flat probe component of `level`,
 * geometry `cols` x `rows` (hard regions at the probe's boundaries), with
 * plane p's block shape, doubling, overlap and scale. */
static int probe_picture(const plane_t *p, const probe_axis_t *cols, const probe_axis_t *rows, int32_t level,
                         int32_t *out) {
  const uint32_t mbs = cols->mbs * rows->mbs, blocks = p->bc * p->br;
  int32_t *flat = (int32_t *)malloc(sizeof(int32_t) * 16 * mbs);
  int32_t *coef = (int32_t *)malloc(sizeof(int32_t) * blocks * mbs);
  uint32_t *offsets = (uint32_t *)malloc(sizeof(uint32_t) * mbs);
  uint32_t *quant = (uint32_t *)malloc(sizeof(uint32_t) * 3 * mbs);
  uint8_t *bands = (uint8_t *)malloc(mbs), *preds = (uint8_t *)malloc(mbs);
  uint32_t *cx = (uint32_t *)malloc(sizeof(uint32_t) * mbs), *cy = (uint32_t *)malloc(sizeof(uint32_t) * mbs);
  uint32_t tw[PROBE_EXTENT + 1], th[PROBE_EXTENT + 1];
  int r = !flat || !coef || !offsets || !quant || !bands || !preds || !cx || !cy;
  if (!r) r = analyze_flat(level, p->bc, p->br, p->scale_after_first, p->arena->overlap, cols, rows, flat);
  if (!r) {
    for (uint32_t m = 0; m < mbs; m++) {
      for (uint32_t k = 0; k < blocks; k++) coef[m * blocks + k] = flat[m * 16 + k];
      offsets[m] = m * blocks;
      quant[3 * m] = quant[3 * m + 1] = quant[3 * m + 2] = 1;
      bands[m] = JXRHIP_BANDS_NO_HIGH_PASS;
      preds[m] = JXRHIP_PRED_NONE;
      cx[m] = m % cols->mbs;
      cy[m] = m / cols->mbs;
    }
    for (uint32_t i = 0; i + 1 < cols->nb; i++) tw[i] = cols->bounds[i + 1] - cols->bounds[i];
    for (uint32_t j = 0; j + 1 < rows->nb; j++) th[j] = rows->bounds[j + 1] - rows->bounds[j];
    jxrhip_coefficient_plane cp = {0, (uint64_t)blocks * mbs, 0, mbs, p->bc, p->br};
    jxrhip_arena a;
    memset(&a, 0, sizeof a);
    a.coefficients = coef;
    a.coefficient_count = (uint64_t)blocks * mbs;
    a.macroblock_count = mbs;
    a.coefficient_offsets = offsets;
    a.quantizers = quant;
    a.bands = bands;
    a.hp_predictions = preds;
    a.coded_x = cx;
    a.coded_y = cy;
    a.planes = &cp;
    a.plane_count = 1;
    a.overlap = p->arena->overlap;
    a.hard_tiles = 1;
    a.tile_column_count = cols->nb - 1;
    a.tile_column_widths = tw;
    a.tile_row_count = rows->nb - 1;
    a.tile_row_heights = th;
    plane_t q;
    r = plane_setup(&q, &a, 0, p->scale_after_first, p->alpha, p->resolution, NULL, 0);
    if (!r) {
      uint32_t st = 0;
      r = reduced_picture(&q, &st);
      if (!r) memcpy(out, q.samples, sizeof(int32_t) * q.sw * q.sh);
      plane_free(&q);
    }
  }
  free(flat);
  free(coef);
  free(offsets);
  free(quant);
  free(bands);
  free(preds);
  free(cx);
  free(cy);
  return r;
}

/* floor(n / d) for d > 0 (i64::div_euclid) */
static int64_t floor_div(int64_t n, int64_t d) {
  int64_t q = n / d;
  return (n % d < 0) ? q - 1 : q;
}

/* SampleScale (sample_scale.rs:157-285): without overlap the gain of the
 * transform stages left out, as a shift; with it, the line through the
 * reduced pictures of flat probes at opposite levels, per probe sample,
 * each sample of the plane taking the probe sample placed like it. */
static int scale_reduced(plane_t *p, uint32_t *stage) {
  const uint32_t n = p->sw * p->sh;
  const uint32_t mb_gain = (p->bc == 4 && p->br == 4) ? 2 : 1;
  const uint32_t bits = p->resolution == JXRHIP_SCALE_QUARTER ? 2 : mb_gain + 2;
  if (p->arena->overlap == 0) {
    for (uint32_t i = 0; i < n; i++) p->samples[i] >>= bits;
    return 0;
  }
  probe_axis_t cols, rows;
  if (probe_axis(p->col_bounds, p->ncol, &cols)) {
    *stage = JXRHIP_STAGE_SAMPLE_SCALE;
    return 1;
  }
  if (probe_axis(p->row_bounds, p->nrow, &rows)) {
    free(cols.pos);
    *stage = JXRHIP_STAGE_SAMPLE_SCALE;
    return 1;
  }
  const uint32_t ux = p->ux, uy = p->uy;
  const uint32_t stride = cols.mbs * ux, length = stride * rows.mbs * uy;
  int64_t *rise = (int64_t *)calloc(length, sizeof(int64_t));
  int64_t *offset = (int64_t *)calloc(length, sizeof(int64_t));
  int32_t *up = (int32_t *)malloc(sizeof(int32_t) * length), *down = (int32_t *)malloc(sizeof(int32_t) * length);
  int r = !rise || !offset || !up || !down;
  int64_t span = 0, count = 0;
  for (int k = 0; k < 4 && !r; k++) {
    r = probe_picture(p, &cols, &rows, PROBE_LEVELS[k], up) || probe_picture(p, &cols, &rows, -PROBE_LEVELS[k], down);
    for (uint32_t i = 0; i < length && !r; i++) {
      rise[i] += (int64_t)up[i] - down[i];
      offset[i] += (int64_t)up[i] + down[i];
    }
    span += 2 * (int64_t)PROBE_LEVELS[k];
    count += 2;
  }
  const int64_t gain = count * span;
  for (uint32_t y = 0; y < p->sh && !r; y++) {
    const uint32_t py = rows.pos[y / uy] * uy + y % uy;
    for (uint32_t x = 0; x < p->sw && !r; x++) {
      const uint32_t px = cols.pos[x / ux] * ux + x % ux;
      const int64_t rs = rise[(uint64_t)py * stride + px], os = offset[(uint64_t)py * stride + px];
      /* FlatTerms (sample_scale.rs:228-245): divisor = count * rise > 0,
       * bias = offset * span - divisor / 2; level = floor((c * gain - bias)
       * / divisor), zero for zero (sample_scale.rs:300-315) */
      const int64_t divisor = count * rs;
      if (divisor <= 0) {
        r = 1;
        break;
      }
      const int64_t bias = os * span - divisor / 2;
      int32_t *v = &p->samples[(uint64_t)y * p->sw + x];
      if (*v == 0) continue;
      const int64_t level = floor_div((int64_t)*v * gain - bias, divisor);
      if (level > INT32_MAX || level < INT32_MIN) {
        r = 1;
        break;
      }
      *v = (int32_t)level;
    }
  }
  free(rise);
  free(offset);
  free(up);
  free(down);
  free(cols.pos);
  free(rows.pos);
  if (r) *stage = JXRHIP_STAGE_SAMPLE_SCALE;
  return r;
}

static int reconstruct_reduced(plane_t *p, uint32_t *stage) {
  if (reduced_picture(p, stage)) return JXRHIP_ERR_KERNEL_ARITHMETIC;
  if (scale_reduced(p, stage)) return JXRHIP_ERR_KERNEL_ARITHMETIC;
  return 0;
}

static int reconstruct(plane_t *p, uint32_t *stage) {
  if (p->resolution != JXRHIP_SCALE_FULL) return reconstruct_reduced(p, stage);
  if (first_level(p, 0)) {
    *stage = JXRHIP_STAGE_FIRST_TRANSFORM;
    return JXRHIP_ERR_KERNEL_ARITHMETIC;
  }
  if (p->arena->overlap == 2 && plane_overlap(p, 1)) {
    *stage = JXRHIP_STAGE_OVERLAP;
    return JXRHIP_ERR_KERNEL_ARITHMETIC;
  }
  if (second_level(p, stage)) return JXRHIP_ERR_KERNEL_ARITHMETIC;
  if (p->arena->overlap >= 1 && plane_overlap(p, 0)) {
    *stage = JXRHIP_STAGE_OVERLAP;
    return JXRHIP_ERR_KERNEL_ARITHMETIC;
  }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* image: planes, output formatting                                       */
/* ---------------------------------------------------------------------- */
/* colour planes of an N-component image (DecodeLimits::max_components,
 * crates/jxr-core/src/request.rs:85), plus alpha */
#define MAX_COLORS 64
#define MAXP (MAX_COLORS + 1)
typedef struct {
  plane_t p[MAXP];
  uint32_t count;      /* planes reconstructed                          */
  uint32_t colors;     /* colour planes (component_count)                */
  int alpha_index;     /* plane holding alpha, -1 for none               */
} image_t;

static void image_free(image_t *im) {
  for (uint32_t i = 0; i < im->count; i++) plane_free(&im->p[i]);
  im->count = 0;
}

static int image_setup(image_t *im, const jxrhip_image_desc *d, char *msg, size_t msglen) {
  memset(im, 0, sizeof *im);
  im->alpha_index = -1;
  if (d->abi_version != JXRHIP_ABI_VERSION) FAIL(JXRHIP_ERR_INVALID_ARGUMENT, "ABI version");
  if (!d->primary) FAIL(JXRHIP_ERR_INVALID_PLAN, "no primary arena");
  if (d->scale > JXRHIP_SCALE_SIXTEENTH) FAIL(JXRHIP_ERR_INVALID_PLAN, "decode scale");
  uint32_t colors = d->component_count;
  uint32_t cf = d->color_format;
  uint32_t expect = cf == JXRHIP_COLOR_LUMA ? 1
                    : (cf == JXRHIP_COLOR_YUV420 || cf == JXRHIP_COLOR_YUV422 || cf == JXRHIP_COLOR_YUV444 ||
                       cf == JXRHIP_COLOR_RGB || cf == JXRHIP_COLOR_RGBE)
                        ? 3
                    : (cf == JXRHIP_COLOR_CMYK || cf == JXRHIP_COLOR_CMYK_DIRECT || cf == JXRHIP_COLOR_YUVK) ? 4
                    : colors;
  if (colors != expect || colors == 0 || colors > MAX_COLORS)
    FAIL(JXRHIP_ERR_INVALID_PLAN, "primary component count does not match the colour format");
  im->colors = colors;
  uint32_t integrated = d->alpha_mode == JXRHIP_ALPHA_INTEGRATED;
  if (d->primary->plane_count != colors + integrated)
    FAIL(JXRHIP_ERR_INVALID_PLAN, "primary coefficient arena component count does not match the image");
  for (uint32_t i = 0; i < colors + integrated; i++) {
    int alpha = integrated && i == colors;
    int scale = !alpha && d->scaled && i != 0;
    int r = plane_setup(&im->p[i], d->primary, i, scale, alpha, d->scale, msg, msglen);
    if (r) {
      image_free(im);
      return r;
    }
    im->count++;
  }
  if (integrated) im->alpha_index = (int)colors;
  if (d->alpha_mode == JXRHIP_ALPHA_SEPARATE) {
    if (!d->separate_alpha || d->separate_alpha->plane_count != 1) {
      image_free(im);
      FAIL(JXRHIP_ERR_INVALID_PLAN, "separate alpha arena is not one independent luma plane");
    }
    int r = plane_setup(&im->p[im->count], d->separate_alpha, 0, 0, 1, d->scale, msg, msglen);
    if (r) {
      image_free(im);
      return r;
    }
    im->alpha_index = (int)im->count;
    im->count++;
  }
  /* chroma geometry must match the colour format */
  for (uint32_t i = 0; i < colors; i++) {
    uint32_t bc = im->p[i].bc, br = im->p[i].br;
    int chroma = i > 0 && i < 3;
    int ok = (cf == JXRHIP_COLOR_YUV420 && chroma) ? (bc == 2 && br == 2)
             : (cf == JXRHIP_COLOR_YUV422 && chroma) ? (bc == 2 && br == 4)
                                                       : (bc == 4 && br == 4);
    if (!ok) {
      image_free(im);
      FAIL(JXRHIP_ERR_INVALID_PLAN, "component macroblock geometry does not match the colour format");
    }
  }
  return 0;
}

typedef struct {
  const jxrhip_image_desc *d;
  image_t *im;
  uint32_t primary_count;
  int has_alpha;
  int reverse;
  uint32_t channels;
  uint32_t bytes;
  channel_scale color_scale[MAXP];
  channel_scale alpha_scale;
  int32_t u16_max;
  int upsample; /* 0 none, 1 = 4:2:0, 2 = 4:2:2 */
  uint32_t luma_mb; /* luma samples across a macroblock at the decode scale */
  int kind;     /* OUT_* */
  uint32_t depth;
} out_ctx;

/* Output kinds: typed samples per channel, one-bit luma, packed RGB words,
 * RGBE words and native-resolution planar YUV. */
enum { OUT_TYPED = 0, OUT_BITS, OUT_PACKED16, OUT_RGB101010, OUT_RGBE, OUT_YUV };

static int layout_has_alpha(uint32_t l) {
  return l == JXRHIP_LAYOUT_LUMA_ALPHA || l == JXRHIP_LAYOUT_YUVA || l == JXRHIP_LAYOUT_RGBA ||
         l == JXRHIP_LAYOUT_BGRA || l == JXRHIP_LAYOUT_CMYKA || l == JXRHIP_LAYOUT_NCOMPONENT_ALPHA;
}

/* sample of a full-resolution plane at coded position (x, y) */
static int32_t plane_at(const plane_t *p, uint32_t x, uint32_t y) {
  uint32_t lx = x - p->mb_ox * p->ux, ly = y - p->mb_oy * p->uy;
  return p->samples[(uint64_t)ly * p->sw + lx];
}

static int32_t chroma_clamped(const plane_t *p, int64_t cx, int64_t cy) {
  if (cx < 0) cx = 0;
  if (cx > (int64_t)p->sw - 1) cx = p->sw - 1;
  if (cy < 0) cy = 0;
  if (cy > (int64_t)p->sh - 1) cy = p->sh - 1;
  return p->samples[(uint64_t)cy * p->sw + cx];
}

/* Full-resolution chroma at coded position (x, y): vertical interpolation
 * of each needed chroma column, then horizontal (reconstruct/chroma.rs:
 * 174-325). The plane's end samples stand in beyond its edges. */
static int32_t chroma_at(const out_ctx *o, const plane_t *p, uint32_t x, uint32_t y) {
  const jxrhip_output_format *f = &o->d->output;
  int64_t lx = (int64_t)x - (int64_t)p->mb_ox * o->luma_mb;
  int64_t ly = (int64_t)y - (int64_t)p->mb_oy * o->luma_mb;
  int64_t cx = lx >> 1;
  int32_t pair[2];
  if (o->upsample == 2) {
    jxro_upsample_pair(chroma_clamped(p, cx - 1, ly), chroma_clamped(p, cx, ly),
                       chroma_clamped(p, cx + 1, ly), f->chroma_centering_x, pair);
    return pair[lx & 1];
  }
  int64_t cy = ly >> 1;
  int32_t vert[3];
  for (int k = -1; k <= 1; k++) {
    jxro_upsample_pair(chroma_clamped(p, cx + k, cy - 1), chroma_clamped(p, cx + k, cy),
                       chroma_clamped(p, cx + k, cy + 1), f->chroma_centering_y, pair);
    vert[k + 1] = pair[ly & 1];
  }
  jxro_upsample_pair(vert[0], vert[1], vert[2], f->chroma_centering_x, pair);
  return pair[lx & 1];
}

/* output_format/color.rs:100-123 */
static int inverse_yuv(int32_t y, int32_t u, int32_t v, int32_t rgb[3]) {
  int err = 0, *e = &err;
  int32_t t = SUB(0, u);
  int32_t g = SUB(y, t >> 1);
  int32_t half = (v >> 1) + (v & 1);
  int32_t r = SUB(ADD(t, g), half);
  int32_t b = ADD(v, r);
  rgb[0] = r;
  rgb[1] = g;
  rgb[2] = b;
  return err;
}
static int inverse_yuvk(const int32_t in[4], int32_t out[4]) {
  int err = 0, *e = &err;
  int32_t black = ADD(in[3], in[0] >> 1);
  int32_t magenta = SUB(SUB(black, in[0]), in[1] >> 1);
  int32_t cyan = ADD(ADD(in[1], magenta), in[2] >> 1);
  int32_t yellow = SUB(cyan, in[2]);
  out[0] = cyan;
  out[1] = magenta;
  out[2] = yellow;
  out[3] = black;
  return err;
}

/* Internal-to-output colour of one pixel (output_format/color.rs:9-47):
 * internal is the reconstructed colour at full resolution (a 4:2:0 or 4:2:2
 * image's upsampled chroma counts as 4:4:4). Packed RGB words swap red and
 * blue when the image does not say they are not swapped. N-component
 * images keep their components as they are. 0 ok, 1 overflow,
 * 2 unsupported combination. */
int jxro_convert(uint32_t internal, uint32_t output, const int32_t in[4], int32_t packed_rgb,
                 int32_t red_blue_not_swapped, int32_t out[4]) {
  uint32_t ic = internal, oc = output;
  int yuv = ic == JXRHIP_COLOR_YUV420 || ic == JXRHIP_COLOR_YUV422 || ic == JXRHIP_COLOR_YUV444;
  out[0] = out[1] = out[2] = out[3] = 0;
  if (ic == JXRHIP_COLOR_LUMA && oc == JXRHIP_COLOR_LUMA) {
    out[0] = in[0];
  } else if (ic == JXRHIP_COLOR_LUMA && oc == JXRHIP_COLOR_RGB) {
    out[0] = out[1] = out[2] = in[0];
  } else if (yuv && (oc == JXRHIP_COLOR_RGB || oc == JXRHIP_COLOR_RGBE)) {
    if (inverse_yuv(in[0], in[1], in[2], out)) return 1;
    if (packed_rgb && !red_blue_not_swapped) {
      int32_t t = out[0];
      out[0] = out[2];
      out[2] = t;
    }
  } else if (ic == JXRHIP_COLOR_YUVK && oc == JXRHIP_COLOR_CMYK) {
    if (inverse_yuvk(in, out)) return 1;
  } else if (ic == JXRHIP_COLOR_YUVK && oc == JXRHIP_COLOR_CMYK_DIRECT) {
    out[0] = in[1];
    out[1] = in[2];
    out[2] = in[3];
    out[3] = in[0];
  } else if (ic == JXRHIP_COLOR_NCOMPONENT && oc == JXRHIP_COLOR_NCOMPONENT) {
    for (int c = 0; c < 4; c++) out[c] = in[c];
  } else {
    return 2;
  }
  return 0;
}

/* Converted colour components of one pixel (packing.rs:465-505). */
static int color_at(const out_ctx *o, uint32_t x, uint32_t y, int32_t out[4]) {
  const jxrhip_image_desc *d = o->d;
  const image_t *im = o->im;
  int32_t in[4] = {0, 0, 0, 0};
  for (uint32_t c = 0; c < im->colors && c < 4; c++) {
    if (o->upsample && (c == 1 || c == 2))
      in[c] = chroma_at(o, &im->p[c], x, y);
    else
      in[c] = plane_at(&im->p[c], x, y);
  }
  const int packed = o->kind == OUT_PACKED16 || o->kind == OUT_RGB101010;
  return jxro_convert(d->color_format, d->output.output_color, in, packed, (int32_t)d->output.red_blue_not_swapped,
                      out);
}

static void store_le(uint8_t *dst, uint64_t v, uint32_t bytes) {
  for (uint32_t b = 0; b < bytes; b++) dst[b] = (uint8_t)(v >> (8 * b));
}

static int output_setup(out_ctx *o, const jxrhip_image_desc *d, image_t *im, char *msg,
                        size_t msglen) {
  memset(o, 0, sizeof *o);
  o->d = d;
  o->im = im;
  const jxrhip_output_format *f = &d->output;
  uint32_t l = f->channel_layout;
  switch (l) {
    case JXRHIP_LAYOUT_LUMA:
    case JXRHIP_LAYOUT_LUMA_ALPHA: o->primary_count = 1; break;
    case JXRHIP_LAYOUT_RGB:
    case JXRHIP_LAYOUT_BGR:
    case JXRHIP_LAYOUT_RGBA:
    case JXRHIP_LAYOUT_BGRA: o->primary_count = 3; break;
    case JXRHIP_LAYOUT_RGBX:
    case JXRHIP_LAYOUT_BGRX:
    case JXRHIP_LAYOUT_CMYK:
    case JXRHIP_LAYOUT_CMYKA: o->primary_count = 4; break;
    case JXRHIP_LAYOUT_NCOMPONENT:
    case JXRHIP_LAYOUT_NCOMPONENT_ALPHA: o->primary_count = im->colors; break;
    case JXRHIP_LAYOUT_YUV:
    case JXRHIP_LAYOUT_YUVA: o->primary_count = 3; break;
    default: FAIL(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "channel layout not supported");
  }
  o->has_alpha = layout_has_alpha(l);
  o->reverse = l == JXRHIP_LAYOUT_BGR || l == JXRHIP_LAYOUT_BGRX || l == JXRHIP_LAYOUT_BGRA;
  o->channels = o->primary_count + (uint32_t)o->has_alpha;
  if (o->channels != f->channel_count) FAIL(JXRHIP_ERR_INVALID_PLAN, "channel count differs from the layout");
  if (o->has_alpha && im->alpha_index < 0) FAIL(JXRHIP_ERR_INVALID_PLAN, "alpha output layout has no prepared alpha plane");
  uint32_t st = f->storage;
  uint32_t dep = f->bit_depth;
  uint32_t adep = f->has_alpha_format ? f->alpha_bit_depth : dep;
  uint32_t oc = f->output_color;
  o->depth = dep;
  o->kind = OUT_TYPED;
  /* validate.rs:137-255: storage, colour, layout and depth combinations of
   * the packed, one-bit and planar YUV outputs */
  if (st == JXRHIP_STORE_BITS) {
    if (dep != JXRHIP_DEPTH_BIT1_WHITE && dep != JXRHIP_DEPTH_BIT1_BLACK) goto bad_depth;
    if (l != JXRHIP_LAYOUT_LUMA || oc != JXRHIP_COLOR_LUMA)
      FAIL(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "one-bit output stores luma only");
    o->kind = OUT_BITS;
    o->bytes = 1;
  } else if (st == JXRHIP_STORE_PACKED16 || st == JXRHIP_STORE_PACKED32) {
    if (l != JXRHIP_LAYOUT_RGB) FAIL(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "packed output takes the RGB layout");
    if (st == JXRHIP_STORE_PACKED16 && (dep == JXRHIP_DEPTH_RGB555 || dep == JXRHIP_DEPTH_RGB565) &&
        oc == JXRHIP_COLOR_RGB)
      o->kind = OUT_PACKED16;
    else if (st == JXRHIP_STORE_PACKED32 && dep == JXRHIP_DEPTH_RGB101010 && oc == JXRHIP_COLOR_RGB)
      o->kind = OUT_RGB101010;
    else if (st == JXRHIP_STORE_PACKED32 && dep == JXRHIP_DEPTH_U8 && oc == JXRHIP_COLOR_RGBE)
      o->kind = OUT_RGBE;
    else
      goto bad_depth;
    o->bytes = st == JXRHIP_STORE_PACKED16 ? 2 : 4;
  } else if (l == JXRHIP_LAYOUT_YUV || l == JXRHIP_LAYOUT_YUVA) {
    if (oc != d->color_format || (oc != JXRHIP_COLOR_YUV420 && oc != JXRHIP_COLOR_YUV422) ||
        f->internal_color != oc)
      FAIL(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "planar YUV output keeps the coded 4:2:0 or 4:2:2 sampling");
    if (f->premultiply_alpha) FAIL(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "premultiplied YUV output");
    if (!d->surface.planar) FAIL(JXRHIP_ERR_INVALID_PLAN, "YUV output is planar");
    /* prepare.rs:185-192: a sixteenth has one chroma sample per macroblock */
    if (d->scale == JXRHIP_SCALE_SIXTEENTH)
      FAIL(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "sixteenth-resolution subsampled planar YUV output");
    o->kind = OUT_YUV;
  } else if (dep == JXRHIP_DEPTH_BIT1_WHITE || dep == JXRHIP_DEPTH_BIT1_BLACK || dep == JXRHIP_DEPTH_RGB555 ||
             dep == JXRHIP_DEPTH_RGB565 || dep == JXRHIP_DEPTH_RGB101010) {
    goto bad_depth;
  }
  if (o->kind == OUT_TYPED || o->kind == OUT_YUV) switch (st) {
    case JXRHIP_STORE_U8: o->bytes = 1; if (dep != JXRHIP_DEPTH_U8) goto bad_depth; break;
    case JXRHIP_STORE_U16: o->bytes = 2; if (dep != JXRHIP_DEPTH_U16 && dep != JXRHIP_DEPTH_U10) goto bad_depth; break;
    case JXRHIP_STORE_I16: o->bytes = 2; if (dep != JXRHIP_DEPTH_I16) goto bad_depth; break;
    case JXRHIP_STORE_I32: o->bytes = 4; if (dep != JXRHIP_DEPTH_I32) goto bad_depth; break;
    case JXRHIP_STORE_F16: o->bytes = 2; if (dep != JXRHIP_DEPTH_F16) goto bad_depth; break;
    case JXRHIP_STORE_F32:
      o->bytes = 4;
      if (dep != JXRHIP_DEPTH_F32 || (o->has_alpha && adep != JXRHIP_DEPTH_F32)) goto bad_depth;
      if (f->mantissa_length > 23 || (o->has_alpha && f->has_alpha_format && f->alpha_mantissa_length > 23))
        FAIL(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "BD32F mantissa length exceeds 23 bits");
      break;
    default: FAIL(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "storage kind not supported");
  }
  if (o->kind == OUT_YUV && st != JXRHIP_STORE_U8 && st != JXRHIP_STORE_U16 && st != JXRHIP_STORE_I16)
    FAIL(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "T.832 output colour and bit depth matrix");
  o->u16_max = dep == JXRHIP_DEPTH_U10 ? 1023 : 65535;
  uint32_t cf = d->color_format;
  /* interpolates_chroma (metadata.rs:253-263): not for native planar
   * output, nor at a sixteenth, where chroma has a sample per macroblock as
   * luma does (reconstructed_color, decode/output.rs:464-482) */
  o->upsample = o->kind == OUT_YUV || d->scale == JXRHIP_SCALE_SIXTEENTH ? 0
                : cf == JXRHIP_COLOR_YUV420                            ? 1
                : cf == JXRHIP_COLOR_YUV422                            ? 2
                                                                       : 0;
  o->luma_mb = d->scale == JXRHIP_SCALE_QUARTER ? 4 : d->scale == JXRHIP_SCALE_SIXTEENTH ? 1 : 16;
  if (o->upsample && (f->chroma_centering_x > 4 || (o->upsample == 1 && f->chroma_centering_y > 4)))
    FAIL(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "unknown chroma centering has no exact reconstruction filter");
  for (uint32_t c = 0; c < MAXP; c++) {
    o->color_scale[c] = resolve_scale(c, f->output_color, dep, f->shift_bits, f->scaled);
    if ((l == JXRHIP_LAYOUT_RGBX || l == JXRHIP_LAYOUT_BGRX) && c == 3) o->color_scale[c].padding = 1;
  }
  o->alpha_scale = resolve_scale(0, JXRHIP_COLOR_LUMA, adep,
                                 f->has_alpha_format ? f->alpha_shift_bits : f->shift_bits,
                                 f->has_alpha_format ? f->alpha_scaled : f->scaled);
  /* surface */
  const jxrhip_surface_layout *s = &d->surface;
  if (s->width != f->crop_width || s->height != f->crop_height || f->crop_width == 0 || f->crop_height == 0)
    FAIL(JXRHIP_ERR_INVALID_PLAN, "output layout and prepared request differ");
  uint32_t need = s->planar ? o->channels : 1;
  if ((o->kind == OUT_BITS || o->kind == OUT_PACKED16 || o->kind == OUT_RGB101010 || o->kind == OUT_RGBE) &&
      s->planar)
    FAIL(JXRHIP_ERR_INVALID_PLAN, "packed and one-bit outputs are interleaved");
  if (s->plane_count != need || need > JXRHIP_MAX_SURFACE_PLANES)
    FAIL(JXRHIP_ERR_INVALID_PLAN, "surface plane count");
  for (uint32_t i = 0; i < need; i++) {
    const jxrhip_surface_plane *sp = &s->planes[i];
    uint32_t ch = s->planar ? 1 : o->channels;
    uint32_t pw = s->width, ph = s->height;
    if (o->kind == OUT_YUV && (i == 1 || i == 2)) {
      /* planar.rs:77-106: chroma planes at the coded sampling */
      pw = s->width / 2;
      ph = cf == JXRHIP_COLOR_YUV420 ? s->height / 2 : s->height;
      if (pw == 0 || ph == 0) FAIL(JXRHIP_ERR_INVALID_PLAN, "planar chroma plane is empty");
    }
    if (sp->width != pw || sp->height != ph || sp->channels != ch)
      FAIL(JXRHIP_ERR_INVALID_PLAN, "surface plane geometry");
    uint64_t row = (uint64_t)sp->width * ch * o->bytes;
    if (o->kind == OUT_BITS) row = ((uint64_t)sp->width + 7) / 8;              /* format.rs:216 */
    if (o->kind == OUT_PACKED16 || o->kind == OUT_RGB101010 || o->kind == OUT_RGBE)
      row = (uint64_t)sp->width * o->bytes;                                      /* one word a pixel */
    if (sp->row_stride_bytes < row) FAIL(JXRHIP_ERR_INVALID_PLAN, "surface row stride");
    if (sp->byte_offset + (uint64_t)(sp->height - 1) * sp->row_stride_bytes + row > s->byte_len)
      FAIL(JXRHIP_ERR_INVALID_PLAN, "surface plane exceeds its byte length");
  }
  /* crop inside every plane's window (plan.rs:282-354) */
  for (uint32_t i = 0; i < im->count; i++) {
    const plane_t *p = &im->p[i];
    /* a plane's samples per luma sample: its macroblock over luma's */
    uint32_t dx = o->luma_mb / p->ux, dy = o->luma_mb / p->uy;
    uint64_t left = f->crop_x / dx, top = f->crop_y / dy;
    uint64_t right = ((uint64_t)f->crop_x + f->crop_width - 1) / dx + 1;
    uint64_t bottom = ((uint64_t)f->crop_y + f->crop_height - 1) / dy + 1;
    uint64_t px = (uint64_t)p->mb_ox * p->ux, py = (uint64_t)p->mb_oy * p->uy;
    if (left < px || top < py || right > px + p->sw || bottom > py + p->sh)
      FAIL(JXRHIP_ERR_INVALID_PLAN, "output crop is outside a sample plane");
  }
  return 0;
bad_depth:
  FAIL(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "bit depth does not match the storage kind");
}

/* ---- packed words (packed_color.rs:9-38, 100-135) ---------------------- */
static uint32_t clip_unsigned(int32_t v, uint32_t maximum) {
  return v < 0 ? 0u : ((uint32_t)v > maximum ? maximum : (uint32_t)v);
}

/* RGBE (rgbe.rs:6-100): each channel as a mantissa and exponent, aligned to
 * the largest exponent with a rounded shift, bytes [R, G, B, E]. */
static void rgbe_component(int32_t v, int32_t *mantissa, int32_t *exponent) {
  if (v <= 0) {
    *mantissa = 0;
    *exponent = 0;
  } else if ((v >> 7) > 1) {
    *mantissa = (v & 0x7f) + 128;
    *exponent = v >> 7;
  } else {
    *mantissa = v;
    *exponent = 1;
  }
}
static int32_t rgbe_rounded_shift(int32_t value, int32_t difference) {
  if (difference >= 31) return 0;
  return (int32_t)((2 * (int64_t)value + 1) >> (difference + 1));
}
uint32_t jxro_rgbe_word(const int32_t rgb[3]) {
  int32_t m[3], e[3], top = 0;
  for (int c = 0; c < 3; c++) {
    rgbe_component(rgb[c], &m[c], &e[c]);
    if (e[c] > top) top = e[c];
  }
  uint32_t word = 0;
  for (int c = 0; c < 3; c++) {
    int32_t v = top > e[c] ? rgbe_rounded_shift(m[c], top - e[c]) : m[c];
    word |= (uint32_t)(v < 0 ? 0 : v > 255 ? 255 : v) << (8 * c);
  }
  return word | (uint32_t)(top > 255 ? 255 : top) << 24;
}

/* One packed word of converted RGB (packed_color.rs:48-98). */
static int packed_word(const out_ctx *o, const int32_t rgb[3], uint32_t *word) {
  int32_t v[3];
  for (int c = 0; c < 3; c++)
    if (apply_scale(&o->color_scale[c], rgb[c], &v[c])) return 1;
  switch (o->kind) {
    case OUT_PACKED16:
      if (o->depth == JXRHIP_DEPTH_RGB565)
        *word = clip_unsigned(v[2], 31) | clip_unsigned(v[1], 63) << 5 | clip_unsigned(v[0], 31) << 11;
      else
        *word = clip_unsigned(v[2], 31) | clip_unsigned(v[1], 31) << 5 | clip_unsigned(v[0], 31) << 10;
      return 0;
    case OUT_RGB101010:
      *word = clip_unsigned(v[2], 1023) | clip_unsigned(v[1], 1023) << 10 | clip_unsigned(v[0], 1023) << 20;
      return 0;
    default:
      *word = jxro_rgbe_word(v);
      return 0;
  }
}

/* One-bit luma rows (packing.rs:364-386): the first pixel of a row is the
 * highest bit of its first byte; padding bits stay zero. */
static int format_bits(const out_ctx *o, uint8_t *out) {
  const jxrhip_output_format *f = &o->d->output;
  const jxrhip_surface_plane *sp = &o->d->surface.planes[0];
  uint32_t row_bytes = (f->crop_width + 7) / 8;
  for (uint32_t y = 0; y < f->crop_height; y++) {
    uint8_t *row = out + sp->byte_offset + (uint64_t)y * sp->row_stride_bytes;
    memset(row, 0, row_bytes);
    for (uint32_t x = 0; x < f->crop_width; x++) {
      int32_t conv[4], v;
      int r = color_at(o, f->crop_x + x, f->crop_y + y, conv);
      if (r) return r;
      if (apply_scale(&o->color_scale[0], conv[0], &v)) return 1;
      uint32_t bit = v < 0 ? 0u : v > 1 ? 1u : (uint32_t)v;
      if (o->depth == JXRHIP_DEPTH_BIT1_BLACK) bit = 1 - bit;
      row[x / 8] |= (uint8_t)(bit << (7 - x % 8));
    }
  }
  return 0;
}

/* Native-resolution planar YUV (planar.rs:77-230): luma and alpha over the
 * crop, each chroma plane over the crop at its coded sampling. */
static int format_yuv(const out_ctx *o, uint8_t *out) {
  const jxrhip_output_format *f = &o->d->output;
  const jxrhip_surface_layout *s = &o->d->surface;
  const image_t *im = o->im;
  int cs420 = o->d->color_format == JXRHIP_COLOR_YUV420;
  for (uint32_t pl = 0; pl < s->plane_count; pl++) {
    const jxrhip_surface_plane *sp = &s->planes[pl];
    int chroma = pl == 1 || pl == 2, alpha = pl == 3;
    const plane_t *src = alpha ? &im->p[im->alpha_index] : &im->p[pl];
    const channel_scale *cs = alpha ? &o->alpha_scale : &o->color_scale[pl];
    uint32_t x0 = chroma ? f->crop_x / 2 : f->crop_x;
    uint32_t y0 = chroma && cs420 ? f->crop_y / 2 : f->crop_y;
    for (uint32_t y = 0; y < sp->height; y++)
      for (uint32_t x = 0; x < sp->width; x++) {
        int32_t v;
        if (apply_scale(cs, plane_at(src, x0 + x, y0 + y), &v)) return 1;
        switch (f->storage) {
          case JXRHIP_STORE_U8: v = v < 0 ? 0 : v > 255 ? 255 : v; break;
          case JXRHIP_STORE_U16: v = v < 0 ? 0 : v > o->u16_max ? o->u16_max : v; break;
          default: v = v < -32768 ? -32768 : v > 32767 ? 32767 : v; break;
        }
        store_le(out + sp->byte_offset + (uint64_t)y * sp->row_stride_bytes + (uint64_t)x * o->bytes,
                 (uint32_t)v, o->bytes);
      }
  }
  return 0;
}

/* Stores every cropped pixel (packing.rs:333-357 ordered stores, then
 * premultiply.rs on the stored values). */
static int format_output(const out_ctx *o, uint8_t *out) {
  const jxrhip_image_desc *d = o->d;
  const jxrhip_output_format *f = &d->output;
  const jxrhip_surface_layout *s = &d->surface;
  const image_t *im = o->im;
  uint32_t st = f->storage;
  if (o->kind == OUT_BITS) return format_bits(o, out);
  if (o->kind == OUT_YUV) return format_yuv(o, out);
  if (o->kind != OUT_TYPED) {
    const jxrhip_surface_plane *sp = &s->planes[0];
    for (uint32_t y = 0; y < f->crop_height; y++)
      for (uint32_t x = 0; x < f->crop_width; x++) {
        int32_t conv[4];
        uint32_t word;
        int r = color_at(o, f->crop_x + x, f->crop_y + y, conv);
        if (r) return r;
        if (packed_word(o, conv, &word)) return 1;
        store_le(out + sp->byte_offset + (uint64_t)y * sp->row_stride_bytes + (uint64_t)x * o->bytes, word,
                 o->bytes);
      }
    return 0;
  }
  int ncomp = d->output.output_color == JXRHIP_COLOR_NCOMPONENT;
  for (uint32_t y = 0; y < f->crop_height; y++)
    for (uint32_t x = 0; x < f->crop_width; x++) {
      uint32_t sx = f->crop_x + x, sy = f->crop_y + y;
      int32_t ordered[MAXP + 1] = {0};
      if (ncomp) {
        for (uint32_t c = 0; c < o->primary_count; c++) ordered[c] = plane_at(&im->p[c], sx, sy);
      } else {
        int32_t conv[4];
        int r = color_at(o, sx, sy, conv);
        if (r) return r;
        if (o->reverse) {
          ordered[0] = conv[2];
          ordered[1] = conv[1];
          ordered[2] = conv[0];
          ordered[3] = 0;
        } else {
          for (int c = 0; c < 4; c++) ordered[c] = conv[c];
        }
      }
      if (o->has_alpha) ordered[o->primary_count] = plane_at(&im->p[im->alpha_index], sx, sy);
      uint64_t stored[MAXP + 1];
      for (uint32_t c = 0; c < o->channels; c++) {
        int alpha = o->has_alpha && c == o->primary_count;
        const channel_scale *cs = alpha ? &o->alpha_scale : &o->color_scale[c];
        int32_t v;
        if (st == JXRHIP_STORE_F16) {
          if (cs->padding) {
            stored[c] = 0;
            continue;
          }
          uint16_t b;
          if (jxro_float16_bits(ordered[c], alpha ? (f->has_alpha_format ? f->alpha_scaled : f->scaled) : f->scaled, &b))
            return 1;
          stored[c] = b;
          continue;
        }
        if (st == JXRHIP_STORE_F32) {
          if (cs->padding) {
            stored[c] = 0;
            continue;
          }
          uint32_t b;
          int hf = alpha && f->has_alpha_format;
          if (jxro_float32_bits(ordered[c], hf ? f->alpha_scaled : f->scaled,
                                hf ? f->alpha_mantissa_length : f->mantissa_length,
                                hf ? f->alpha_exponent_bias : f->exponent_bias, &b))
            return 1;
          stored[c] = b;
          continue;
        }
        if (apply_scale(cs, ordered[c], &v)) return 1;
        switch (st) {
          case JXRHIP_STORE_U8: v = v < 0 ? 0 : v > 255 ? 255 : v; break;
          case JXRHIP_STORE_U16: v = v < 0 ? 0 : v > o->u16_max ? o->u16_max : v; break;
          case JXRHIP_STORE_I16: v = v < -32768 ? -32768 : v > 32767 ? 32767 : v; break;
          default: break;
        }
        stored[c] = (uint32_t)v;
      }
      if (f->premultiply_alpha && o->has_alpha) {
        uint32_t ai = o->channels - 1;
        for (uint32_t c = 0; c < ai; c++) {
          if (st == JXRHIP_STORE_U8 || st == JXRHIP_STORE_U16) {
            uint32_t r;
            if (jxro_premultiply_unsigned((uint32_t)stored[c], (uint32_t)stored[ai],
                                          st == JXRHIP_STORE_U8 ? 255 : 65535, &r))
              return 1;
            stored[c] = r;
          } else if (st == JXRHIP_STORE_I16) {
            int32_t r;
            if (jxro_premultiply_signed((int16_t)stored[c], (int16_t)stored[ai], 32767, &r)) return 1;
            stored[c] = (uint32_t)r;
          } else if (st == JXRHIP_STORE_I32) {
            int32_t r;
            if (jxro_premultiply_signed((int32_t)stored[c], (int32_t)stored[ai], INT32_MAX, &r)) return 1;
            stored[c] = (uint32_t)r;
          } else if (st == JXRHIP_STORE_F16) {
            uint16_t r;
            if (premultiply_sm15((uint16_t)stored[c], (uint16_t)stored[ai], &r)) return 1;
            stored[c] = r;
          } else if (st == JXRHIP_STORE_F32) {
            float v, a;
            uint32_t vb = (uint32_t)stored[c], ab = (uint32_t)stored[ai];
            memcpy(&v, &vb, 4);
            memcpy(&a, &ab, 4);
            /* f32::clamp: NaN stays NaN */
            float ac = a < 0.0f ? 0.0f : (a > 1.0f ? 1.0f : a);
            float r = v * ac;
            memcpy(&vb, &r, 4);
            stored[c] = vb;
          }
        }
      }
      for (uint32_t c = 0; c < o->channels; c++) {
        uint64_t at;
        if (s->planar)
          at = s->planes[c].byte_offset + (uint64_t)y * s->planes[c].row_stride_bytes + (uint64_t)x * o->bytes;
        else
          at = s->planes[0].byte_offset + (uint64_t)y * s->planes[0].row_stride_bytes +
               ((uint64_t)x * o->channels + c) * o->bytes;
        store_le(out + at, stored[c], o->bytes);
      }
    }
  return 0;
}

int jxro_decode(const jxrhip_image_desc *d, uint8_t *out, uint64_t out_len, uint32_t *stage,
                char *msg, size_t msglen) {
  uint32_t st = 0;
  if (stage) *stage = 0;
  image_t im;
  int r = image_setup(&im, d, msg, msglen);
  if (r) return r;
  out_ctx o;
  r = output_setup(&o, d, &im, msg, msglen);
  if (r) {
    image_free(&im);
    return r;
  }
  if (out_len < d->surface.byte_len) {
    image_free(&im);
    FAIL(JXRHIP_ERR_BUFFER_TOO_SMALL, "output buffer smaller than the surface");
  }
  for (uint32_t i = 0; i < im.count; i++) {
    r = reconstruct(&im.p[i], &st);
    if (r) {
      image_free(&im);
      if (stage) *stage = st;
      FAIL(r, "arithmetic overflow in plane %u stage %u", i, st);
    }
  }
  r = format_output(&o, out);
  image_free(&im);
  if (r == 2) FAIL(JXRHIP_ERR_UNSUPPORTED_OUTPUT, "internal-to-output colour conversion");
  if (r) {
    if (stage) *stage = JXRHIP_STAGE_OUTPUT;
    FAIL(JXRHIP_ERR_KERNEL_ARITHMETIC, "arithmetic overflow in output formatting");
  }
  return 0;
}

int jxro_reconstruct_plane(const jxrhip_image_desc *d, uint32_t index, int32_t *samples,
                           uint64_t capacity, uint32_t *width, uint32_t *height,
                           uint32_t *origin_x, uint32_t *origin_y, uint32_t *stage) {
  image_t im;
  char *msg = NULL;
  size_t msglen = 0;
  int r = image_setup(&im, d, msg, msglen);
  if (r) return r;
  if (index >= im.count) {
    image_free(&im);
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  plane_t *p = &im.p[index];
  uint32_t st = 0;
  r = reconstruct(p, &st);
  if (stage) *stage = st;
  if (!r) {
    *width = p->sw;
    *height = p->sh;
    *origin_x = p->mb_ox * p->ux;
    *origin_y = p->mb_oy * p->uy;
    if ((uint64_t)p->sw * p->sh > capacity)
      r = JXRHIP_ERR_BUFFER_TOO_SMALL;
    else
      memcpy(samples, p->samples, (size_t)p->sw * p->sh * sizeof(int32_t));
  }
  image_free(&im);
  return r;
}

/* ---------------------------------------------------------------------- */
/* threaded batch decode (per-image split)                                 */
/* ---------------------------------------------------------------------- */
typedef struct {
  const jxrhip_image_desc *const *descs;
  uint8_t *const *outs;
  const uint64_t *lens;
  uint32_t count;
  uint32_t next;
  int first_error;
  pthread_mutex_t lock;
} batch_t;

static void *batch_worker(void *arg) {
  batch_t *b = (batch_t *)arg;
  for (;;) {
    pthread_mutex_lock(&b->lock);
    uint32_t i = b->next++;
    pthread_mutex_unlock(&b->lock);
    if (i >= b->count) break;
    uint32_t st;
    int r = jxro_decode(b->descs[i], b->outs[i], b->lens[i], &st, NULL, 0);
    if (r) {
      pthread_mutex_lock(&b->lock);
      if (!b->first_error) b->first_error = r;
      pthread_mutex_unlock(&b->lock);
    }
  }
  return NULL;
}

int jxro_decode_many(const jxrhip_image_desc *const *descs, uint8_t *const *outs,
                     const uint64_t *out_lens, uint32_t count, uint32_t threads) {
  batch_t b;
  b.descs = descs;
  b.outs = outs;
  b.lens = out_lens;
  b.count = count;
  b.next = 0;
  b.first_error = 0;
  pthread_mutex_init(&b.lock, NULL);
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  pthread_t tid[256];
  for (uint32_t t = 0; t < threads; t++) pthread_create(&tid[t], NULL, batch_worker, &b);
  for (uint32_t t = 0; t < threads; t++) pthread_join(tid[t], NULL);
  pthread_mutex_destroy(&b.lock);
  return b.first_error;
}
