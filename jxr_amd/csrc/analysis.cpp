// analysis.cpp — forward analysis of seeded images into coefficient arenas.
//
// The encoder side of the transforms the reconstruction inverts: colour
// transform and bias removal (crates/jxr-native/src/output_format/color.rs:
// 74-84, scaling.rs:129-148), chroma decimation (crates/jxr-math/src/
// sampling.rs:72-97), sample- and low-pass-level overlap prefilters
// (crates/jxr-math/src/overlap.rs:59-263, reconstruct/overlap.rs:124-161,
// reconstruct/subsampled_overlap.rs:110-152), forward core transforms
// (crates/jxr-math/src/transform.rs:29-131), quantization
// (crates/jxr-math/src/quantization.rs:31-50) and high-pass prediction
// residuals (the forward of crates/jxr-math/src/prediction.rs:53-97). Entropy
// coding is not modelled: the arena is what the reference's entropy decoder
// would hand the reconstruction.

#include <algorithm>
#include <atomic>
#include <memory>
#include <utility>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/jxr_hip_encode.h"
#include "encoded.h"

extern "C" void jxrhip_set_error(const char *message);

namespace {

using i32 = int32_t;
using i64 = int64_t;

// ---- forward operators --------------------------------------------------
inline void t2x2h(i32 *v, i32 r) {
  v[0] += v[3];
  v[1] -= v[2];
  i32 first = (v[0] - v[1] + r) >> 1;
  i32 second = v[2];
  v[2] = first - v[3];
  v[3] = first - second;
  v[0] -= v[3];
  v[1] += v[2];
}

inline void forward_todd(i32 *v) {
  v[0] += v[3];
  v[1] -= v[2];
  v[3] = ((v[0] + 1) >> 1) - v[3];
  v[2] += (v[1] + 1) >> 1;
  v[3] -= (v[2] * 3 + 4) >> 3;
  v[2] += (v[3] * 3 + 4) >> 3;
  v[1] -= (v[0] * 3 + 4) >> 3;
  v[0] += (v[1] * 3 + 4) >> 3;
  v[2] -= (v[0] + 1) >> 1;
  v[3] += v[1] >> 1;
  v[0] += v[2];
  v[1] -= v[3];
}

inline void forward_todd_odd(i32 *v) {
  v[1] = -v[1];
  v[2] = -v[2];
  v[3] += v[0];
  v[2] -= v[1];
  i32 first = v[3] >> 1, second = v[2] >> 1;
  v[0] -= first;
  v[1] += second;
  v[0] += (v[1] * 3 + 4) >> 3;
  v[1] -= (v[0] * 3 + 3) >> 2;
  v[0] += (v[1] * 3 + 3) >> 3;
  v[1] -= second;
  v[0] += first;
  v[2] += v[1];
  v[3] -= v[0];
}

template <class F>
inline void grp(i32 *c, int a, int b, int d, int e, F f) {
  i32 v[4] = {c[a], c[b], c[d], c[e]};
  f(v);
  c[a] = v[0];
  c[b] = v[1];
  c[d] = v[2];
  c[e] = v[3];
}

const int PERM[16] = {0, 8, 4, 13, 2, 15, 3, 14, 1, 12, 5, 9, 7, 11, 6, 10};

void forward_core(i32 *c) {
  auto z = [](i32 *v) { t2x2h(v, 0); };
  grp(c, 4, 7, 8, 11, z);
  grp(c, 1, 2, 13, 14, z);
  grp(c, 5, 6, 9, 10, z);
  grp(c, 0, 3, 12, 15, z);
  grp(c, 10, 11, 14, 15, forward_todd_odd);
  grp(c, 8, 12, 9, 13, forward_todd);
  grp(c, 2, 3, 6, 7, forward_todd);
  grp(c, 0, 1, 4, 5, [](i32 *v) { t2x2h(v, 1); });
  i32 in[16];
  std::memcpy(in, c, sizeof in);
  for (int i = 0; i < 16; i++) c[i] = in[PERM[i]];
}

void forward_chroma420(i32 *v) {
  std::swap(v[1], v[2]);
  t2x2h(v, 0);
}

void forward_chroma422(i32 *v) {
  std::swap(v[5], v[6]);
  i32 s[4] = {v[4], v[6], v[5], v[7]};
  t2x2h(s, 0);
  v[4] = s[0];
  v[6] = s[1];
  v[5] = s[2];
  v[7] = s[3];
  std::swap(v[1], v[2]);
  t2x2h(v, 0);
  v[4] -= v[0];
  v[0] += (v[4] + 1) >> 1;
}

inline void fwd_rotate(i32 &a, i32 &b) {
  b -= (a + 1) >> 1;
  a += (b + 1) >> 1;
}
inline void fwd_scale(i32 &a, i32 &b) {
  b += a >> 10;
  b -= a >> 7;
  b -= (a * 3) >> 4;
  a -= (b * 3) >> 3;
  b = (a >> 1) - b;
  a -= b;
}

void forward_overlap4(i32 *v) {
  v[1] += v[2];
  v[0] += v[3];
  v[2] -= (v[1] + 1) >> 1;
  v[3] -= (v[0] + 1) >> 1;
  fwd_rotate(v[2], v[3]);
  v[2] = -v[2];
  v[3] = -v[3];
  v[1] -= v[2];
  v[0] -= v[3];
  v[2] += v[1] >> 1;
  v[3] += v[0] >> 1;
  v[1] -= (v[2] * 3 + 4) >> 3;
  v[0] -= (v[3] * 3 + 4) >> 3;
  fwd_scale(v[1], v[2]);
  fwd_scale(v[0], v[3]);
  v[2] += (v[1] + 1) >> 1;
  v[3] += (v[0] + 1) >> 1;
  v[1] -= v[2];
  v[0] -= v[3];
}

void fwd_hadamard_pre(i32 *v) {
  v[1] -= v[2];
  v[0] += v[3];
  std::swap(v[2], v[3]);
  v[2] = ((v[0] - v[1]) >> 1) - v[2];
  v[3] += v[1] >> 1;
  v[0] -= (v[3] * 3 + 4) >> 3;
  v[1] += v[2];
}

void fwd_todd_odd_pre(i32 *v) {
  v[3] += v[0];
  v[2] -= v[1];
  i32 first = v[3] >> 1, second = v[2] >> 1;
  v[0] -= first;
  v[1] += second;
  v[0] += (v[1] * 3 + 4) >> 3;
  v[1] -= (v[0] * 3 + 2) >> 2;
  v[0] += (v[1] * 3 + 6) >> 3;
  v[1] -= second;
  v[0] += first;
  v[2] += v[1];
  v[3] -= v[0];
}

void forward_overlap4x4(i32 *v) {
  static const int G[4][4] = {{0, 3, 12, 15}, {1, 2, 13, 14}, {4, 7, 8, 11}, {5, 6, 9, 10}};
  static const int S[4][2] = {{0, 15}, {1, 14}, {4, 11}, {5, 10}};
  static const int R[4][2] = {{13, 12}, {9, 8}, {7, 3}, {6, 2}};
  for (auto &g : G) grp(v, g[0], g[1], g[2], g[3], fwd_hadamard_pre);
  for (auto &s : S) fwd_scale(v[s[0]], v[s[1]]);
  grp(v, 10, 11, 14, 15, fwd_todd_odd_pre);
  for (auto &r : R) fwd_rotate(v[r[0]], v[r[1]]);
  for (auto &g : G) grp(v, g[0], g[1], g[2], g[3], [](i32 *x) { t2x2h(x, 0); });
}

void forward_overlap2(i32 &a, i32 &b) {
  b -= (a + 2) >> 2;
  a -= b >> 13;
  a -= b >> 9;
  a -= b >> 5;
  a -= (b + 1) >> 1;
  b -= (a + 2) >> 2;
}

void forward_overlap2x2(i32 *v) {
  v[1] += v[2];
  v[0] += v[3];
  v[2] -= (v[1] + 1) >> 1;
  v[3] -= (v[0] + 1) >> 1;
  forward_overlap2(v[0], v[1]);
  v[2] += (v[1] + 1) >> 1;
  v[3] += (v[0] + 1) >> 1;
  v[1] -= v[2];
  v[0] -= v[3];
}

// ---- prefilter schedules -----------------------------------------------
struct Regions {
  std::vector<uint32_t> cols, rows;  // boundaries incl. 0 and extent (units of the level)
};

void prefilter_full(std::vector<i32> &s, uint32_t stride, const Regions &reg) {
  auto f4 = [&](size_t a, size_t b, size_t c, size_t d) {
    i32 v[4] = {s[a], s[b], s[c], s[d]};
    forward_overlap4(v);
    s[a] = v[0];
    s[b] = v[1];
    s[c] = v[2];
    s[d] = v[3];
  };
  for (size_t j = 0; j + 1 < reg.rows.size(); j++)
    for (size_t i = 0; i + 1 < reg.cols.size(); i++) {
      uint32_t L = reg.cols[i], R = reg.cols[i + 1], T = reg.rows[j], B = reg.rows[j + 1];
      for (uint32_t y = T + 2; y + 2 < B; y += 4)
        for (uint32_t x = L + 2; x + 2 < R; x += 4) {
          i32 v[16];
          for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) v[r * 4 + c] = s[(size_t)(y + r) * stride + x + c];
          forward_overlap4x4(v);
          for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) s[(size_t)(y + r) * stride + x + c] = v[r * 4 + c];
        }
      for (uint32_t y = T + 2; y + 2 < B; y += 4)
        for (uint32_t x : {L, L + 1, R - 2, R - 1})
          f4((size_t)y * stride + x, (size_t)(y + 1) * stride + x, (size_t)(y + 2) * stride + x,
             (size_t)(y + 3) * stride + x);
      for (uint32_t x = L + 2; x + 2 < R; x += 4)
        for (uint32_t y : {T, T + 1, B - 2, B - 1}) {
          size_t b = (size_t)y * stride + x;
          f4(b, b + 1, b + 2, b + 3);
        }
      const uint32_t cx[4] = {L, R - 2, L, R - 2}, cy[4] = {T, T, B - 2, B - 2};
      for (int k = 0; k < 4; k++) {
        size_t b = (size_t)cy[k] * stride + cx[k];
        f4(b, b + 1, b + stride, b + stride + 1);
      }
    }
}

// The decoder's ordered subsampled schedule, run last to first with every
// step inverted.
void prefilter_subsampled(std::vector<i32> &s, uint32_t stride, const Regions &reg) {
  auto at = [&](uint32_t x, uint32_t y) -> i32 & { return s[(size_t)y * stride + x]; };
  auto boundary_row = [&](uint32_t L, uint32_t R, uint32_t row) {
    // decoder: c0 -= n0; c1 -= n1; pairs; c0 += n0; c1 += n1
    at(R - 1, row) -= at(R - 2, row);
    at(L, row) -= at(L + 1, row);
    for (uint32_t x = L + 2; x < R; x += 2) forward_overlap2(at(x - 1, row), at(x, row));
    at(R - 1, row) += at(R - 2, row);
    at(L, row) += at(L + 1, row);
  };
  for (size_t j = 0; j + 1 < reg.rows.size(); j++)
    for (size_t i = 0; i + 1 < reg.cols.size(); i++) {
      uint32_t L = reg.cols[i], R = reg.cols[i + 1], T = reg.rows[j], B = reg.rows[j + 1];
      boundary_row(L, R, B - 1);
      for (uint32_t y = T + 2; y < B; y += 2) {
        forward_overlap2(at(R - 1, y - 1), at(R - 1, y));
        forward_overlap2(at(L, y - 1), at(L, y));
        for (uint32_t x = L + 2; x < R; x += 2) {
          i32 v[4] = {at(x - 1, y - 1), at(x, y - 1), at(x - 1, y), at(x, y)};
          forward_overlap2x2(v);
          at(x - 1, y - 1) = v[0];
          at(x, y - 1) = v[1];
          at(x - 1, y) = v[2];
          at(x, y) = v[3];
        }
      }
      boundary_row(L, R, T);
    }
}

// ---- quantization ---------------------------------------------------------
uint32_t quant_map(uint32_t qp, bool scaled, uint32_t shift) {
  if (qp == 0) return 1;
  uint32_t m, e;
  if (scaled) {
    if (qp < 16) {
      m = qp;
      e = shift;
    } else {
      m = 16 + qp % 16;
      e = (qp >> 4) - 1 + shift;
    }
  } else if (qp < 32) {
    m = (qp + 3) >> 2;
    e = 0;
  } else if (qp < 48) {
    m = (17 + qp % 16) / 2;
    e = (qp >> 4) - 2;
  } else {
    m = 16 + qp % 16;
    e = (qp >> 4) - 3;
  }
  return m << e;
}

// Quantizer::quantize (quantization.rs:27-48): the magnitude plus `offset`,
// divided by the step toward zero; an offset of a step or more counts as
// one less than the step.
inline i32 quantize(i32 c, uint32_t step, uint32_t offset) {
  if (offset >= step) offset = step - 1;
  uint32_t mag = c < 0 ? 0u - (uint32_t)c : (uint32_t)c;
  uint32_t carry = (mag % step) >= (step - offset) ? 1u : 0u;
  if (offset == 0) carry = 0;
  uint32_t q = mag / step + carry;
  return c < 0 ? -(i32)q : (i32)q;
}

inline uint32_t mix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// ---- sample preparation ------------------------------------------------
int32_t depth_bias(uint32_t kind, uint32_t shift_bits) {
  if (kind == JXRHIP_INPUT_U8) return 128;
  if (kind == JXRHIP_INPUT_U16) return 32768 >> shift_bits;
  return 0;
}

struct PlaneIn {
  uint32_t w = 0, h = 0;  // coded size of this component
  uint32_t bc = 4, br = 4;
  std::vector<i32> s;
};

using jxr::ArenaOut;

std::atomic<uint64_t> g_identity{1};

void downsample_line(const i32 *in, size_t stride_in, uint32_t length, i32 *out,
                     size_t stride_out, uint32_t count) {
  static const i64 W[5] = {1, 4, 6, 4, 1};
  uint32_t last = length ? length - 1 : 0;
  for (uint32_t index = 0; index < count; index++) {
    i64 centre = 2 * (i64)index, sum = 8;
    for (int t = 0; t < 5; t++) {
      i64 pos = centre + t - 2;
      if (pos < 0)
        pos = -pos;
      else if (pos > (i64)last)
        pos = (i64)last - (pos - (i64)last) < 0 ? 0 : (i64)last - (pos - (i64)last);
      if (pos > (i64)last) pos = last;
      sum += W[t] * in[(size_t)pos * stride_in];
    }
    out[(size_t)index * stride_out] = (i32)(sum >> 4);
  }
}

// One component's quantized macroblocks in arena (tile-major) order: block k
// of a macroblock holds its quantized DC/low-pass value at 16k and its
// quantized high-pass coefficients at 16k+1..16k+15.
struct PlaneQ {
  uint32_t bc = 4, br = 4;
  std::vector<i32> q;          // [macroblock][blocks * 16]
  std::vector<uint32_t> steps; // [macroblock][3]
  std::vector<uint8_t> qps;    // [macroblock][3]
  std::vector<uint32_t> cx, cy;
};

PlaneQ quantize_plane(const PlaneIn &p, uint32_t component, const jxrhip_encode_params &prm,
                      bool scale_after_first, uint32_t qp_dc, uint32_t qp_lp, uint32_t qp_hp,
                      const std::vector<uint32_t> &tcol, const std::vector<uint32_t> &trow) {
  const uint32_t bc = p.bc, br = p.br, blocks = bc * br;
  const uint32_t mbx = p.w / (bc * 4), mby = p.h / (br * 4);
  std::vector<i32> s = p.s;
  Regions sreg, lreg;
  auto bounds = [](const std::vector<uint32_t> &sizes, uint32_t scale, bool hard, uint32_t total) {
    std::vector<uint32_t> b{0};
    if (hard) {
      uint32_t acc = 0;
      for (size_t i = 0; i + 1 < sizes.size(); i++) {
        acc += sizes[i];
        b.push_back(acc * scale);
      }
    }
    b.push_back(total * scale);
    return b;
  };
  sreg.cols = bounds(tcol, bc * 4, prm.hard_tiles != 0, mbx);
  sreg.rows = bounds(trow, br * 4, prm.hard_tiles != 0, mby);
  lreg.cols = bounds(tcol, bc, prm.hard_tiles != 0, mbx);
  lreg.rows = bounds(trow, br, prm.hard_tiles != 0, mby);
  if (prm.overlap >= 1) prefilter_full(s, p.w, sreg);
  const uint32_t lw = mbx * bc, lh = mby * br;
  std::vector<i32> low((size_t)lw * lh);
  std::vector<i32> high((size_t)lw * lh * 16);
  for (uint32_t by = 0; by < lh; by++)
    for (uint32_t bx = 0; bx < lw; bx++) {
      i32 v[16];
      for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) v[r * 4 + c] = s[(size_t)(by * 4 + r) * p.w + bx * 4 + c];
      forward_core(v);
      low[(size_t)by * lw + bx] = v[0];
      std::memcpy(&high[((size_t)by * lw + bx) * 16], v, sizeof v);
    }
  if (prm.overlap == 2) {
    if (bc == 4)
      prefilter_full(low, lw, lreg);
    else
      prefilter_subsampled(low, lw, lreg);
  }
  const bool scaled = prm.scaled != 0;
  const uint32_t lb_shift = component == 0 ? 1 : 0;
  // Tile-major arena order (tile_decode/output.rs:232-299).
  std::vector<uint32_t> tx{0}, ty{0};
  for (uint32_t w : tcol) tx.push_back(tx.back() + w);
  for (uint32_t h : trow) ty.push_back(ty.back() + h);
  PlaneQ out;
  out.bc = bc;
  out.br = br;
  const size_t mbs = (size_t)mbx * mby;
  out.q.resize(mbs * blocks * 16);
  out.steps.reserve(mbs * 3);
  out.qps.reserve(mbs * 3);
  out.cx.reserve(mbs);
  out.cy.reserve(mbs);
  size_t m = 0;
  for (size_t tj = 0; tj + 1 < ty.size(); tj++)
    for (size_t ti = 0; ti + 1 < tx.size(); ti++) {
      uint32_t tile_vary = prm.vary_quantizers ? mix(prm.seed * 31u + (uint32_t)(tj * 977 + ti)) % 5 : 0;
      for (uint32_t my = ty[tj]; my < ty[tj + 1]; my++)
        for (uint32_t mx = tx[ti]; mx < tx[ti + 1]; mx++, m++) {
          uint32_t mb_vary =
              prm.vary_quantizers ? mix(prm.seed * 7919u + my * 65599u + mx) % 3 : 0;
          uint32_t qd = qp_dc ? qp_dc + tile_vary : 0;
          uint32_t ql = qp_lp ? qp_lp + tile_vary + mb_vary : 0;
          uint32_t qh = qp_hp ? qp_hp + tile_vary + 2 * mb_vary : 0;
          qd = qd > 255 ? 255 : qd;
          ql = ql > 255 ? 255 : ql;
          qh = qh > 255 ? 255 : qh;
          // a band that is not coded has no parameters of its own: it takes
          // those of the band below, as a decoder does
          // (tile_decode/quantizer.rs parse_low_pass / parse_high_pass)
          if (prm.bands == JXRHIP_BANDS_DC_ONLY) ql = qd;
          if (prm.bands <= JXRHIP_BANDS_NO_HIGH_PASS) qh = ql;
          uint32_t sdc = quant_map(qd, scaled, lb_shift);
          uint32_t slp = quant_map(ql, scaled, lb_shift);
          uint32_t shp = quant_map(qh, scaled, 1);
          i32 lp[16] = {0};
          for (uint32_t r = 0; r < br; r++)
            for (uint32_t c = 0; c < bc; c++) {
              i32 v = low[(size_t)(my * br + r) * lw + mx * bc + c];
              lp[r * bc + c] = v;
            }
          if (bc == 4)
            forward_core(lp);
          else if (br == 4)
            forward_chroma422(lp);
          else
            forward_chroma420(lp);
          i32 *q = &out.q[m * blocks * 16];
          for (uint32_t k = 0; k < blocks; k++) {
            uint32_t bx = mx * bc + k % bc, by = my * br + k / bc;
            const i32 *h = &high[((size_t)by * lw + bx) * 16];
            // reconstruction doubles a scaled component's DC and low-pass
            // band after its first level: they are divided by twice their
            // step, rounding to nearest (analyze.rs:500-530)
            const uint32_t step = k == 0 ? sdc : slp;
            const uint32_t divisor = scale_after_first ? 2 * step : step;
            q[k * 16] = quantize(lp[k], divisor, divisor / 2);
            // the high-pass band's zero bin is wider: an offset of 3/8 step
            // (analyze.rs:532-536)
            for (int j = 1; j < 16; j++) q[k * 16 + j] = quantize(h[j], shp, shp / 8 * 3);
          }
          out.steps.insert(out.steps.end(), {sdc, slp, shp});
          out.qps.insert(out.qps.end(), {(uint8_t)qd, (uint8_t)ql, (uint8_t)qh});
          out.cx.push_back(mx);
          out.cy.push_back(my);
        }
    }
  return out;
}

inline uint32_t sat_add(uint32_t a, uint32_t b) { return a > UINT32_MAX - b ? UINT32_MAX : a + b; }
inline uint32_t sat_times4(uint32_t a) { return a > UINT32_MAX / 4 ? UINT32_MAX : a * 4; }
inline uint32_t magnitude(i32 v) { return v < 0 ? 0u - (uint32_t)v : (uint32_t)v; }

// The high-pass direction of every macroblock of one coded plane, from its
// quantized low-pass coefficients, as a decoder derives it
// (ComponentSyntax::high_pass_mode, tile_decode/syntax.rs:173-205): the
// luma's energy across (positions 1, 2, 3) and down (4, 8, 12), plus, for a
// YUV plane, its two chroma components' at their sampling's positions; four
// times more energy one way picks the other direction. Every component of
// the plane shares it; a plane without a high-pass band has none.
std::vector<uint8_t> hp_directions(const std::vector<const PlaneQ *> &planes, uint32_t color, uint32_t bands) {
  const size_t mbs = planes[0]->cx.size();
  std::vector<uint8_t> dirs(mbs, JXRHIP_PRED_NONE);
  if (bands < JXRHIP_BANDS_NO_FLEXBITS) return dirs;
  const bool yuv = color == JXRHIP_COLOR_YUV420 || color == JXRHIP_COLOR_YUV422 || color == JXRHIP_COLOR_YUV444;
  std::vector<int> across, down;
  if (color == JXRHIP_COLOR_YUV420) {
    across = {1};
    down = {2};
  } else if (color == JXRHIP_COLOR_YUV422) {
    across = {1, 5};
    down = {2, 6};
  } else {
    across = {1};
    down = {4};
  }
  for (size_t m = 0; m < mbs; m++) {
    auto lp = [&](size_t c, int position) {
      const PlaneQ &pq = *planes[c];
      return magnitude(pq.q[(m * pq.bc * pq.br + position) * 16]);
    };
    uint32_t hor = 0, ver = 0;
    for (int position : {1, 2, 3}) hor = sat_add(hor, lp(0, position));
    for (int position : {4, 8, 12}) ver = sat_add(ver, lp(0, position));
    if (yuv && planes.size() >= 3)
      for (size_t c = 1; c < 3; c++) {
        for (int position : across) hor = sat_add(hor, lp(c, position));
        for (int position : down) ver = sat_add(ver, lp(c, position));
      }
    if (sat_times4(hor) < ver)
      dirs[m] = JXRHIP_PRED_FROM_LEFT;
    else if (sat_times4(ver) < hor)
      dirs[m] = JXRHIP_PRED_FROM_TOP;
  }
  return dirs;
}

// The arena of one component: its macroblocks' DC/low-pass values and their
// high-pass residuals after the shared direction's prediction is removed
// (the forward of the decoder's running sums, prediction.rs:53-97).
void emit_plane(const PlaneQ &pq, const std::vector<uint8_t> &dirs, uint32_t bands, ArenaOut &out) {
  const uint32_t bc = pq.bc, br = pq.br, blocks = bc * br;
  jxrhip_coefficient_plane cp{};
  cp.coefficient_offset = out.coefficients.size();
  cp.macroblock_offset = out.offsets.size();
  cp.block_columns = bc;
  cp.block_rows = br;
  for (size_t m = 0; m < pq.cx.size(); m++) {
    const i32 *q = &pq.q[m * blocks * 16];
    const uint8_t pred = dirs[m];
    i32 res[256];
    std::memcpy(res, q, sizeof(i32) * blocks * 16);
    if (pred == JXRHIP_PRED_FROM_LEFT) {
      for (uint32_t r = 0; r < br; r++)
        for (uint32_t c = 1; c < bc; c++)
          for (int pos : {4, 8, 12})
            res[(r * bc + c) * 16 + pos] = q[(r * bc + c) * 16 + pos] - q[(r * bc + c - 1) * 16 + pos];
    } else if (pred == JXRHIP_PRED_FROM_TOP) {
      for (uint32_t r = 1; r < br; r++)
        for (uint32_t c = 0; c < bc; c++)
          for (int pos : {1, 2, 3})
            res[(r * bc + c) * 16 + pos] = q[(r * bc + c) * 16 + pos] - q[((r - 1) * bc + c) * 16 + pos];
    }
    out.offsets.push_back((uint32_t)out.coefficients.size());
    out.quantizers.insert(out.quantizers.end(), &pq.steps[m * 3], &pq.steps[m * 3] + 3);
    out.qps.insert(out.qps.end(), &pq.qps[m * 3], &pq.qps[m * 3] + 3);
    out.bands.push_back((uint8_t)bands);
    out.preds.push_back(pred);
    out.coded_x.push_back(pq.cx[m]);
    out.coded_y.push_back(pq.cy[m]);
    if (bands == JXRHIP_BANDS_DC_ONLY) {
      out.coefficients.push_back(res[0]);
    } else if (bands == JXRHIP_BANDS_NO_HIGH_PASS) {
      for (uint32_t k = 0; k < blocks; k++) out.coefficients.push_back(res[k * 16]);
    } else {
      out.coefficients.insert(out.coefficients.end(), res, res + blocks * 16);
    }
  }
  cp.coefficient_count = out.coefficients.size() - cp.coefficient_offset;
  cp.macroblock_count = out.offsets.size() - cp.macroblock_offset;
  out.planes.push_back(cp);
}

void finish_arena(ArenaOut &o, const jxrhip_encode_params &prm) {
  jxrhip_arena &a = o.arena;
  std::memset(&a, 0, sizeof a);
  a.coefficients = o.coefficients.data();
  a.coefficient_count = o.coefficients.size();
  a.macroblock_count = o.offsets.size();
  a.coefficient_offsets = o.offsets.data();
  a.quantizers = o.quantizers.data();
  a.bands = o.bands.data();
  a.hp_predictions = o.preds.data();
  a.coded_x = o.coded_x.data();
  a.coded_y = o.coded_y.data();
  a.planes = o.planes.data();
  a.plane_count = (uint32_t)o.planes.size();
  a.overlap = prm.overlap;
  a.hard_tiles = prm.hard_tiles;
  a.tile_column_count = (uint32_t)o.tile_w.size();
  a.tile_column_widths = o.tile_w.data();
  a.tile_row_count = (uint32_t)o.tile_h.size();
  a.tile_row_heights = o.tile_h.data();
  a.identity = g_identity.fetch_add(1);
}

}  // namespace

namespace jxr {

// DC and low-pass coefficients of a flat component of `level` at unit
// steps, of mbx x mby macroblocks of bc x br blocks whose regions end at the
// macroblock boundaries `cols` x `rows` (from 0 to the extent): the forward
// stages full reconstruction inverts (sample prefilter, block transforms,
// low-pass prefilter, first-level transforms; analyze.rs:56-86), which the
// sample scale of a reduced decode probes (sample_scale.rs:157-213). A
// component whose low bands reconstruction doubles has them divided by two,
// its magnitude rounded half up (analyze.rs:500-530 at unit steps).
// `out` receives bc * br values per macroblock, raster order.
void flat_probe_coefficients(int32_t level, uint32_t bc, uint32_t br, bool doubled, uint32_t overlap,
                             const std::vector<uint32_t> &cols, const std::vector<uint32_t> &rows,
                             std::vector<int32_t> &out) {
  const uint32_t mbx = cols.back(), mby = rows.back(), blocks = bc * br;
  const uint32_t sw = mbx * bc * 4, lw = mbx * bc, lh = mby * br;
  std::vector<i32> s((size_t)sw * mby * br * 4, level);
  Regions sreg, lreg;
  for (uint32_t c : cols) {
    sreg.cols.push_back(c * bc * 4);
    lreg.cols.push_back(c * bc);
  }
  for (uint32_t r : rows) {
    sreg.rows.push_back(r * br * 4);
    lreg.rows.push_back(r * br);
  }
  if (overlap >= 1) prefilter_full(s, sw, sreg);
  std::vector<i32> low((size_t)lw * lh);
  for (uint32_t by = 0; by < lh; by++)
    for (uint32_t bx = 0; bx < lw; bx++) {
      i32 v[16];
      for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) v[r * 4 + c] = s[(size_t)(by * 4 + r) * sw + bx * 4 + c];
      forward_core(v);
      low[(size_t)by * lw + bx] = v[0];
    }
  if (overlap == 2) {
    if (bc == 4)
      prefilter_full(low, lw, lreg);
    else
      prefilter_subsampled(low, lw, lreg);
  }
  out.assign((size_t)mbx * mby * blocks, 0);
  for (uint32_t my = 0; my < mby; my++)
    for (uint32_t mx = 0; mx < mbx; mx++) {
      i32 v[16] = {0};
      for (uint32_t r = 0; r < br; r++)
        for (uint32_t c = 0; c < bc; c++) v[r * bc + c] = low[(size_t)(my * br + r) * lw + mx * bc + c];
      if (bc == 4)
        forward_core(v);
      else if (br == 4)
        forward_chroma422(v);
      else
        forward_chroma420(v);
      for (uint32_t k = 0; k < blocks; k++) {
        i32 q = v[k];
        if (doubled) {
          const uint32_t mag = q < 0 ? 0u - (uint32_t)q : (uint32_t)q;
          const uint32_t h = mag / 2 + (mag & 1u);
          q = q < 0 ? -(i32)h : (i32)h;
        }
        out[((size_t)my * mbx + mx) * blocks + k] = q;
      }
    }
}

}  // namespace jxr


extern "C" uint32_t jxrhip_quant_step(uint32_t qp, uint32_t scaled, uint32_t scaled_shift) {
  return quant_map(qp > 255 ? 255 : qp, scaled != 0, scaled_shift);
}

// ---- quality scale (crates/jxr-native/src/encode/quality.rs) --------------
// A quality names a fidelity calibrated on 8-bit pictures: a level in half
// parameters of luma, split between luma (rounded up) and the colour
// differences (rounded down), with band offsets, a chroma offset by
// sampling, and for other sample ranges the parameter whose step is as
// large against the range as the calibrated step is against 255.
namespace {

constexpr uint32_t kReferenceRange = 255;  // quality.rs:27
constexpr uint32_t kOneSample = 4;         // step of one sample value (quality.rs:32)
constexpr uint32_t kFinestLossy = 5;       // quality.rs:35
constexpr uint32_t kWholeStepsBelow = 32;  // quality.rs:38

// Step of the first component's high-pass band at `parameter` in scaled
// arithmetic: only ratios of steps are compared (quality.rs:387-392).
unsigned __int128 scale_step(uint32_t parameter) { return quant_map(parameter, true, 1); }

// Whether the next step up is farther from `target` than `nearest`, i.e.
// target lies below their geometric mean (quality.rs:382-384).
bool past_target(unsigned __int128 candidate, unsigned __int128 nearest, unsigned __int128 target) {
  return candidate * nearest > target * target;
}

// Level of a lossy quality in half luma parameters (quality.rs:416-425).
uint32_t quality_level(uint32_t q) {
  if (q >= 99) return 10;
  if (q >= 93) return 16 + 8 * (98 - q);
  if (q >= 91) return 64 + 3 * (92 - q);
  if (q >= 76) return 70 + 2 * (90 - q);
  if (q >= 11) return 100 + (75 - q);
  return 165 + 6 * (10 - q);
}

// Parameter of a colour difference whose channels take `p`: sixteen more
// at full resolution, eight at 4:2:2, four at 4:2:0, in proportion below
// sixteen (quality.rs:448-461).
uint32_t difference_parameter(uint32_t p, uint32_t color_format) {
  uint32_t offset = color_format == JXRHIP_COLOR_YUV420 ? 4 : color_format == JXRHIP_COLOR_YUV422 ? 8 : 16;
  if (p < 16) offset = p * offset / 16;
  return std::min<uint32_t>(p + offset, 255);
}

// DC keeps half the excess of the level over 64, at most sixteen
// parameters, below the AC parameter (quality.rs:470-479).
void band_parameters(uint32_t level, uint32_t parameter, uint32_t out[3]) {
  const uint32_t excess = (level > 64 ? level - 64 : 0) / 2;
  out[0] = parameter - std::min<uint32_t>(excess, 16);
  out[1] = out[2] = parameter;
}

// The parameter whose step is nearest range / 255 times the step of
// `reference`, for a range of at least 255 (quality.rs:369-380).
uint32_t scaled_parameter(uint32_t reference, uint32_t range) {
  const unsigned __int128 target = scale_step(reference) * range;
  uint32_t nearest = reference;
  unsigned __int128 nearest_step = scale_step(reference);
  for (uint32_t p = reference; p <= 255; p++) {
    const unsigned __int128 c = scale_step(p);
    if (past_target(c * kReferenceRange, nearest_step * kReferenceRange, target)) break;
    nearest = p;
    nearest_step = c;
  }
  return nearest;
}

// A range below 255: 5 and the multiples of four below 32, any parameter
// from 32; a step down to one sample value is lossless, and every band and
// component follows the primary parameter (quality.rs:340-363).
uint32_t reduced_parameter(uint32_t primary, uint32_t range) {
  const unsigned __int128 target = scale_step(primary) * range;
  uint32_t nearest = kOneSample;
  unsigned __int128 nearest_step = scale_step(kOneSample);
  for (uint32_t p = kFinestLossy; p <= 255; p++) {
    if (!(p == kFinestLossy || p >= kWholeStepsBelow || p % 4 == 0)) continue;
    const unsigned __int128 c = scale_step(p);
    if (past_target(c * kReferenceRange, nearest_step * kReferenceRange, target)) break;
    nearest = p;
    nearest_step = c;
  }
  return nearest == kOneSample ? 0u : nearest;
}

// Overlap mode of a lossy quality: none at 99, second level at 6 and below
// (quality.rs:493-499).
uint32_t quality_overlap(uint32_t q) { return q >= 99 ? 0u : q >= 7 ? 1u : 2u; }

}  // namespace

extern "C" jxrhip_status jxrhip_quality_parameters(uint32_t quality, uint32_t color_format, uint32_t rgb_source,
                                                   uint32_t range, uint32_t luma[3], uint32_t chroma[3],
                                                   uint32_t *overlap) {
  return jxrhip_quality_parameters_scaled(quality, color_format, rgb_source, range, 0, 0, luma, chroma, overlap);
}

extern "C" jxrhip_status jxrhip_quality_parameters_scaled(uint32_t quality, uint32_t color_format,
                                                          uint32_t rgb_source, uint32_t range, uint32_t is_float,
                                                          uint32_t lowest_quality, uint32_t luma[3],
                                                          uint32_t chroma[3], uint32_t *overlap) {
  if (!luma || !chroma || !overlap) {
    jxrhip_set_error("null argument");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  if (quality > 100 || range == 0) {
    jxrhip_set_error("quality above 100 or empty sample range");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  const bool difference = rgb_source && color_format != JXRHIP_COLOR_LUMA;
  if (quality >= 100) {  // lossless, and no overlap unless asked for (quality.rs:127-128,300-302)
    for (int b = 0; b < 3; b++) luma[b] = chroma[b] = 0;
    *overlap = 0;
    return JXRHIP_OK;
  }
  // every quality below the scale's lowest codes as it does (quality.rs:303)
  const uint32_t q = std::max(quality, std::min<uint32_t>(lowest_quality, 99));
  const uint32_t level = quality_level(q);
  const uint32_t primary = (level + 1) / 2;
  if (range < kReferenceRange) {
    const uint32_t p = reduced_parameter(primary, range);
    for (int b = 0; b < 3; b++) {
      luma[b] = p;
      chroma[b] = p && difference ? difference_parameter(p, color_format) : p;
    }
    // lossless coding takes no overlap (quality.rs:126-128)
    *overlap = p == 0 ? 0u : quality_overlap(q);
    return JXRHIP_OK;
  }
  // float patterns keep the differences at the primary's whole levels
  // (quality.rs:312-315)
  const uint32_t diff = is_float ? primary : level / 2;
  band_parameters(primary, primary, luma);
  if (difference)
    band_parameters(diff, difference_parameter(diff, color_format), chroma);
  else
    band_parameters(primary, primary, chroma);
  for (int b = 0; b < 3; b++) {
    luma[b] = scaled_parameter(luma[b], range);
    chroma[b] = scaled_parameter(chroma[b], range);
  }
  *overlap = quality_overlap(q);
  return JXRHIP_OK;
}

namespace {
constexpr int32_t kFloatOctavesLog2 = 3;     // eight octaves for the 8-bit range (quality.rs:41-51)
constexpr int32_t kLogarithmicFromBias = 6;  // quality.rs:53-61
constexpr int32_t kAlphaOctavesLog2 = 1;     // two octaves for alpha (quality.rs:63-72)
constexpr uint32_t kRgbeLowestQuality = 10;  // quality.rs:74-81

// The coded distance of 2^octaves_log2 octaves of float patterns with
// `mantissa_bits` mantissa bits; a shift of 32 or more gives none, and the
// range is at least one (quality.rs:259-274).
uint32_t pattern_range(uint32_t mantissa_bits, int32_t octaves_log2) {
  const uint32_t magnitude = (uint32_t)(octaves_log2 < 0 ? -(int64_t)octaves_log2 : octaves_log2);
  const uint32_t octave = 1u << std::min<uint32_t>(mantissa_bits, 23);
  uint32_t range = 0;
  if (magnitude < 32) range = octaves_log2 >= 0 ? octave << magnitude : octave >> magnitude;
  return std::max<uint32_t>(range, 1);
}
}  // namespace

extern "C" jxrhip_status jxrhip_sample_scale(uint32_t depth, uint32_t output_color, uint32_t shift_bits,
                                             uint32_t mantissa_length, int32_t exponent_bias, uint32_t alpha,
                                             uint32_t *range, uint32_t *is_float, uint32_t *lowest_quality) {
  if (!range || !is_float || !lowest_quality) {
    jxrhip_set_error("null argument");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  *is_float = 0;
  *lowest_quality = 0;
  const bool f16 = depth == JXRHIP_DEPTH_F16, f32 = depth == JXRHIP_DEPTH_F32;
  if (alpha && (f16 || f32)) {
    // the patterns of two octaves, or those up to one if fewer
    // (quality.rs:238-257)
    uint32_t bits = f16 ? 10u : mantissa_length, one = 0x3c00;
    if (f32) {
      int32_t sample = 0;
      if (jxrhip_float32_sample(0x3f800000u, 0, mantissa_length, exponent_bias, &sample) != JXRHIP_OK)
        return JXRHIP_ERR_INVALID_ARGUMENT;
      one = sample > 0 ? (uint32_t)sample : 0u;
    }
    *range = std::max<uint32_t>(std::min<uint32_t>(pattern_range(bits, kAlphaOctavesLog2), one), 1);
    *is_float = 1;
    return JXRHIP_OK;
  }
  if (!alpha && depth == JXRHIP_DEPTH_U8 && output_color == JXRHIP_COLOR_RGBE) {
    // an exponent above a 7-bit mantissa (quality.rs:216-220)
    *range = pattern_range(7, kFloatOctavesLog2);
    *is_float = 1;
    *lowest_quality = kRgbeLowestQuality;
    return JXRHIP_OK;
  }
  if (f16 || f32) {
    // a 32-bit float coded linearly below 2^(1 - bias) halves the range
    // with each step of bias below six (quality.rs:221-230)
    const int32_t linear = f32 ? std::max<int32_t>(kLogarithmicFromBias - exponent_bias, 0) : 0;
    *range = pattern_range(f16 ? 10u : mantissa_length, kFloatOctavesLog2 - linear);
    *is_float = 1;
    return JXRHIP_OK;
  }
  // integers span their range, fixed point zero to one, less the dropped
  // bits (quality.rs:276-291, output_policy.rs:99-108)
  switch (depth) {
    case JXRHIP_DEPTH_U16: *range = 65535u >> std::min<uint32_t>(shift_bits, 31); break;
    case JXRHIP_DEPTH_I16: *range = (1u << 13) >> std::min<uint32_t>(shift_bits, 31); break;
    case JXRHIP_DEPTH_I32: *range = (1u << 24) >> std::min<uint32_t>(shift_bits, 31); break;
    case JXRHIP_DEPTH_BIT1_WHITE:
    case JXRHIP_DEPTH_BIT1_BLACK: *range = 1; break;
    case JXRHIP_DEPTH_U10:
    case JXRHIP_DEPTH_RGB101010: *range = 1023; break;
    case JXRHIP_DEPTH_RGB555: *range = 31; break;
    case JXRHIP_DEPTH_RGB565: *range = 63; break;  // red and blue coded at twice their five bits
    default: *range = 255; break;
  }
  *range = std::max<uint32_t>(*range, 1);
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_float32_sample(uint32_t bits, uint32_t scaled, uint32_t length,
                                               int32_t exponent_bias, int32_t *sample) {
  if (length > 23 || !sample) {
    jxrhip_set_error("BD32F mantissa length greater than 23");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  i64 exponent = (bits >> 23) & 0xff;
  uint64_t mantissa = bits & 0x7fffff;
  i64 magnitude;
  auto round_shift = [](uint64_t v, uint32_t sh) -> uint64_t {
    if (sh == 0) return v;
    if (sh >= 64) return 0;
    return (v + ((uint64_t)1 << (sh - 1))) >> sh;
  };
  if (exponent == 0 && mantissa == 0) {
    magnitude = 0;
  } else {
    if (exponent == 0) {
      exponent = 1;
      while (mantissa < (1u << 23)) {
        mantissa <<= 1;
        exponent -= 1;
      }
    } else {
      mantissa |= 1u << 23;
    }
    i64 coded = exponent - 127 + exponent_bias;
    if (coded >= 1) {
      uint64_t fraction = round_shift(mantissa & 0x7fffff, 23 - length);
      magnitude = (coded << length) + (i64)fraction;
    } else {
      i64 sh = (i64)(23 - length) + 1 - coded;
      magnitude = (i64)round_shift(mantissa, sh > 63 ? 64u : (uint32_t)sh);
    }
  }
  i64 signed_value = (bits >> 31) ? -magnitude : magnitude;
  if (signed_value > INT32_MAX || signed_value < INT32_MIN) {
    jxrhip_set_error("floating-point sample outside i32");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  // F32 depth: no bias; scaled streams carry three more bits.
  i64 v = signed_value * (scaled ? 8 : 1);
  if (v > INT32_MAX || v < INT32_MIN) {
    jxrhip_set_error("scaled floating-point sample outside i32");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  *sample = (i32)v;
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_encode(const void *pixels, uint64_t pixel_bytes,
                                       const jxrhip_encode_params *prm, jxrhip_encoded **out) {
  if (!pixels || !prm || !out) {
    jxrhip_set_error("null argument");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  const uint32_t W = prm->width, H = prm->height, C = prm->input_channels;
  const uint32_t cf = prm->color_format;
  const uint32_t esize = prm->input_kind == JXRHIP_INPUT_U8 ? 1 : prm->input_kind == JXRHIP_INPUT_U16 ? 2 : 4;
  if (W == 0 || H == 0 || C == 0 || C > 4 || prm->overlap > 2 || prm->bands > 3 ||
      (uint64_t)W * H * C * esize != pixel_bytes) {
    jxrhip_set_error("encode parameters do not describe the pixel buffer");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  const bool gray = C <= 2;
  const bool has_alpha = C == 2 || C == 4;
  if ((gray && cf != JXRHIP_COLOR_LUMA) ||
      (!gray && cf != JXRHIP_COLOR_YUV420 && cf != JXRHIP_COLOR_YUV422 && cf != JXRHIP_COLOR_YUV444) ||
      (has_alpha && prm->alpha_mode == JXRHIP_ALPHA_NONE)) {
    jxrhip_set_error("colour format does not match the input channels");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  const uint32_t mbx = (W + 15) / 16, mby = (H + 15) / 16;
  std::vector<uint32_t> tcol, trow;
  if (prm->tile_column_count) tcol.assign(prm->tile_column_widths, prm->tile_column_widths + prm->tile_column_count);
  else tcol = {mbx};
  if (prm->tile_row_count) trow.assign(prm->tile_row_heights, prm->tile_row_heights + prm->tile_row_count);
  else trow = {mby};
  uint64_t sc = 0, sr = 0;
  for (auto v : tcol) sc += v;
  for (auto v : trow) sr += v;
  if (sc != mbx || sr != mby) {
    jxrhip_set_error("tile sizes do not add up to the coded macroblock grid");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  const uint32_t CW = mbx * 16, CH = mby * 16;
  const int scale = prm->scaled ? 3 : 0;
  // Reconstructed-sample value of one input channel value: the inverse of
  // output bias, rounding and shift (ChannelScale::unapply, scaling.rs:129-148).
  auto fetch = [&](uint32_t x, uint32_t y, uint32_t c, bool alpha) -> i32 {
    x = x < W ? x : W - 1;
    y = y < H ? y : H - 1;
    size_t idx = ((size_t)y * W + x) * C + c;
    (void)alpha;
    if (prm->input_kind == JXRHIP_INPUT_U8) {
      i32 v = ((const uint8_t *)pixels)[idx];
      return (v << scale) - (128 << scale);
    }
    if (prm->input_kind == JXRHIP_INPUT_U16) {
      i32 v = ((const uint16_t *)pixels)[idx];
      uint32_t sb = prm->shift_bits;
      return ((v >> sb) << scale) - ((32768 >> sb) << scale);
    }
    uint32_t bits;
    std::memcpy(&bits, (const uint8_t *)pixels + idx * 4, 4);
    i32 s = 0;
    jxrhip_float32_sample(bits, prm->scaled, prm->mantissa_length, prm->exponent_bias, &s);
    return s;
  };
  std::unique_ptr<jxrhip_encoded> enc(new (std::nothrow) jxrhip_encoded());
  if (!enc) {
    jxrhip_set_error("out of host memory");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  std::vector<PlaneIn> comps;
  PlaneIn full;
  full.w = CW;
  full.h = CH;
  if (gray) {
    full.s.resize((size_t)CW * CH);
    for (uint32_t y = 0; y < CH; y++)
      for (uint32_t x = 0; x < CW; x++) full.s[(size_t)y * CW + x] = fetch(x, y, 0, false);
    comps.push_back(full);
  } else {
    PlaneIn Y = full, U = full, V = full;
    Y.s.resize((size_t)CW * CH);
    U.s.resize((size_t)CW * CH);
    V.s.resize((size_t)CW * CH);
    for (uint32_t y = 0; y < CH; y++)
      for (uint32_t x = 0; x < CW; x++) {
        i32 r = fetch(x, y, 0, false), g = fetch(x, y, 1, false), b = fetch(x, y, 2, false);
        // forward of inverse_yuv (output_format/color.rs:74-84)
        i32 v = b - r;
        i32 t = r - g + ((v >> 1) + (v & 1));
        size_t i = (size_t)y * CW + x;
        Y.s[i] = g + (t >> 1);
        U.s[i] = -t;
        V.s[i] = v;
      }
    comps.push_back(Y);
    if (cf == JXRHIP_COLOR_YUV444) {
      comps.push_back(U);
      comps.push_back(V);
    } else {
      for (PlaneIn *src : {&U, &V}) {
        PlaneIn d;
        d.w = CW / 2;
        d.bc = 2;
        d.br = cf == JXRHIP_COLOR_YUV420 ? 2 : 4;
        d.h = cf == JXRHIP_COLOR_YUV420 ? CH / 2 : CH;
        std::vector<i32> half((size_t)d.w * CH);
        for (uint32_t y = 0; y < CH; y++)
          downsample_line(&src->s[(size_t)y * CW], 1, CW, &half[(size_t)y * d.w], 1, d.w);
        d.s.resize((size_t)d.w * d.h);
        if (cf == JXRHIP_COLOR_YUV420) {
          for (uint32_t x = 0; x < d.w; x++)
            downsample_line(&half[x], d.w, CH, &d.s[x], d.w, d.h);
        } else {
          d.s = half;
        }
        comps.push_back(std::move(d));
      }
    }
  }
  PlaneIn alpha_plane;
  if (has_alpha) {
    alpha_plane = full;
    alpha_plane.s.resize((size_t)CW * CH);
    for (uint32_t y = 0; y < CH; y++)
      for (uint32_t x = 0; x < CW; x++) alpha_plane.s[(size_t)y * CW + x] = fetch(x, y, C - 1, true);
  }
  ArenaOut &P = enc->primary;
  P.tile_w = tcol;
  P.tile_h = trow;
  std::vector<PlaneQ> quantized;
  for (uint32_t c = 0; c < comps.size(); c++) {
    const bool chroma = c != 0 && prm->chroma_qp_set;
    quantized.push_back(quantize_plane(comps[c], c, *prm, prm->scaled && c != 0,
                                       chroma ? prm->qp_chroma_dc : prm->qp_dc,
                                       chroma ? prm->qp_chroma_lp : prm->qp_lp,
                                       chroma ? prm->qp_chroma_hp : prm->qp_hp, tcol, trow));
  }
  {
    std::vector<const PlaneQ *> group;
    for (const PlaneQ &pq : quantized) group.push_back(&pq);
    const std::vector<uint8_t> dirs = hp_directions(group, cf, prm->bands);
    for (const PlaneQ &pq : quantized) emit_plane(pq, dirs, prm->bands, P);
  }
  PlaneQ alpha_q;
  if (has_alpha) {
    alpha_q = quantize_plane(alpha_plane, 0, *prm, false, prm->alpha_qp, prm->alpha_qp, prm->alpha_qp, tcol, trow);
  }
  const std::vector<uint8_t> alpha_dirs =
      has_alpha ? hp_directions({&alpha_q}, JXRHIP_COLOR_LUMA, prm->bands) : std::vector<uint8_t>();
  if (has_alpha && prm->alpha_mode == JXRHIP_ALPHA_INTEGRATED) emit_plane(alpha_q, alpha_dirs, prm->bands, P);
  finish_arena(P, *prm);
  if (has_alpha && prm->alpha_mode == JXRHIP_ALPHA_SEPARATE) {
    ArenaOut &A = enc->alpha;
    A.tile_w = tcol;
    A.tile_h = trow;
    emit_plane(alpha_q, alpha_dirs, prm->bands, A);
    finish_arena(A, *prm);
    enc->has_alpha = true;
  }
  enc->params = *prm;
  enc->params.tile_column_widths = nullptr;
  enc->params.tile_row_heights = nullptr;
  enc->params.tile_column_count = (uint32_t)tcol.size();
  enc->params.tile_row_count = (uint32_t)trow.size();
  *out = enc.release();
  return JXRHIP_OK;
}

extern "C" void jxrhip_encoded_destroy(jxrhip_encoded *e) { delete e; }

extern "C" const jxrhip_arena *jxrhip_encoded_arena(const jxrhip_encoded *e, uint32_t which) {
  if (!e) return nullptr;
  if (which == 0) return &e->primary.arena;
  if (which == 1 && e->has_alpha) return &e->alpha.arena;
  return nullptr;
}
