// codestream.cpp — JPEG XR codestream reader and writer (host side).
//
// Reader: Annex-A container and T.832 header parsing, the tile index and
// profile area, and the adaptive-VLC entropy decoding of spatial- and
// frequency-mode tile packets into the coefficient arenas the reconstruction
// takes (crates/jxr-native/src/{codestream,annex_a,tail}.rs, header/,
// tile_decode/, entropy/). Writer: the same syntax driven the other way, from
// the arenas of the forward analysis (crates/jxr-native/src/header/write.rs,
// tile_encode/, entropy/encode/, annex_a/writer.rs).
//
// Both directions share one set of adaptive state machines (code-table
// selectors, adaptive scans, coefficient-normalisation models, coded-block-
// pattern predictors), so a writer and a reader that see the same symbols
// stay in step by construction. Tiles are independent: each is decoded on its
// own host thread straight into its windows of the arena.

#include <algorithm>
#include <array>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/jxr_hip_codestream.h"
#include "encoded.h"

extern "C" void jxrhip_set_error(const char *message);

namespace {

using i32 = int32_t;
using i64 = int64_t;
using u8 = uint8_t;
using u16 = uint16_t;
using u32 = uint32_t;
using u64 = uint64_t;

// ---- failures -------------------------------------------------------------
// Thrown inside this file only; every extern "C" entry point catches it and
// returns its status.
struct Failure {
  jxrhip_status status;
  std::string message;
};
[[noreturn]] void syntax_error(const std::string &what) { throw Failure{JXRHIP_ERR_SYNTAX, what}; }
[[noreturn]] void unsupported(const std::string &what) { throw Failure{JXRHIP_ERR_UNSUPPORTED, what}; }
[[noreturn]] void overflow(const std::string &what) {
  throw Failure{JXRHIP_ERR_OVERFLOW, "arithmetic overflow: " + what};
}

i32 checked_add(i32 a, i32 b, const char *what) {
  const i64 v = (i64)a + b;
  if (v < INT32_MIN || v > INT32_MAX) overflow(what);
  return (i32)v;
}
i32 checked_sub(i32 a, i32 b, const char *what) {
  const i64 v = (i64)a - b;
  if (v < INT32_MIN || v > INT32_MAX) overflow(what);
  return (i32)v;
}
// a left shift that drops the bits it moves out, as a two's-complement
// machine word does
inline i32 wrap_shl(i32 v, unsigned s) { return (i32)((u32)v << s); }
inline u32 magnitude(i32 v) { return v < 0 ? 0u - (u32)v : (u32)v; }
inline u32 sat_add(u32 a, u32 b) { return a > UINT32_MAX - b ? UINT32_MAX : a + b; }
inline u32 sat_mul(u32 a, u32 b) {
  const u64 v = (u64)a * b;
  return v > UINT32_MAX ? UINT32_MAX : (u32)v;
}

// ---- bits -----------------------------------------------------------------
// Most-significant-bit-first reading of one packet (or of the headers); a
// read past its last byte is a truncation. The cache holds `avail_` bits
// at its top; below them it may hold the stream's next bits, loaded early
// by an eight-byte refill, which a later refill ORs in again unchanged.
class BitReader {
 public:
  BitReader(const u8 *data, size_t bytes) : p_(data), nbytes_(bytes) {}
  size_t position() const { return next_ * 8 - (size_t)avail_; }
  size_t remaining() const { return (nbytes_ - next_) * 8 + (size_t)avail_; }
  // n <= 32
  __attribute__((always_inline)) u32 bits(unsigned n) {
    if (n == 0) return 0;
    if ((int)n > avail_) {
      refill();
      if ((int)n > avail_)
        syntax_error("truncated: " + std::to_string(n) + " bits wanted at bit " + std::to_string(position()));
    }
    const u32 v = (u32)(cache_ >> (64 - n));
    cache_ <<= n;
    avail_ -= (int)n;
    return v;
  }
  u64 bits64(unsigned n) {  // n <= 64
    if (n <= 32) return bits(n);
    const u64 hi = bits(n - 32);
    return (hi << 32) | bits(32);
  }
  __attribute__((always_inline)) bool bit() { return bits(1) != 0; }
  // the next eight bits, zeros past the end of the data
  __attribute__((always_inline)) unsigned peek8() {
    if (avail_ < 8) refill();
    return (unsigned)(cache_ >> 56);
  }
  // A flexbits value of `n` bits (n <= 15) and, when `may_sign` and the
  // value is not zero, the sign bit after it; false (nothing read) when
  // the packet holds too few bits for both, where a caller reads them one
  // at a time.
  __attribute__((always_inline)) bool magnitude_sign(unsigned n, bool may_sign, u32 &mag, bool &negative) {
    if ((int)n + 1 > avail_) {
      refill();
      if ((int)n + 1 > avail_) return false;
    }
    mag = n ? (u32)(cache_ >> (64 - n)) : 0u;
    const bool take = may_sign && mag != 0;
    negative = take && ((cache_ << n) >> 63) != 0;
    const unsigned used = n + (take ? 1u : 0u);
    cache_ <<= used;
    avail_ -= (int)used;
    return true;
  }
  // n bits that peek8 has shown and remaining() covers
  __attribute__((always_inline)) void skip(unsigned n) {
    cache_ <<= n;
    avail_ -= (int)n;
  }
  void align() {
    const unsigned r = (unsigned)(position() & 7);
    if (r) bits(8 - r);
  }
  void align_zero(const char *what) {
    while (position() & 7)
      if (bit()) syntax_error(std::string("non-zero alignment bit in ") + what);
  }

 private:
  void refill() {
    if (next_ + 8 <= nbytes_) {
      u64 w;
      std::memcpy(&w, p_ + next_, 8);
      cache_ |= __builtin_bswap64(w) >> avail_;
      const int take = (63 - avail_) >> 3;
      next_ += (size_t)take;
      avail_ += take * 8;
    } else {
      while (avail_ <= 56 && next_ < nbytes_) {
        cache_ |= (u64)p_[next_++] << (56 - avail_);
        avail_ += 8;
      }
    }
  }
  const u8 *p_;
  size_t nbytes_, next_ = 0;
  u64 cache_ = 0;
  int avail_ = 0;
};

class BitWriter {
 public:
  void put(u32 value, unsigned n) {  // n <= 32
    if (n == 0) return;
    acc_ = (acc_ << n) | (n == 32 ? value : (value & ((1u << n) - 1)));
    nacc_ += n;
    while (nacc_ >= 8) {
      out_.push_back((u8)(acc_ >> (nacc_ - 8)));
      nacc_ -= 8;
    }
    acc_ &= (nacc_ ? ((1ull << nacc_) - 1) : 0);
  }
  void put64(u64 value, unsigned n) {
    if (n > 32) {
      put((u32)(value >> 32), n - 32);
      put((u32)value, 32);
    } else {
      put((u32)value, n);
    }
  }
  void flag(bool b) { put(b ? 1 : 0, 1); }
  void align() {
    if (nacc_) put(0, 8 - nacc_);
  }
  size_t bit_length() const { return out_.size() * 8 + nacc_; }
  std::vector<u8> &bytes() {
    align();
    return out_;
  }

 private:
  std::vector<u8> out_;
  u64 acc_ = 0;
  unsigned nacc_ = 0;
};

// ---- prefix codes ---------------------------------------------------------
struct Code {
  u16 bits;
  u8 len;
  u8 sym;
};

// A prefix code resolved through one peek of eight bits, and indexed by
// symbol for writing.
struct Vlc {
  u16 lut[256] = {};
  u16 enc_bits[16] = {};
  u8 enc_len[16] = {};
  u8 max_len = 0;
  const char *name;
  Vlc(const char *n, std::initializer_list<Code> codes) : name(n) {
    for (const Code &c : codes) {
      const unsigned span = 1u << (8 - c.len), first = (unsigned)c.bits << (8 - c.len);
      for (unsigned s = first; s < first + span; s++) lut[s] = (u16)((c.len << 8) | c.sym);
      enc_bits[c.sym] = c.bits;
      enc_len[c.sym] = c.len;
      max_len = std::max(max_len, c.len);
    }
  }
};

__attribute__((always_inline)) inline u32 read_vlc(BitReader &r, const Vlc &t) {
  const u16 e = t.lut[r.peek8()];
  const unsigned len = e >> 8;
  if (e == 0 || len > r.remaining()) {
    if (r.remaining() < t.max_len)
      syntax_error(std::string("truncated ") + t.name + " at bit " + std::to_string(r.position()));
    syntax_error(std::string("invalid ") + t.name + " code at bit " + std::to_string(r.position()));
  }
  r.skip(len);
  return e & 0xff;
}

void write_vlc(BitWriter &w, const Vlc &t, u32 sym) {
  if (sym >= 16 || t.enc_len[sym] == 0) overflow(std::string("symbol outside ") + t.name);
  w.put(t.enc_bits[sym], t.enc_len[sym]);
}

// The normative code tables of T.832 (Tables 52 and 76-82, and the YUV and
// coded-block-pattern tables), as crates/jxr-native/src/entropy/vlc.rs and
// tile_decode/{yuv,cbphp}.rs list them: the codestream fixes every bit.
const Vlc kAbsLevel[2] = {
    {"ABS_LEVEL_INDEX", {{0b01, 2, 0}, {0b10, 2, 1}, {0b11, 2, 2}, {0b001, 3, 3}, {0b0001, 4, 4}, {0b00000, 5, 5}, {0b00001, 5, 6}}},
    {"ABS_LEVEL_INDEX", {{0b1, 1, 0}, {0b01, 2, 1}, {0b001, 3, 2}, {0b0001, 4, 3}, {0b00001, 5, 4}, {0b000000, 6, 5}, {0b000001, 6, 6}}}};
const Vlc kRunValue[3] = {{"RUN_VALUE", {{0b1, 1, 1}, {0b0, 1, 2}}},
                          {"RUN_VALUE", {{0b1, 1, 1}, {0b01, 2, 2}, {0b00, 2, 3}}},
                          {"RUN_VALUE", {{0b1, 1, 1}, {0b01, 2, 2}, {0b001, 3, 3}, {0b000, 3, 4}}}};
const Vlc kRunIndex{"RUN_INDEX", {{0b1, 1, 0}, {0b01, 2, 1}, {0b001, 3, 2}, {0b0000, 4, 3}, {0b0001, 4, 4}}};
const Vlc kIndexA[4] = {
    {"INDEX_A", {{0b1, 1, 0}, {0b00000, 5, 1}, {0b001, 3, 2}, {0b00001, 5, 3}, {0b01, 2, 4}, {0b0001, 4, 5}}},
    {"INDEX_A", {{0b01, 2, 0}, {0b0000, 4, 1}, {0b10, 2, 2}, {0b0001, 4, 3}, {0b11, 2, 4}, {0b001, 3, 5}}},
    {"INDEX_A", {{0b0000, 4, 0}, {0b0001, 4, 1}, {0b01, 2, 2}, {0b10, 2, 3}, {0b11, 2, 4}, {0b001, 3, 5}}},
    {"INDEX_A", {{0b00000, 5, 0}, {0b00001, 5, 1}, {0b01, 2, 2}, {0b1, 1, 3}, {0b0001, 4, 4}, {0b001, 3, 5}}}};
const Vlc kIndexB{"INDEX_B", {{0b0, 1, 0}, {0b10, 2, 2}, {0b110, 3, 1}, {0b111, 3, 3}}};
const Vlc kFirstIndex[5] = {
    {"FIRST_INDEX", {{0b00001, 5, 0}, {0b000001, 6, 1}, {0b0000000, 7, 2}, {0b0000001, 7, 3}, {0b00100, 5, 4}, {0b010, 3, 5},
                     {0b00101, 5, 6}, {0b1, 1, 7}, {0b00110, 5, 8}, {0b0001, 4, 9}, {0b00111, 5, 10}, {0b011, 3, 11}}},
    {"FIRST_INDEX", {{0b0010, 4, 0}, {0b00010, 5, 1}, {0b000000, 6, 2}, {0b000001, 6, 3}, {0b0011, 4, 4}, {0b010, 3, 5},
                     {0b00011, 5, 6}, {0b11, 2, 7}, {0b011, 3, 8}, {0b100, 3, 9}, {0b00001, 5, 10}, {0b101, 3, 11}}},
    {"FIRST_INDEX", {{0b11, 2, 0}, {0b001, 3, 1}, {0b0000000, 7, 2}, {0b0000001, 7, 3}, {0b00001, 5, 4}, {0b010, 3, 5},
                     {0b0000010, 7, 6}, {0b011, 3, 7}, {0b100, 3, 8}, {0b101, 3, 9}, {0b0000011, 7, 10}, {0b0001, 4, 11}}},
    {"FIRST_INDEX", {{0b001, 3, 0}, {0b11, 2, 1}, {0b0000000, 7, 2}, {0b00001, 5, 3}, {0b00010, 5, 4}, {0b010, 3, 5},
                     {0b0000001, 7, 6}, {0b011, 3, 7}, {0b00011, 5, 8}, {0b100, 3, 9}, {0b000001, 6, 10}, {0b101, 3, 11}}},
    {"FIRST_INDEX", {{0b010, 3, 0}, {0b1, 1, 1}, {0b0000001, 7, 2}, {0b0001, 4, 3}, {0b0000010, 7, 4}, {0b011, 3, 5},
                     {0b00000000, 8, 6}, {0b0010, 4, 7}, {0b0000011, 7, 8}, {0b0011, 4, 9}, {0b00000001, 8, 10}, {0b00001, 5, 11}}}};
const Vlc kDcPattern{"VAL_DC_YUV", {{0b10, 2, 0}, {0b001, 3, 1}, {0b00001, 5, 2}, {0b0001, 4, 3}, {0b11, 2, 4}, {0b010, 3, 5}, {0b00000, 5, 6}, {0b011, 3, 7}}};
const Vlc kLpPattern{"CBPLP_YUV1", {{0b0, 1, 0}, {0b100, 3, 1}, {0b1010, 4, 2}, {0b1011, 4, 3}, {0b1100, 4, 4}, {0b1101, 4, 5}, {0b1110, 4, 6}, {0b1111, 4, 7}}};
const Vlc kLpPatternSub{"CBPLP_YUV1", {{0b0, 1, 0}, {0b10, 2, 1}, {0b110, 3, 2}, {0b111, 3, 3}}};
const Vlc kNumCbphp[2] = {{"NUM_CBPHP", {{0b1, 1, 0}, {0b01, 2, 1}, {0b001, 3, 2}, {0b0000, 4, 3}, {0b0001, 4, 4}}},
                          {"NUM_CBPHP", {{0b1, 1, 0}, {0b000, 3, 1}, {0b001, 3, 2}, {0b010, 3, 3}, {0b011, 3, 4}}}};
const Vlc kNumCbphpYuv[2] = {
    {"NUM_BLKCBPHP_YUV", {{0b010, 3, 0}, {0b00000, 5, 1}, {0b0010, 4, 2}, {0b00001, 5, 3}, {0b00010, 5, 4}, {0b1, 1, 5},
                          {0b011, 3, 6}, {0b00011, 5, 7}, {0b0011, 4, 8}}},
    {"NUM_BLKCBPHP_YUV", {{0b1, 1, 0}, {0b001, 3, 1}, {0b010, 3, 2}, {0b0001, 4, 3}, {0b000001, 6, 4}, {0b011, 3, 5},
                          {0b00001, 5, 6}, {0b0000000, 7, 7}, {0b0000001, 7, 8}}}};
const Vlc kChroma{"CHR_CBPHP", {{0b1, 1, 0}, {0b01, 2, 1}, {0b00, 2, 2}}};
const Vlc kChromaBlocks{"NUM_CH_BLK", {{0b1, 1, 0}, {0b01, 2, 1}, {0b000, 3, 2}, {0b001, 3, 3}}};
const Vlc kRefTwo{"REF_CBPHP1", {{0b00, 2, 3}, {0b01, 2, 5}, {0b100, 3, 6}, {0b101, 3, 9}, {0b110, 3, 10}, {0b111, 3, 12}}};

// Remaps, refinement lengths and adaptation deltas of the same tables.
const i32 kAbsRemap[6] = {2, 3, 4, 6, 10, 14};
const u8 kAbsFixed[6] = {0, 0, 1, 2, 2, 2};
const u8 kRunRemap[15] = {1, 2, 3, 5, 7, 1, 2, 3, 5, 7, 1, 2, 3, 4, 5};
const int8_t kRunBin[15] = {-1, -1, -1, -1, 2, 2, 2, 1, 1, 1, 1, 0, 0, 0, 0};
const u8 kRunFixed[15] = {0, 0, 1, 1, 3, 0, 0, 1, 1, 2, 0, 0, 0, 0, 1};
const int8_t kAbsLevelDelta[7] = {1, 0, -1, -1, -1, -1, -1};
const int8_t kFirstIndexDelta[4][12] = {{1, 1, 1, 1, 1, 0, 0, -1, 2, 1, 0, 0},
                                        {2, 2, -1, -1, -1, 0, -2, -1, 0, 0, -2, -1},
                                        {-1, 1, 0, 2, 0, 0, 0, 0, -2, 0, 1, 1},
                                        {0, 1, 0, 1, -2, 0, -1, -1, -2, -1, -2, -2}};
const int8_t kIndexDelta[3][6] = {{-1, 1, 1, 1, 0, 1}, {-2, 0, 0, 2, 0, 0}, {-1, -1, 0, 1, -2, 0}};
const int8_t kCbphpDelta[5] = {0, -1, 0, 1, 1};
const int8_t kCbphpDeltaYuv[9] = {2, 2, 1, 1, -1, -2, -2, -2, -3};
const u16 kBlockOutput[16] = {0, 15, 3, 12, 1, 2, 4, 8, 5, 6, 9, 10, 7, 11, 13, 14};
const u8 kBlockOffset[6] = {0, 4, 2, 8, 12, 1};
const u8 kBlockBits[6] = {0, 2, 1, 2, 2, 0};
const int kGroupShift422[4] = {0, 1, 4, 5};
// Adaptive-scan start orders and totals (T.832 Tables 107-114).
const u8 kScanOrder0[16] = {0, 4, 1, 5, 8, 2, 9, 6, 12, 3, 10, 13, 7, 14, 11, 15};
const u8 kScanOrder1[16] = {0, 1, 2, 5, 4, 3, 6, 9, 8, 7, 12, 15, 13, 10, 11, 14};
const i32 kScanTotals[16] = {0, 32, 30, 28, 26, 24, 22, 20, 18, 16, 14, 12, 10, 8, 6, 4};
// Refinement order of a block's coefficients: down each column in turn.
const int kTranspose[16] = {0, 4, 8, 12, 1, 5, 9, 13, 2, 6, 10, 14, 3, 7, 11, 15};
const int kHierarchical[16] = {0, 1, 4, 5, 2, 3, 6, 7, 8, 9, 12, 13, 10, 11, 14, 15};
const int kRaster[8] = {0, 1, 2, 3, 4, 5, 6, 7};
// Coefficient-normalisation weights (T.832 Tables 115 and 116).
const i32 kLumaWeight[3] = {240, 12, 1};
const i32 kOtherWeight[3][16] = {{0, 240, 120, 80, 60, 48, 40, 34, 30, 27, 24, 22, 20, 18, 17, 16},
                                 {0, 12, 6, 4, 3, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1},
                                 {0, 16, 8, 5, 4, 3, 3, 2, 2, 2, 2, 1, 1, 1, 1, 1}};
const i32 kSubsampledWeight[2][3] = {{120, 37, 2}, {120, 18, 1}};

// ---- adaptive code-table selection ----------------------------------------
// Each decoded symbol moves a discriminant by its delta; at each adaptation
// point a discriminant past eight moves the selection one table along
// (entropy/adaptive.rs).
struct Adaptive {
  i32 d[2] = {0, 0};
  int table = 0;
  int di[2] = {0, 0};
  static Adaptive two() { return Adaptive{}; }
  static Adaptive several() {
    Adaptive a;
    a.table = 1;
    a.di[1] = 1;
    return a;
  }
  void observe_abs(u32 sym) { d[0] += kAbsLevelDelta[sym]; }
  void observe_one(const int8_t *delta, u32 sym) { d[0] += delta[sym]; }
  void observe_first(u32 sym) {
    d[0] += kFirstIndexDelta[di[0]][sym];
    d[1] += kFirstIndexDelta[di[1]][sym];
  }
  void observe_index(u32 sym) {
    d[0] += kIndexDelta[di[0]][sym];
    d[1] += kIndexDelta[di[1]][sym];
  }
  void adapt_two_tables() {
    if (d[0] < -8 && table != 0) {
      table--;
      d[0] = 0;
    } else if (d[0] > 8 && table != 1) {
      table++;
      d[0] = 0;
    } else {
      d[0] = std::clamp(d[0], -64, 64);
    }
  }
  void adapt_several(int max_table) {
    bool changed = false;
    if (d[0] < -8 && table != 0) {
      table--;
      changed = true;
    } else if (d[1] > 8 && table != max_table) {
      table++;
      changed = true;
    }
    if (changed) {
      d[0] = d[1] = 0;
      if (table == max_table)
        di[0] = di[1] = table - 1;
      else if (table == 0)
        di[0] = di[1] = 0;
      else {
        di[0] = table - 1;
        di[1] = table;
      }
    } else {
      d[0] = std::clamp(d[0], -64, 64);
      d[1] = std::clamp(d[1], -64, 64);
    }
  }
};

struct DcVlc {
  Adaptive abs[2];
  void adapt() {
    abs[0].adapt_two_tables();
    abs[1].adapt_two_tables();
  }
};

struct AcVlc {
  Adaptive first[2] = {Adaptive::several(), Adaptive::several()};
  Adaptive index[2][2] = {{Adaptive::several(), Adaptive::several()}, {Adaptive::several(), Adaptive::several()}};
  Adaptive abs[2];
  void adapt() {
    first[0].adapt_several(4);
    first[1].adapt_several(4);
    for (auto &cls : index)
      for (auto &a : cls) a.adapt_several(3);
    abs[0].adapt_two_tables();
    abs[1].adapt_two_tables();
  }
};

// ---- adaptive scan -------------------------------------------------------
struct Scan {
  u8 order[16];
  i32 totals[16];
  explicit Scan(const u8 *start) {
    std::memcpy(order, start, 16);
    std::memcpy(totals, kScanTotals, sizeof totals);
  }
  void reset_totals() { std::memcpy(totals, kScanTotals, sizeof totals); }
  // a coefficient found at scan index i: its count may overtake the index
  // before it, which then swaps places
  void count(int i) {
    i &= 15;
    totals[i]++;
    if (i > 1 && totals[i] > totals[i - 1]) {
      std::swap(totals[i], totals[i - 1]);
      std::swap(order[i], order[i - 1]);
    }
  }
};

// ---- coefficient-normalisation model (entropy/model.rs) -------------------
enum Band { kDc = 0, kLp = 1, kHp = 2 };

enum Kind { kLuma, kYuv420, kYuv422, kYuv444, kYuvK, kNComponent };

struct Model {
  int band;
  i32 state[2] = {0, 0};
  u8 bits[2];
  explicit Model(int b) : band(b) {
    const u8 initial = b == kDc ? 8 : b == kLp ? 4 : 0;
    bits[0] = bits[1] = initial;
  }
  static void update_one(i32 &state, u8 &bits, i32 mean) {
    i32 delta = (mean - 70) >> 2;
    if (delta <= -8) {
      delta = std::max(delta + 4, -16);
      state += delta;
      if (state < -8) {
        if (bits == 0) {
          state = -8;
        } else {
          state = 0;
          bits--;
        }
      }
    } else if (delta >= 8) {
      delta = std::min(delta - 4, 15);
      state += delta;
      if (state > 8) {
        if (bits >= 15) {
          bits = 15;
          state = 8;
        } else {
          state = 0;
          bits++;
        }
      }
    }
  }
  void update(i32 lap0, i32 lap1, Kind kind, int comps) {
    const i64 l0 = (i64)lap0 * kLumaWeight[band];
    if (l0 > INT32_MAX || l0 < INT32_MIN) overflow("coefficient model");
    int count = 1;
    i64 l1 = lap1;
    if (kind == kYuv420) {
      l1 *= kSubsampledWeight[0][band];
      count = 2;
    } else if (kind == kYuv422) {
      l1 *= kSubsampledWeight[1][band];
      count = 2;
    } else if (kind != kLuma) {
      if (comps < 2 || comps > 16) unsupported("coefficient-model component count");
      l1 *= kOtherWeight[band][comps - 1];
      if (l1 > INT32_MAX || l1 < INT32_MIN) overflow("coefficient model");
      if (band == kHp) l1 = (i32)l1 >> 4;
      count = 2;
    }
    if (l1 > INT32_MAX || l1 < INT32_MIN) overflow("coefficient model");
    update_one(state[0], bits[0], (i32)l0);
    if (count == 2) update_one(state[1], bits[1], (i32)l1);
  }
};

// ---- what sets one internal colour format's macroblock syntax apart -------
struct Syntax {
  Kind kind = kLuma;
  int comps = 1;
  bool joint() const { return kind == kYuv420 || kind == kYuv422 || kind == kYuv444; }
  bool joint_chroma() const { return kind == kYuv420 || kind == kYuv422; }
  int cols(int c) const { return c > 0 && joint_chroma() ? 2 : 4; }
  int rows(int c) const { return c > 0 && kind == kYuv420 ? 2 : 4; }
  int blocks(int c) const { return cols(c) * rows(c); }
  // component c's DC codes with the chroma tables only in a YUV plane
  int dc_class(int c) const { return joint() && c != 0 ? 1 : 0; }
  const int *block_order(int c, int &count) const {
    count = blocks(c);
    return count <= 8 ? kRaster : kHierarchical;
  }
};

bool syntax_of(u32 internal_code, u32 components, Syntax &s) {
  switch (internal_code) {
    case 0: s.kind = kLuma; s.comps = 1; return true;
    case 1: s.kind = kYuv420; s.comps = 3; return true;
    case 2: s.kind = kYuv422; s.comps = 3; return true;
    case 3: s.kind = kYuv444; s.comps = 3; return true;
    case 4: s.kind = kYuvK; s.comps = 4; return true;
    case 6:
      if (components < 2 || components > 16) return false;
      s.kind = kNComponent;
      s.comps = (int)components;
      return true;
    default: return false;
  }
}

// ---- coefficient syntax ---------------------------------------------------
i32 read_abs_level(BitReader &r, Adaptive &a) {
  const u32 idx = read_vlc(r, kAbsLevel[a.table]);
  a.observe_abs(idx);
  if (idx < 6) return kAbsRemap[idx] + (i32)r.bits(kAbsFixed[idx]);
  u32 fixed = r.bits(4) + 4;
  if (fixed == 19) {
    fixed += r.bits(2);
    if (fixed == 22) fixed += r.bits(3);
  }
  return (i32)((1u << fixed) + 2 + r.bits(fixed));
}

void write_abs_level(BitWriter &w, Adaptive &a, u32 v) {
  // v >= 2
  u32 idx;
  if (v < 18) {
    idx = v == 2 ? 0 : v == 3 ? 1 : v < 6 ? 2 : v < 10 ? 3 : v < 14 ? 4 : 5;
  } else {
    idx = 6;
  }
  write_vlc(w, kAbsLevel[a.table], idx);
  a.observe_abs(idx);
  if (idx < 6) {
    w.put(v - (u32)kAbsRemap[idx], kAbsFixed[idx]);
    return;
  }
  const u32 rest = v - 2;
  u32 fixed = 31 - __builtin_clz(rest);  // rest >= 16
  if (fixed > 29) overflow("coefficient level beyond the escape code");
  if (fixed < 19) {
    w.put(fixed - 4, 4);
  } else {
    w.put(15, 4);
    if (fixed < 22) {
      w.put(fixed - 19, 2);
    } else {
      w.put(3, 2);
      w.put(fixed - 22, 3);
    }
  }
  w.put(rest - (1u << fixed), fixed);
}

// (run << 3) | bits of every eight-bit prefix, for each maximum run 1..14:
// a run code and its refinement take at most seven bits, so one peek
// resolves a run whenever eight bits remain.
struct RunTable {
  u8 entry[15][256];
  RunTable() {
    std::memset(entry, 0, sizeof entry);
    for (int max_run = 1; max_run < 15; max_run++)
      for (unsigned prefix = 0; prefix < 256; prefix++) {
        unsigned run, len;
        if (max_run == 1) {
          run = 1;
          len = 0;
        } else if (max_run < 5) {
          const u16 e = kRunValue[max_run - 2].lut[prefix];
          run = e & 0xff;
          len = e >> 8;
        } else {
          const u16 e = kRunIndex.lut[prefix];
          const unsigned clen = e >> 8;
          const int t = (int)(e & 0xff) + 5 * kRunBin[max_run];
          const unsigned fixed = kRunFixed[t];
          const unsigned refinement = fixed ? ((prefix << clen) & 0xff) >> (8 - fixed) : 0;
          run = kRunRemap[t] + refinement;
          len = clen + fixed;
        }
        entry[max_run][prefix] = (u8)((run << 3) | len);
      }
  }
};
const RunTable kRunTable;

u32 read_run_slow(BitReader &r, int max_run);

__attribute__((always_inline)) inline u32 read_run(BitReader &r, int max_run) {
  if (max_run >= 1 && max_run <= 14) {
    const u8 e = kRunTable.entry[max_run][r.peek8()];
    if ((size_t)(e & 7) <= r.remaining()) {
      r.skip(e & 7);
      return e >> 3;
    }
  }
  return read_run_slow(r, max_run);
}

u32 read_run_slow(BitReader &r, int max_run) {
  if (max_run <= 0 || max_run > 14) syntax_error("coefficient run with no room");
  if (max_run == 1) return 1;
  if (max_run < 5) return read_vlc(r, kRunValue[max_run - 2]);
  const u32 s = read_vlc(r, kRunIndex);
  const int t = (int)s + 5 * kRunBin[max_run];
  return kRunRemap[t] + r.bits(kRunFixed[t]);
}

void write_run(BitWriter &w, u32 run, int max_run) {
  if (max_run == 1) return;
  if (max_run < 5) {
    write_vlc(w, kRunValue[max_run - 2], run);
    return;
  }
  for (u32 s = 0; s < 5; s++) {
    const int t = (int)s + 5 * kRunBin[max_run];
    if (run >= kRunRemap[t] && run < kRunRemap[t] + (1u << kRunFixed[t])) {
      write_vlc(w, kRunIndex, s);
      w.put(run - kRunRemap[t], kRunFixed[t]);
      return;
    }
  }
  overflow("coefficient run outside its table");
}

i32 signed_level(i32 mag, bool negative) {
  if (negative && mag == INT32_MIN) overflow("coefficient sign");
  return negative ? -mag : mag;
}

// The run/level syntax of one LP or HP block (T.832 Table 53 and its
// FIRST_INDEX/INDEX_A/INDEX_B symbols, entropy/coefficients.rs): calls
// place(scan_index, level) for each non-zero coefficient in coded order and
// returns how many there were.
template <class Place>
int read_block(BitReader &r, AcVlc &st, int cls, int start, Place place) {
  if (start < 1 || start > 15) syntax_error("block start location");
  const u32 first = read_vlc(r, kFirstIndex[st.first[cls].table]);
  st.first[cls].observe_first(first);
  bool negative = r.bit();
  u32 successors = first >> 2;
  bool context = (first & 1) && (successors & 1);
  int location = start;
  i32 mag = (first & 2) ? read_abs_level(r, st.abs[context]) : 1;
  u32 run = (first & 1) ? 0 : read_run(r, 15 - location);
  location += (int)run + 1;
  if (location > 16) syntax_error("coefficient location past the block");
  place(location - 1, signed_level(mag, negative));
  int count = 1;
  while (successors) {
    run = (successors & 1) ? 0 : read_run(r, std::max(0, 15 - location));
    location += (int)run + 1;
    if (location > 16) syntax_error("coefficient location past the block");
    u32 index;
    if (location < 15) {
      Adaptive &a = st.index[cls][context];
      index = read_vlc(r, kIndexA[a.table]);
      a.observe_index(index);
    } else if (location == 15) {
      index = read_vlc(r, kIndexB);
    } else {
      index = r.bit() ? 1 : 0;
    }
    successors = index >> 1;
    context = context && (successors & 1);
    negative = r.bit();
    mag = (index & 1) ? read_abs_level(r, st.abs[context]) : 1;
    place(location - 1, signed_level(mag, negative));
    count++;
  }
  return count;
}

// The writer's side: `at` and `level` list a block's non-zero coefficients
// in increasing scan index, all within start..15.
void write_block(BitWriter &w, AcVlc &st, int cls, int start, const int *at, const i32 *level, int n) {
  int location = start;
  auto successors_after = [&](int k) -> u32 { return k + 1 == n ? 0u : at[k + 1] == at[k] + 1 ? 1u : 2u; };
  u32 run = (u32)(at[0] - location);
  u32 successors = successors_after(0);
  u32 mag = magnitude(level[0]);
  const u32 first = (successors << 2) | (mag > 1 ? 2u : 0u) | (run == 0 ? 1u : 0u);
  write_vlc(w, kFirstIndex[st.first[cls].table], first);
  st.first[cls].observe_first(first);
  w.flag(level[0] < 0);
  bool context = run == 0 && (successors & 1);
  if (mag > 1) write_abs_level(w, st.abs[context], mag);
  if (run) write_run(w, run, 15 - location);
  location = at[0] + 1;
  for (int k = 1; k < n; k++) {
    run = (u32)(at[k] - location);
    if (run) write_run(w, run, 15 - location);
    location = at[k] + 1;
    successors = successors_after(k);
    mag = magnitude(level[k]);
    const u32 index = (successors << 1) | (mag > 1 ? 1u : 0u);
    if (location < 15) {
      Adaptive &a = st.index[cls][context];
      write_vlc(w, kIndexA[a.table], index);
      a.observe_index(index);
    } else if (location == 15) {
      write_vlc(w, kIndexB, index);
    } else {
      w.flag(mag > 1);
    }
    context = context && (successors & 1);
    w.flag(level[k] < 0);
    if (mag > 1) write_abs_level(w, st.abs[context], mag);
  }
}

// A flexbits refinement of a coefficient whose coded part is `coded`: its
// low `bits`, with a sign of its own only when the coded part is zero.
__attribute__((always_inline)) inline i32 read_flex(BitReader &r, i32 coded, unsigned bits) {
  const i32 value = (i32)r.bits(bits);
  if (coded > 0) return value;
  if (coded < 0) return -value;
  if (value == 0) return 0;
  return r.bit() ? -value : value;
}

void write_flex(BitWriter &w, i32 coded, u32 value, unsigned bits, bool negative) {
  w.put(value, bits);
  if (coded == 0 && value != 0) w.flag(negative);
}

// ---- coded-block patterns --------------------------------------------------
// CBPLP of a YUV plane (tile_decode/yuv.rs CbplpState).
struct Cbplp {
  bool full = true;  // 4:4:4: three components' bits; else luma + joint chroma
  int zero = 1, maxm = 1;
  u32 top() const { return full ? 7u : 3u; }
  void observe(u32 p) {
    zero = std::clamp(zero + 1 - (p == 0 ? 4 : 0), -8, 7);
    maxm = std::clamp(maxm + 1 - (p == top() ? 4 : 0), -8, 7);
  }
  u32 read(BitReader &r) {
    u32 p;
    if (zero <= 0 || maxm < 0) {
      p = read_vlc(r, full ? kLpPattern : kLpPatternSub);
      if (maxm < zero) p = top() - p;
    } else {
      p = r.bits(full ? 3 : 2);
    }
    observe(p);
    return p;
  }
  void write(BitWriter &w, u32 p) {
    if (zero <= 0 || maxm < 0) {
      write_vlc(w, full ? kLpPattern : kLpPatternSub, maxm < zero ? top() - p : p);
    } else {
      w.put(p, full ? 3 : 2);
    }
    observe(p);
  }
};

// Prediction of a component's coded-block pattern from its neighbours and
// its class's recent density (tile_decode/cbphp.rs PredictionState).
struct PatternPredictor {
  int state = 0, ones = -4, zeroes = 4;
  static u16 seed(int left, int top, int lb, int tb) {
    if (left < 0 && top < 0) return 1;
    if (left < 0) return (u16)((top >> tb) & 1);
    return (u16)((left >> lb) & 1);
  }
  void update(int count) {
    ones = std::clamp(ones + count - 3, -16, 15);
    zeroes = std::clamp(zeroes + 16 - count - 3, -16, 15);
    state = ones < 0 ? (ones < zeroes ? 1 : 2) : zeroes < 0 ? 2 : 0;
  }
  // 0: 4:4:4 (sixteen blocks); 1: 4:2:2 chroma (eight); 2: 4:2:0 chroma (four)
  u16 predict(int form, u16 residual, int left, int top) {
    u16 a = residual;
    if (form == 0) {
      if (state == 0) {
        a ^= seed(left, top, 5, 10);
        a ^= 0x0002 & (a << 1);
        a ^= 0x0010 & (a << 3);
        a ^= 0x0020 & (a << 1);
        a ^= (a & 0x0033) << 2;
        a ^= (a & 0x00cc) << 6;
        a ^= (a & 0x3300) << 2;
      } else if (state == 2) {
        a ^= 0xffff;
      }
      update(__builtin_popcount(a));
    } else if (form == 1) {
      if (state == 0) {
        a ^= seed(left, top, 1, 6);
        a ^= (a & 0x01) << 1;
        a ^= (a & 0x03) << 2;
        a ^= (a & 0x0c) << 2;
        a ^= (a & 0x30) << 2;
      } else if (state == 2) {
        a ^= 0x00ff;
      }
      update(std::min(16, __builtin_popcount(a) * 2));
    } else {
      if (state == 0) {
        a ^= seed(left, top, 1, 2);
        a ^= 0x02 & (a << 1);
        a ^= (a & 0x03) << 2;
      } else if (state == 2) {
        a ^= 0x000f;
      }
      update(std::min(16, __builtin_popcount(a) * 4));
    }
    return a;
  }
  u16 residual(int form, u16 actual, int left, int top) {
    u16 r = actual;
    if (form == 0) {
      if (state == 0) {
        r ^= (r & 0x3300) << 2;
        r ^= (r & 0x00cc) << 6;
        r ^= (r & 0x0033) << 2;
        r ^= 0x0020 & (r << 1);
        r ^= 0x0010 & (r << 3);
        r ^= 0x0002 & (r << 1);
        r ^= seed(left, top, 5, 10);
      } else if (state == 2) {
        r ^= 0xffff;
      }
      update(__builtin_popcount(actual));
    } else if (form == 1) {
      if (state == 0) {
        r ^= (r & 0x30) << 2;
        r ^= (r & 0x0c) << 2;
        r ^= (r & 0x03) << 2;
        r ^= (r & 0x01) << 1;
        r ^= seed(left, top, 1, 6);
      } else if (state == 2) {
        r ^= 0x00ff;
      }
      update(std::min(16, __builtin_popcount(actual) * 2));
    } else {
      if (state == 0) {
        r ^= (r & 0x03) << 2;
        r ^= 0x02 & (r << 1);
        r ^= seed(left, top, 1, 2);
      } else if (state == 2) {
        r ^= 0x000f;
      }
      update(std::min(16, __builtin_popcount(actual) * 4));
    }
    return r;
  }
};

inline u16 spread422(u32 blocks) { return (u16)((blocks & 1) | ((blocks & 2) << 1)); }
inline u16 gather422(u16 p) { return (u16)((p & 1) | ((p >> 1) & 2)); }

u16 read_groups(BitReader &r, u32 count) {
  switch (count) {
    case 0: return 0;
    case 1: return (u16)(1u << r.bits(2));
    case 2: return (u16)read_vlc(r, kRefTwo);
    case 3: return (u16)(0x0f ^ (1u << r.bits(2)));
    case 4: return 0x0f;
    default: syntax_error("CBPHP group count");
  }
}

void write_groups(BitWriter &w, u32 groups) {
  switch (__builtin_popcount(groups)) {
    case 0:
    case 4: return;
    case 1: w.put((u32)__builtin_ctz(groups), 2); return;
    case 2: write_vlc(w, kRefTwo, groups); return;
    case 3: w.put((u32)__builtin_ctz(groups ^ 0x0f), 2); return;
  }
}

// (value, increment) of a non-zero nibble of blocks in the NUM_BLKCBPHP code
void block_value(u16 nibble, u32 &value, u32 &increment) {
  int code = 0;
  for (int c = 0; c < 16; c++)
    if (kBlockOutput[c] == (nibble & 0x0f)) code = c;
  value = code == 1 ? 5 : code <= 3 ? 2 : code <= 7 ? 1 : code <= 11 ? 3 : 4;
  increment = (u32)code - kBlockOffset[value];
}

// CBPHP of a macroblock (tile_decode/cbphp.rs CbphpState): per component for
// planes other than YUV, jointly for YUV.
struct Cbphp {
  Adaptive groups, blocks;
  PatternPredictor prediction[2];
  std::vector<std::vector<u16>> previous;  // [component][x] of the row above
  Cbphp(int width, int comps) : previous((size_t)comps, std::vector<u16>((size_t)width, 0)) {}
  void adapt() {
    groups.adapt_two_tables();
    blocks.adapt_two_tables();
  }
  int left(int c, int x) const { return x != 0 ? (int)previous[c][x - 1] : -1; }
  int top(int c, int x, int y) const { return y != 0 ? (int)previous[c][x] : -1; }

  u16 read_component(BitReader &r, int x, int y, int c) {
    const u32 count = read_vlc(r, kNumCbphp[groups.table]);
    groups.observe_one(kCbphpDelta, count);
    const u16 g = read_groups(r, count);
    u16 residual = 0;
    for (int group = 0; group < 4; group++) {
      if (!(g & (1 << group))) continue;
      const u32 nb = read_vlc(r, kNumCbphp[blocks.table]);
      blocks.observe_one(kCbphpDelta, nb);
      const u32 value = nb + 1;
      const u32 code = kBlockOffset[value] + r.bits(kBlockBits[value]);
      residual |= (u16)(kBlockOutput[code] << (group * 4));
    }
    const u16 actual = prediction[c != 0].predict(0, residual, left(c, x), top(c, x, y));
    previous[c][x] = actual;
    return actual;
  }

  void write_component(BitWriter &w, int x, int y, int c, u16 actual) {
    const u16 residual = prediction[c != 0].residual(0, actual, left(c, x), top(c, x, y));
    previous[c][x] = actual;
    u32 g = 0;
    for (int group = 0; group < 4; group++)
      if ((residual >> (group * 4)) & 0x0f) g |= 1u << group;
    const u32 count = (u32)__builtin_popcount(g);
    write_vlc(w, kNumCbphp[groups.table], count);
    groups.observe_one(kCbphpDelta, count);
    write_groups(w, g);
    for (int group = 0; group < 4; group++) {
      if (!(g & (1u << group))) continue;
      u32 value, increment;
      block_value((u16)((residual >> (group * 4)) & 0x0f), value, increment);
      write_vlc(w, kNumCbphp[blocks.table], value - 1);
      blocks.observe_one(kCbphpDelta, value - 1);
      w.put(increment, kBlockBits[value]);
    }
  }

  // sampling form of the chroma components: 0 = 4:4:4, 1 = 4:2:2, 2 = 4:2:0
  void read_yuv(BitReader &r, int x, int y, int form, u16 out[3]) {
    const u32 count = read_vlc(r, kNumCbphp[groups.table]);
    groups.observe_one(kCbphpDelta, count);
    const u16 g = read_groups(r, count);
    u16 residual[3] = {0, 0, 0};
    for (int group = 0; group < 4; group++) {
      if (!(g & (1 << group))) continue;
      const u32 nb = read_vlc(r, kNumCbphpYuv[blocks.table]);
      blocks.observe_one(kCbphpDeltaYuv, nb);
      u32 value = nb + 1, chroma = 0;
      if (value >= 6) {
        chroma = read_vlc(r, kChroma) + 1;
        if (value >= 9) value += read_vlc(r, kChroma);
        value -= 6;
      }
      const u32 code = kBlockOffset[value] + r.bits(kBlockBits[value]);
      residual[0] |= (u16)((kBlockOutput[code] & 0x0f) << (group * 4));
      for (int c = 0; c < 2; c++) {
        if (!(chroma & (1u << c))) continue;
        if (form == 0) {
          const u32 n = read_vlc(r, kChromaBlocks) + 1;
          residual[c + 1] |= (u16)(read_groups(r, n) << (group * 4));
        } else if (form == 1) {
          const u32 b = read_vlc(r, kChroma) + 1;
          residual[c + 1] |= (u16)(spread422(b) << kGroupShift422[group]);
        } else {
          residual[c + 1] |= (u16)(1u << group);
        }
      }
    }
    for (int c = 0; c < 3; c++) {
      const int f = c == 0 ? 0 : form;
      out[c] = prediction[c != 0].predict(f, residual[c], left(c, x), top(c, x, y));
      previous[c][x] = out[c];
    }
  }

  void write_yuv(BitWriter &w, int x, int y, int form, const u16 actual[3]) {
    u16 residual[3];
    for (int c = 0; c < 3; c++) {
      const int f = c == 0 ? 0 : form;
      residual[c] = prediction[c != 0].residual(f, actual[c], left(c, x), top(c, x, y));
      previous[c][x] = actual[c];
    }
    auto luma = [&](int group) { return (u16)((residual[0] >> (group * 4)) & 0x0f); };
    auto chroma = [&](int c, int group) -> u16 {
      if (form == 0) return (u16)((residual[c] >> (group * 4)) & 0x0f);
      if (form == 1) return gather422((u16)(residual[c] >> kGroupShift422[group]));
      return (u16)((residual[c] >> group) & 1);
    };
    u32 g = 0;
    for (int group = 0; group < 4; group++)
      if (luma(group) || chroma(1, group) || chroma(2, group)) g |= 1u << group;
    const u32 count = (u32)__builtin_popcount(g);
    write_vlc(w, kNumCbphp[groups.table], count);
    groups.observe_one(kCbphpDelta, count);
    write_groups(w, g);
    for (int group = 0; group < 4; group++) {
      if (!(g & (1u << group))) continue;
      const u16 l = luma(group);
      const u16 ch[2] = {chroma(1, group), chroma(2, group)};
      const u32 mask = (ch[0] ? 1u : 0u) | (ch[1] ? 2u : 0u);
      u32 value = 0, increment = 0;
      if (l) block_value(l, value, increment);
      u32 sym;
      int inc = -1;
      if (mask == 0) {
        sym = value - 1;
      } else if (value <= 2) {
        sym = value + 5;
      } else {
        sym = 8;
        inc = (int)value - 3;
      }
      write_vlc(w, kNumCbphpYuv[blocks.table], sym);
      blocks.observe_one(kCbphpDeltaYuv, sym);
      if (mask) write_vlc(w, kChroma, mask - 1);
      if (inc >= 0) write_vlc(w, kChroma, (u32)inc);
      w.put(increment, kBlockBits[value]);
      for (int c = 0; c < 2; c++) {
        if (!ch[c]) continue;
        if (form == 0) {
          write_vlc(w, kChromaBlocks, (u32)__builtin_popcount(ch[c]) - 1);
          write_groups(w, ch[c]);
        } else if (form == 1) {
          write_vlc(w, kChroma, (u32)ch[c] - 1);
        }
      }
    }
  }
};

// ---- quantizers ------------------------------------------------------------
// One QP set: a parameter per component (COMPONENT_MODE 0 uniform, 1 luma
// and chroma, 2 each; tile_decode/quantizer.rs read_quantizer_set).
std::vector<u8> read_qp_set(BitReader &r, int comps) {
  const u32 mode = comps == 1 ? 0 : r.bits(2);
  std::vector<u8> v((size_t)comps);
  if (mode == 0) {
    const u8 p = (u8)r.bits(8);
    std::fill(v.begin(), v.end(), p);
  } else if (mode == 1) {
    const u8 luma = (u8)r.bits(8), chroma = (u8)r.bits(8);
    std::fill(v.begin(), v.end(), chroma);
    v[0] = luma;
  } else if (mode == 2) {
    for (auto &p : v) p = (u8)r.bits(8);
  } else {
    unsupported("reserved QP component mode");
  }
  return v;
}

void write_qp_set(BitWriter &w, const u8 *v, int comps) {
  if (comps == 1) {
    w.put(v[0], 8);
    return;
  }
  bool chroma_uniform = true;
  for (int c = 2; c < comps; c++) chroma_uniform = chroma_uniform && v[c] == v[1];
  if (chroma_uniform && v[0] == v[1]) {
    w.put(0, 2);
    w.put(v[0], 8);
  } else if (chroma_uniform) {
    w.put(1, 2);
    w.put(v[0], 8);
    w.put(v[1], 8);
  } else {
    w.put(2, 2);
    for (int c = 0; c < comps; c++) w.put(v[c], 8);
  }
}

u32 qp_index_bits(size_t sets) { return sets <= 3 ? 1 : sets <= 5 ? 2 : sets <= 9 ? 3 : 4; }

u32 read_qp_index(BitReader &r, size_t sets) {
  if (sets == 0 || sets > 16) syntax_error("QP table length");
  if (!r.bit()) return 0;
  const u32 index = r.bits(qp_index_bits(sets)) + 1;
  if (index >= sets) syntax_error("QP index past its table");
  return index;
}

void write_qp_index(BitWriter &w, u32 index, size_t sets) {
  w.flag(index != 0);
  if (index) w.put(index - 1, qp_index_bits(sets));
}

// ---- headers ---------------------------------------------------------------
struct PlaneHeader {
  u32 internal = 0;  // INTERNAL_CLR_FMT
  bool scaled = false;
  u32 bands_code = 0;  // BANDS_PRESENT (0 all .. 3 DC only)
  int comps = 1;
  u32 cx = 0, cy = 0;  // chroma centering
  u32 shift_bits = 0, mantissa = 0;
  int exponent_bias = 0;
  bool has_dc = false, has_lp = false, has_hp = false;  // image-level QPs
  std::vector<u8> dc, lp, hp;
  Syntax syn;
};

struct ImageHeader {
  u32 width = 0, height = 0;
  bool hard = false, tiling = false, frequency = false, index = false, short_header = false, long_word = false,
       windowing = false, trim = false, rb_not_swapped = false, premultiplied = false, alpha_plane = false;
  u32 orientation = 0, overlap = 0, out_color = 0, out_depth = 0;
  std::vector<u32> tile_w, tile_h;  // explicit sizes (the last of each inferred)
  u32 margins[4] = {0, 0, 0, 0};     // top, left, bottom, right
};

struct Headers {
  ImageHeader image;
  PlaneHeader primary, alpha;
  size_t bytes_consumed = 0;
};

constexpr u64 kSignature = 0x574D50484F544F00ull;  // "WMPHOTO\0"

// band presence: BANDS_PRESENT code -> JXRHIP_BANDS_*
u32 bands_of_code(u32 code) { return 3 - code; }
bool has_lp(u32 bands) { return bands >= JXRHIP_BANDS_NO_HIGH_PASS; }
bool has_hp(u32 bands) { return bands >= JXRHIP_BANDS_NO_FLEXBITS; }
bool has_flex(u32 bands) { return bands == JXRHIP_BANDS_ALL; }

bool reserved_depth(u32 code) { return code == 5 || (code >= 11 && code <= 14); }
bool depth_has_shift(u32 code) { return code == 2 || code == 3 || code == 6; }

u32 inferred_margin(u32 d) { return d % 16 ? 16 - d % 16 : 0; }

void parse_plane_header(BitReader &r, u32 depth, PlaneHeader &p) {
  p.internal = r.bits(3);
  if (p.internal == 5 || p.internal == 7) syntax_error("reserved INTERNAL_CLR_FMT " + std::to_string(p.internal));
  p.scaled = r.bit();
  p.bands_code = r.bits(4);
  if (p.bands_code > 3) syntax_error("reserved BANDS_PRESENT " + std::to_string(p.bands_code));
  auto centering = [&]() {
    const u32 v = r.bits(3);
    return v == 5 || v == 6 ? 7u : v;
  };
  switch (p.internal) {
    case 6: {
      const u32 n = r.bits(4);
      if (n == 15) {
        p.comps = (int)r.bits(12) + 16;
      } else {
        r.bits(4);
        p.comps = (int)n + 1;
      }
      break;
    }
    case 1:
      r.bits(1);
      p.cx = centering();
      r.bits(1);
      p.cy = centering();
      p.comps = 3;
      break;
    case 2:
      r.bits(1);
      p.cx = centering();
      r.bits(4);
      p.comps = 3;
      break;
    case 3:
      r.bits(4);
      r.bits(4);
      p.comps = 3;
      break;
    case 4: p.comps = 4; break;
    default: p.comps = 1; break;
  }
  if (depth_has_shift(depth)) p.shift_bits = r.bits(8);
  if (depth == 7) {
    p.mantissa = r.bits(8);
    p.exponent_bias = (int8_t)r.bits(8);
  }
  p.has_dc = r.bit();
  if (p.has_dc) p.dc = read_qp_set(r, p.comps);
  const u32 bands = bands_of_code(p.bands_code);
  if (has_lp(bands)) {
    r.bits(1);
    p.has_lp = r.bit();
    if (p.has_lp) p.lp = read_qp_set(r, p.comps);
    if (has_hp(bands)) {
      r.bits(1);
      p.has_hp = r.bit();
      if (p.has_hp) p.hp = read_qp_set(r, p.comps);
    }
  }
  r.align_zero("the plane header");
  syntax_of(p.internal, (u32)p.comps, p.syn);
}

Headers parse_headers(const u8 *data, size_t size) {
  BitReader r(data, size);
  Headers h;
  ImageHeader &im = h.image;
  if (r.bits64(64) != kSignature) syntax_error("invalid GDI signature");
  if (r.bits(4) != 1) syntax_error("reserved RESERVED_B");
  im.hard = r.bit();
  r.bits(3);
  im.tiling = r.bit();
  im.frequency = r.bit();
  im.orientation = r.bits(3);
  im.index = r.bit();
  im.overlap = r.bits(2);
  if (im.overlap == 3) syntax_error("reserved OVERLAP_MODE 3");
  im.short_header = r.bit();
  im.long_word = r.bit();
  im.windowing = r.bit();
  im.trim = r.bit();
  r.bits(1);
  im.rb_not_swapped = r.bit();
  im.premultiplied = r.bit();
  im.alpha_plane = r.bit();
  im.out_color = r.bits(4);
  if (im.out_color > 8) syntax_error("reserved OUTPUT_CLR_FMT " + std::to_string(im.out_color));
  im.out_depth = r.bits(4);
  if (reserved_depth(im.out_depth)) syntax_error("reserved OUTPUT_BITDEPTH " + std::to_string(im.out_depth));
  const unsigned dim_bits = im.short_header ? 16 : 32;
  const u64 w1 = r.bits(dim_bits), h1 = r.bits(dim_bits);
  if (w1 == UINT32_MAX || h1 == UINT32_MAX) overflow("image dimension");
  im.width = (u32)w1 + 1;
  im.height = (u32)h1 + 1;
  if (im.tiling) {
    const u32 cols = r.bits(12), rows = r.bits(12);
    const unsigned tb = im.short_header ? 8 : 16;
    for (u32 i = 0; i < cols; i++) im.tile_w.push_back(r.bits(tb));
    for (u32 i = 0; i < rows; i++) im.tile_h.push_back(r.bits(tb));
  }
  if (im.windowing) {
    for (auto &m : im.margins) m = r.bits(6);
  } else {
    im.margins[2] = inferred_margin(im.height);
    im.margins[3] = inferred_margin(im.width);
  }
  parse_plane_header(r, im.out_depth, h.primary);
  if (im.alpha_plane) parse_plane_header(r, im.out_depth, h.alpha);
  h.bytes_consumed = (r.position() + 7) / 8;
  return h;
}

// Tile index and profile area between the headers and the packets
// (tail.rs CodestreamDirectory::parse).
struct Directory {
  std::vector<u64> offsets;
  u32 profile = 111, level = 255;
  bool has_profile = false;
  size_t data_offset = 0;
};

struct ByteCursor {
  const u8 *p;
  size_t size, pos;
  u32 byte() {
    if (pos >= size) syntax_error("truncated tile index or profile area");
    return p[pos++];
  }
  u32 u16be() {
    const u32 a = byte();
    return (a << 8) | byte();
  }
  u64 vlw() {
    const u32 first = byte();
    if (first < 0xfb) return (u64)first * 256 + byte();
    if (first == 0xfb) {
      u64 v = 0;
      for (int i = 0; i < 4; i++) v = (v << 8) | byte();
      return v;
    }
    if (first == 0xfc) {
      u64 v = 0;
      for (int i = 0; i < 8; i++) v = (v << 8) | byte();
      return v;
    }
    return 0;
  }
};

void write_vlw(std::vector<u8> &out, u64 v) {
  if (v < 0xfb00) {
    out.push_back((u8)(v >> 8));
    out.push_back((u8)v);
  } else if (v <= UINT32_MAX) {
    out.push_back(0xfb);
    for (int s = 24; s >= 0; s -= 8) out.push_back((u8)(v >> s));
  } else {
    out.push_back(0xfc);
    for (int s = 56; s >= 0; s -= 8) out.push_back((u8)(v >> s));
  }
}

size_t packets_per_tile(const Headers &h) {
  if (!h.image.frequency) return 1;
  return (size_t)(4 - h.primary.bands_code);  // DC, then LP, HP, flexbits as coded
}

Directory parse_directory(const u8 *data, size_t size, const Headers &h, size_t tiles) {
  Directory d;
  ByteCursor c{data, size, h.bytes_consumed};
  if (h.bytes_consumed > size) syntax_error("headers past the end of the codestream");
  const bool requires_index = h.image.frequency || h.image.tiling;
  if (requires_index && !h.image.index) syntax_error("tiled or frequency-mode codestream without a tile index");
  if (h.image.index) {
    if (c.u16be() != 1) syntax_error("INDEX_TABLE_STARTCODE");
    const size_t n = tiles * packets_per_tile(h);
    for (size_t i = 0; i < n; i++) d.offsets.push_back(c.vlw());
  }
  const u64 subsequent = c.vlw();
  if (subsequent) {
    if (subsequent < 4) syntax_error("SubsequentBytes below one profile record");
    if (subsequent > size - c.pos) syntax_error("profile area past the end of the codestream");
    const size_t end = c.pos + (size_t)subsequent;
    bool first = true;
    for (;;) {
      if (c.pos + 4 > end) syntax_error("truncated PROFILE_LEVEL_INFO");
      const u32 profile = c.byte(), level = c.byte(), tail = c.u16be();
      if (tail & ~1u) syntax_error("PROFILE_LEVEL_INFO reserved bits");
      if (first) {
        d.profile = profile;
        d.level = level;
        d.has_profile = true;
        first = false;
      }
      if (tail & 1) break;
    }
    c.pos = end;
  }
  d.data_offset = c.pos;
  return d;
}

// ---- Annex-A container (annex_a.rs) ---------------------------------------
struct AnnexA {
  u32 width = 0, height = 0;
  u8 guid[16] = {};
  size_t offset = 0, bytes = 0, alpha_offset = 0, alpha_bytes = 0;
  bool alpha = false;
  bool has_orientation = false;
  u32 orientation = 0;
};

u32 le16(const u8 *p) { return (u32)p[0] | ((u32)p[1] << 8); }
u32 le32(const u8 *p) { return (u32)p[0] | ((u32)p[1] << 8) | ((u32)p[2] << 16) | ((u32)p[3] << 24); }

AnnexA parse_annex_a(const u8 *b, size_t n) {
  if (n < 8) syntax_error("truncated Annex-A header");
  if (!(b[0] == 'I' && b[1] == 'I' && b[2] == 0xBC && b[3] == 1)) syntax_error("invalid signature");
  const size_t ifd = le32(b + 4);
  if (ifd & 1) syntax_error("FIRST_IFD_OFFSET alignment");
  if (ifd + 2 > n) syntax_error("IFD past the end of the file");
  const u32 count = le16(b + ifd);
  if (count == 0) syntax_error("empty Annex-A directory");
  if (ifd + 2 + (size_t)count * 12 > n) syntax_error("IFD entries past the end of the file");
  struct Entry {
    u32 tag, type, count, value;
  };
  std::vector<Entry> e;
  for (u32 i = 0; i < count; i++) {
    const u8 *q = b + ifd + 2 + (size_t)i * 12;
    Entry x{le16(q), le16(q + 2), le32(q + 4), le32(q + 8)};
    if (!e.empty() && x.tag <= e.back().tag) syntax_error("unsorted Annex-A tags");
    e.push_back(x);
  }
  auto find = [&](u32 tag) -> const Entry * {
    for (const Entry &x : e)
      if (x.tag == tag) return &x;
    return nullptr;
  };
  auto scalar = [&](const Entry *x) -> u32 {
    if (x->count != 1 || !(x->type == 1 || x->type == 3 || x->type == 4)) syntax_error("invalid Annex-A entry");
    return x->type == 1 ? (x->value & 0xff) : x->type == 3 ? (x->value & 0xffff) : x->value;
  };
  auto required = [&](u32 tag) {
    const Entry *x = find(tag);
    if (!x) syntax_error("missing Annex-A field " + std::to_string(tag));
    return x;
  };
  // odd codestream lengths that writers rounded up to even stay within the
  // input (annex_a.rs checked_codestream_range)
  auto range = [&](size_t off, size_t len) -> size_t {
    if (off > n) syntax_error("Annex-A range past the end of the file");
    if (len <= n - off) return len;
    if (len % 2 == 0 && len == n - off + 1) return n - off;
    syntax_error("Annex-A range past the end of the file");
  };
  AnnexA a;
  const Entry *pf = required(0xBC01);
  if (pf->type != 1 || pf->count != 16 || (size_t)pf->value + 16 > n) syntax_error("invalid PIXEL_FORMAT");
  std::memcpy(a.guid, b + pf->value, 16);
  a.width = scalar(required(0xBC80));
  a.height = scalar(required(0xBC81));
  a.offset = scalar(required(0xBCC0));
  size_t bytes = scalar(required(0xBCC1));
  if (bytes == 0) {
    if (a.offset > n) syntax_error("IMAGE_OFFSET past the end of the file");
    bytes = n - a.offset;
  }
  a.bytes = range(a.offset, bytes);
  const Entry *ao = find(0xBCC2), *ab = find(0xBCC3);
  if ((ao == nullptr) != (ab == nullptr)) syntax_error("alpha offset without its byte count");
  if (ao) {
    const size_t off = scalar(ao);
    size_t len = scalar(ab);
    if (!(off == 0 && len == 0)) {
      // some writers store the alpha end offset in ALPHA_BYTE_COUNT
      if (len == n && off > 0 && off < len) len -= off;
      a.alpha_bytes = range(off, len);
      a.alpha_offset = off;
      a.alpha = true;
      if (a.alpha_offset < a.offset + a.bytes && a.offset < a.alpha_offset + a.alpha_bytes)
        syntax_error("overlapping primary and alpha codestream ranges");
    }
  }
  if (const Entry *t = find(0xBC02)) {
    a.orientation = scalar(t);
    a.has_orientation = true;
  }
  return a;
}

// ---- one parsed input --------------------------------------------------------
struct Stream {
  const u8 *data = nullptr;  // into the owning jxrhip_codestream's bytes
  size_t size = 0;
  Headers h;
  Directory dir;
  std::vector<u32> cols, rows;  // complete tile grid, macroblocks
};

std::vector<u32> complete_tiles(const std::vector<u32> &explicit_sizes, u32 total) {
  std::vector<u32> out;
  u64 used = 0;
  for (u32 s : explicit_sizes) {
    if (s == 0) syntax_error("zero tile size");
    used += s;
    out.push_back(s);
  }
  if (used >= total) syntax_error("tile sizes past the image");
  out.push_back(total - (u32)used);
  return out;
}

Stream parse_stream(const u8 *data, size_t size) {
  Stream s;
  s.data = data;
  s.size = size;
  s.h = parse_headers(data, size);
  const ImageHeader &im = s.h.image;
  const u64 ew = (u64)im.width + im.margins[1] + im.margins[3];
  const u64 eh = (u64)im.height + im.margins[0] + im.margins[2];
  if (ew % 16 || eh % 16) syntax_error("extended image dimension not a multiple of 16");
  s.cols = complete_tiles(im.tile_w, (u32)(ew / 16));
  s.rows = complete_tiles(im.tile_h, (u32)(eh / 16));
  s.dir = parse_directory(data, size, s.h, s.cols.size() * s.rows.size());
  return s;
}

}  // namespace

struct jxrhip_codestream {
  std::vector<u8> bytes;
  Stream primary;
  bool has_separate = false;
  Stream alpha;
  bool annex = false;
  AnnexA container;
};

namespace {

// ---- decoding a tile --------------------------------------------------------
// Tile-local macroblock record: the DC and low-pass values (absolute, after
// prediction) and the LP quantizer index, which DC, LP and HP-direction
// prediction read.
struct Record {
  i32 low[16];
  u8 lp_index;
};

struct Slot {
  int dc_mode = JXRHIP_PRED_NONE;
  u8 lp_index = 0, hp_index = 0;
  u8 hp_bits[2] = {0, 0};
  int hp_dir = JXRHIP_PRED_NONE;
};

// The QP sets of one plane in one tile, and whether LP and HP inherit.
struct TileQuant {
  int comps = 1;
  std::vector<u8> dc, lp, hp;  // [set][component]
  bool lp_inherits = true, hp_inherits = true;
  size_t lp_sets() const { return lp.size() / (size_t)comps; }
  size_t hp_sets() const { return hp.size() / (size_t)comps; }
};

std::vector<u8> read_sets(BitReader &r, int comps, size_t count) {
  std::vector<u8> v;
  for (size_t i = 0; i < count; i++) {
    const std::vector<u8> s = read_qp_set(r, comps);
    v.insert(v.end(), s.begin(), s.end());
  }
  return v;
}

void read_dc_quant(BitReader &r, const PlaneHeader &p, TileQuant &q) {
  q.comps = p.comps;
  q.dc = p.has_dc ? p.dc : read_qp_set(r, p.comps);
  q.lp = q.hp = q.dc;
  q.lp_inherits = q.hp_inherits = true;
}
void read_lp_quant(BitReader &r, const PlaneHeader &p, TileQuant &q) {
  if (!has_lp(bands_of_code(p.bands_code))) {
    q.lp = q.dc;
    q.lp_inherits = true;
    return;
  }
  if (p.has_lp) {
    q.lp = p.lp;
    q.lp_inherits = false;
    return;
  }
  q.lp_inherits = r.bit();
  if (q.lp_inherits) {
    q.lp = q.dc;
  } else {
    q.lp = read_sets(r, p.comps, r.bits(4) + 1);
  }
}
void read_hp_quant(BitReader &r, const PlaneHeader &p, TileQuant &q) {
  if (!has_hp(bands_of_code(p.bands_code))) {
    q.hp = q.lp;
    q.hp_inherits = true;
    return;
  }
  if (p.has_hp) {
    q.hp = p.hp;
    q.hp_inherits = false;
    return;
  }
  q.hp_inherits = r.bit();
  if (q.hp_inherits) {
    q.hp = q.lp;
  } else {
    q.hp = read_sets(r, p.comps, r.bits(4) + 1);
  }
}

// Where one plane of one tile stores its macroblocks: per component, the
// first coefficient and the first macroblock record of the tile's window.
struct PlaneOut {
  jxrhip_arena *arena;  // pointers below are the decoded object's vectors
  int32_t *coef[16];
  uint32_t coef_base[16];  // arena index of coef[c][0]
  size_t mb_base[16];      // arena macroblock index of the tile's first
  int cpm[16];             // coefficients per macroblock
  uint32_t *offsets, *quantizers, *coded_x, *coded_y;
  uint8_t *bands, *preds;
};

// The reader (and, mirrored, the writer) of one coded plane of one tile.
struct PlaneCoder {
  const PlaneHeader *hdr;
  Syntax syn;
  u32 bands;  // the bands this decode reads (a reduced decode keeps fewer)
  int w, h;   // tile size in macroblocks
  TileQuant q;
  DcVlc dc_vlc;
  AcVlc lp_vlc, hp_vlc;
  Scan lp_scan{kScanOrder0}, hp_h{kScanOrder0}, hp_v{kScanOrder1};
  Model dc_model{kDc}, lp_model{kLp}, hp_model{kHp};
  Cbplp cbplp;
  Cbphp cbphp;
  std::vector<Record> rec;  // [component][macroblock]
  std::vector<Slot> slot;   // [macroblock]

  PlaneCoder(const PlaneHeader *p, u32 b, int width, int height)
      : hdr(p), syn(p->syn), bands(b), w(width), h(height), cbphp(width, p->syn.comps),
        rec((size_t)p->syn.comps * width * height), slot((size_t)width * height) {
    cbplp.full = syn.kind == kYuv444;
  }
  Record &R(int c, int m) { return rec[(size_t)c * w * h + m]; }
  static bool adapts_after(int x, int width) { return x + 1 == width || x % 16 == 0; }

  // DC prediction mode and predictors from the macroblocks to the left,
  // above and above-left (ComponentSyntax::dc_prediction, syntax.rs:118-171)
  int dc_prediction(int x, int y, i32 pred[16]) {
    int mode;
    if (x == 0 && y == 0)
      mode = JXRHIP_PRED_NONE;
    else if (x == 0)
      mode = JXRHIP_PRED_FROM_TOP;
    else if (y == 0)
      mode = JXRHIP_PRED_FROM_LEFT;
    else {
      int deciding = 1;
      u32 first_weight = 1;
      if (syn.kind == kYuv420) {
        deciding = 3;
        first_weight = 8;
      } else if (syn.kind == kYuv422) {
        deciding = 3;
        first_weight = 4;
      } else if (syn.kind == kYuv444 || syn.kind == kYuvK) {
        deciding = 3;
        first_weight = 2;
      }
      u32 hor = 0, ver = 0;
      for (int c = 0; c < deciding; c++) {
        const u32 weight = c == 0 ? first_weight : 1;
        const i64 tl = R(c, (y - 1) * w + x - 1).low[0];
        const u32 left = (u32)std::llabs(tl - R(c, y * w + x - 1).low[0]);
        const u32 top = (u32)std::llabs(tl - R(c, (y - 1) * w + x).low[0]);
        hor = sat_add(hor, sat_mul(left, weight));
        ver = sat_add(ver, sat_mul(top, weight));
      }
      if (sat_mul(hor, 4) < ver)
        mode = JXRHIP_PRED_FROM_TOP;
      else if (sat_mul(ver, 4) < hor)
        mode = JXRHIP_PRED_FROM_LEFT;
      else
        mode = JXRHIP_PRED_FROM_TOP_LEFT;
    }
    for (int c = 0; c < syn.comps; c++) {
      switch (mode) {
        case JXRHIP_PRED_NONE: pred[c] = 0; break;
        case JXRHIP_PRED_FROM_LEFT: pred[c] = R(c, y * w + x - 1).low[0]; break;
        case JXRHIP_PRED_FROM_TOP: pred[c] = R(c, (y - 1) * w + x).low[0]; break;
        default: {
          // subsampled chroma rounds the mean of its neighbours up
          const i64 sum = (i64)R(c, y * w + x - 1).low[0] + R(c, (y - 1) * w + x).low[0];
          pred[c] = (i32)((sum + (c != 0 && syn.joint_chroma() ? 1 : 0)) >> 1);
        }
      }
    }
    return mode;
  }

  // Low-pass prediction of component c from the neighbour the DC came from,
  // when that neighbour used the same LP quantizer (spatial.rs:84-215).
  // remove: the writer's direction, in reverse order.
  void predict_low(int c, i32 *low, int x, int y, int dc_mode, u8 lp_index, bool remove) {
    int mode = JXRHIP_PRED_NONE;
    const Record *n = nullptr;
    if (dc_mode == JXRHIP_PRED_FROM_LEFT && R(c, y * w + x - 1).lp_index == lp_index) {
      mode = JXRHIP_PRED_FROM_LEFT;
      n = &R(c, y * w + x - 1);
    } else if (dc_mode == JXRHIP_PRED_FROM_TOP && R(c, (y - 1) * w + x).lp_index == lp_index) {
      mode = JXRHIP_PRED_FROM_TOP;
      n = &R(c, (y - 1) * w + x);
    }
    // steps: target, source, from own (else the neighbour)
    struct Step {
      int target, source;
      bool own;
    };
    Step steps[3];
    int count = 0;
    const int cols = syn.cols(c), rows = syn.rows(c);
    if (cols == 2 && rows == 2) {
      if (mode == JXRHIP_PRED_FROM_LEFT) steps[count++] = {2, 2, false};
      if (mode == JXRHIP_PRED_FROM_TOP) steps[count++] = {1, 1, false};
    } else if (cols == 2) {
      if (mode == JXRHIP_PRED_FROM_LEFT) {
        steps[count++] = {4, 4, false};
        steps[count++] = {2, 2, false};
        steps[count++] = {6, 6, false};
      } else if (mode == JXRHIP_PRED_FROM_TOP) {
        // the lower half follows the upper half whenever DC comes from above
        steps[count++] = {4, 4, false};
        steps[count++] = {1, 5, false};
        steps[count++] = {5, 1, true};
      } else if (dc_mode == JXRHIP_PRED_FROM_TOP) {
        steps[count++] = {5, 1, true};
      }
    } else {
      if (mode == JXRHIP_PRED_FROM_LEFT) {
        steps[count++] = {4, 4, false};
        steps[count++] = {8, 8, false};
        steps[count++] = {12, 12, false};
      } else if (mode == JXRHIP_PRED_FROM_TOP) {
        steps[count++] = {1, 1, false};
        steps[count++] = {2, 2, false};
        steps[count++] = {3, 3, false};
      }
    }
    for (int k = 0; k < count; k++) {
      const Step &s = steps[remove ? count - 1 - k : k];
      const i32 v = s.own ? low[s.source] : n->low[s.source];
      low[s.target] = remove ? checked_sub(low[s.target], v, "LP prediction") : checked_add(low[s.target], v, "LP prediction");
    }
  }

  // the direction HP prediction takes, from the low-pass coefficients
  // (syntax.rs:173-205)
  int hp_direction(int m) {
    auto energy = [&](int c, std::initializer_list<int> positions) {
      u32 e = 0;
      for (int p : positions) e = sat_add(e, magnitude(R(c, m).low[p]));
      return e;
    };
    u32 hor = energy(0, {1, 2, 3}), ver = energy(0, {4, 8, 12});
    if (syn.kind != kLuma && syn.kind != kNComponent) {
      for (int c = 1; c < 3; c++) {
        if (syn.kind == kYuv420) {
          hor = sat_add(hor, energy(c, {1}));
          ver = sat_add(ver, energy(c, {2}));
        } else if (syn.kind == kYuv422) {
          hor = sat_add(hor, energy(c, {1, 5}));
          ver = sat_add(ver, energy(c, {2, 6}));
        } else {
          hor = sat_add(hor, energy(c, {1}));
          ver = sat_add(ver, energy(c, {4}));
        }
      }
    }
    if (sat_mul(hor, 4) < ver) return JXRHIP_PRED_FROM_LEFT;
    if (sat_mul(ver, 4) < hor) return JXRHIP_PRED_FROM_TOP;
    return JXRHIP_PRED_NONE;
  }

  // ---- reading -----------------------------------------------------------
  void read_lp_index(BitReader &r, int m) {
    slot[m].lp_index = (q.lp_inherits || q.lp_sets() == 1) ? 0 : (u8)read_qp_index(r, q.lp_sets());
  }
  void read_hp_index(BitReader &r, int m) {
    Slot &s = slot[m];
    s.hp_index = q.hp_inherits ? s.lp_index : q.hp_sets() == 1 ? 0 : (u8)read_qp_index(r, q.hp_sets());
  }

  void read_dc(BitReader &r, int x, int y) {
    const int m = y * w + x;
    const int comps = syn.comps;
    const u32 pattern = syn.joint() ? read_vlc(r, kDcPattern) : 0;
    i32 residual[16];
    i32 lap[2] = {0, 0};
    for (int c = 0; c < comps; c++) {
      const bool present = syn.joint() ? (pattern & (4u >> c)) != 0 : r.bit();
      lap[c != 0] += present;
      const unsigned bits = dc_model.bits[c != 0];
      i32 v = 0;
      if (present) {
        const i32 level = read_abs_level(r, dc_vlc.abs[syn.dc_class(c)]);
        v = level - 1;
      }
      if (bits) v = checked_add(wrap_shl(v, bits), (i32)r.bits(bits), "DC refinement");
      if (v != 0 && r.bit()) v = signed_level(v, true);
      residual[c] = v;
    }
    dc_model.update(lap[0], lap[1], syn.kind, comps);
    i32 pred[16];
    const int mode = dc_prediction(x, y, pred);
    for (int c = 0; c < comps; c++) {
      Record &rc = R(c, m);
      std::memset(rc.low, 0, sizeof rc.low);
      rc.low[0] = checked_add(residual[c], pred[c], "DC prediction");
      rc.lp_index = 0;
    }
    slot[m].dc_mode = mode;
    if (adapts_after(x, w)) dc_vlc.adapt();
  }

  void read_lp_refinement(BitReader &r, i32 *low, const int *positions, int count, unsigned bits) {
    for (int k = 0; k < count; k++) {
      const int p = positions[k];
      const i32 flex = read_flex(r, low[p], bits);
      low[p] = checked_add(wrap_shl(low[p], bits), flex, "LP refinement");
    }
  }

  void read_lp(BitReader &r, int x, int y) {
    const int m = y * w + x;
    if (x % 16 == 0) lp_scan.reset_totals();
    const Slot &s = slot[m];
    const int comps = syn.comps;
    const int separate = syn.joint_chroma() ? 1 : comps;
    u32 pattern = 0;
    if (syn.joint()) {
      pattern = cbplp.read(r);
    } else {
      for (int c = 0; c < separate; c++) pattern |= (u32)r.bit() << c;
    }
    i32 lap[2] = {0, 0};
    for (int c = 0; c < separate; c++) {
      i32 low[16];
      std::memcpy(low, R(c, m).low, sizeof low);
      if (pattern & (1u << c)) {
        lap[c != 0] += read_block(r, lp_vlc, c != 0, 1, [&](int index, i32 level) {
          low[lp_scan.order[index] & 15] = level;
          lp_scan.count(index);
        });
      }
      read_lp_refinement(r, low, kTranspose + 1, 15, lp_model.bits[c != 0]);
      if (!syn.joint_chroma()) predict_low(c, low, x, y, s.dc_mode, s.lp_index, false);
      std::memcpy(R(c, m).low, low, sizeof low);
      R(c, m).lp_index = s.lp_index;
    }
    if (syn.joint_chroma()) {
      const bool is420 = syn.kind == kYuv420;
      const int start = is420 ? 10 : 2, count = 16 - start;
      i32 *chroma[2] = {R(1, m).low, R(2, m).low};
      if (pattern & 2) {
        i32 temp[14] = {0};
        lap[1] = read_block(r, lp_vlc, 1, start, [&](int index, i32 level) { temp[index - start] = level; });
        for (int i = 0; i < count; i++) {
          int c, position;
          joint_slot(i, is420, c, position);
          chroma[c - 1][position] = temp[i];
        }
      }
      static const int k420[3] = {2, 1, 3};
      static const int k422[7] = {2, 1, 3, 4, 6, 5, 7};
      const int *positions = is420 ? k420 : k422;
      const int n = is420 ? 3 : 7;
      const unsigned bits = lp_model.bits[1];
      for (int k = 0; k < n; k++)
        for (int c = 0; c < 2; c++) read_lp_refinement(r, chroma[c], positions + k, 1, bits);
      for (int c = 0; c < 3; c++) {
        predict_low(c, R(c, m).low, x, y, s.dc_mode, s.lp_index, false);
        R(c, m).lp_index = s.lp_index;
      }
    }
    lp_model.update(lap[0], lap[1], syn.kind, comps);
    if (adapts_after(x, w)) lp_vlc.adapt();
  }

  // index i of the joint 4:2:0 / 4:2:2 chroma block: its component (1 or 2)
  // and LP position (yuv.rs joint_chroma_slot)
  static void joint_slot(int i, bool is420, int &component, int &position) {
    static const int remap[7] = {4, 1, 2, 3, 5, 6, 7};
    static const int t420[4] = {0, 2, 1, 3};
    static const int t422[8] = {0, 2, 1, 3, 4, 6, 5, 7};
    position = is420 ? t420[remap[(i >> 1) + 1]] : t422[remap[i >> 1]];
    component = (i & 1) + 1;
  }

  // reads the HP band of macroblock m; `hp[c]` is component c's
  // destination: blocks of 16, position 0 of each left untouched
  void read_hp(BitReader &r, int x, int y, int32_t *const *hp, bool spatial_flex, bool spatial_scale, unsigned trim) {
    const int m = y * w + x;
    if (x % 16 == 0) {
      hp_h.reset_totals();
      hp_v.reset_totals();
    }
    const int dir = hp_direction(m);
    u16 patterns[16] = {0};
    if (syn.joint()) {
      read_yuv_patterns(r, x, y, patterns);
    } else {
      for (int c = 0; c < syn.comps; c++) patterns[c] = cbphp.read_component(r, x, y, c);
    }
    const u8 bits[2] = {hp_model.bits[0], hp_model.bits[1]};
    Scan &scan = dir == JXRHIP_PRED_FROM_TOP ? hp_v : hp_h;
    i32 lap[2] = {0, 0};
    for (int c = 0; c < syn.comps; c++) {
      u32 coded = patterns[c];
      int count;
      const int *order = syn.block_order(c, count);
      for (int k = 0; k < count; k++) {
        i32 *blk = hp[c] + order[k] * 16;
        std::memset(blk + 1, 0, 15 * sizeof(i32));
        if (coded & 1) {
          lap[c != 0] += read_block(r, hp_vlc, c != 0, 1, [&](int index, i32 level) {
            blk[scan.order[index] & 15] = level;
            scan.count(index);
          });
        }
        coded >>= 1;
        if (spatial_flex)
          refine_block(r, blk, bits[c != 0], trim);
        else if (spatial_scale)
          scale_block(blk, bits[c != 0]);
      }
    }
    hp_model.update(lap[0], lap[1], syn.kind, syn.comps);
    slot[m].hp_dir = dir;
    slot[m].hp_bits[0] = bits[0];
    slot[m].hp_bits[1] = bits[1];
    if (adapts_after(x, w)) {
      hp_vlc.adapt();
      cbphp.adapt();
    }
  }

  void read_yuv_patterns(BitReader &r, int x, int y, u16 *patterns) {
    const int form = syn.kind == kYuv444 ? 0 : syn.kind == kYuv422 ? 1 : 2;
    cbphp.read_yuv(r, x, y, form, patterns);
  }

  static void refine_block(BitReader &r, i32 *blk, unsigned bits, unsigned trim) {
    if (bits > 15 || trim > 15) syntax_error("flexbits width");
    const unsigned left = bits > trim ? bits - trim : 0;
    if (left == 0) {
      for (int p = 1; p < 16; p++) blk[p] = wrap_shl(blk[p], bits);
      return;
    }
    for (int k = 1; k < 16; k++) {
      const int p = kTranspose[k];
      const i32 coded = blk[p];
      u32 mag;
      bool negative;
      i32 flex;
      if (r.magnitude_sign(left, coded == 0, mag, negative))
        flex = (coded < 0 || negative) ? -(i32)mag : (i32)mag;
      else
        flex = read_flex(r, coded, left);
      const i64 v = (i64)wrap_shl(coded, bits) + wrap_shl(flex, trim);
      if (v < INT32_MIN || v > INT32_MAX) overflow("HP refinement");
      blk[p] = (i32)v;
    }
  }
  static void scale_block(i32 *blk, unsigned bits) {
    for (int p = 1; p < 16; p++) blk[p] = wrap_shl(blk[p], bits);
  }

  void read_flex_packet(BitReader &r, int m, int32_t *const *hp, unsigned trim) {
    for (int c = 0; c < syn.comps; c++) {
      int count;
      const int *order = syn.block_order(c, count);
      for (int k = 0; k < count; k++) refine_block(r, hp[c] + order[k] * 16, slot[m].hp_bits[c != 0], trim);
    }
  }
  void scale_hp(int m, int32_t *const *hp) {
    for (int c = 0; c < syn.comps; c++) {
      int count;
      const int *order = syn.block_order(c, count);
      for (int k = 0; k < count; k++) scale_block(hp[c] + order[k] * 16, slot[m].hp_bits[c != 0]);
    }
  }

  // ---- writing ---------------------------------------------------------
  void write_lp_index(BitWriter &wr, int m) {
    if (!(q.lp_inherits || q.lp_sets() == 1)) write_qp_index(wr, slot[m].lp_index, q.lp_sets());
  }
  void write_hp_index(BitWriter &wr, int m) {
    if (!q.hp_inherits && q.hp_sets() != 1) write_qp_index(wr, slot[m].hp_index, q.hp_sets());
  }

  // low[c] of macroblock m: absolute DC/LP values (positions 0..blocks-1)
  void write_dc(BitWriter &wr, int x, int y, const i32 *const *low) {
    const int m = y * w + x;
    const int comps = syn.comps;
    i32 pred[16];
    const int mode = dc_prediction(x, y, pred);
    i32 residual[16];
    bool present[16];
    u32 pattern = 0;
    for (int c = 0; c < comps; c++) {
      residual[c] = checked_sub(low[c][0], pred[c], "DC residual");
      present[c] = (magnitude(residual[c]) >> dc_model.bits[c != 0]) != 0;
      if (present[c]) pattern |= 4u >> c;
    }
    if (syn.joint()) write_vlc(wr, kDcPattern, pattern);
    i32 lap[2] = {0, 0};
    for (int c = 0; c < comps; c++) {
      if (!syn.joint()) wr.flag(present[c]);
      lap[c != 0] += present[c];
      const unsigned bits = dc_model.bits[c != 0];
      const u32 mag = magnitude(residual[c]);
      if (present[c]) write_abs_level(wr, dc_vlc.abs[syn.dc_class(c)], (mag >> bits) + 1);
      wr.put(mag & ((1u << bits) - 1), bits);
      if (mag) wr.flag(residual[c] < 0);
    }
    dc_model.update(lap[0], lap[1], syn.kind, comps);
    for (int c = 0; c < comps; c++) {
      Record &rc = R(c, m);
      std::memset(rc.low, 0, sizeof rc.low);
      rc.low[0] = low[c][0];
      rc.lp_index = 0;
    }
    slot[m].dc_mode = mode;
    if (adapts_after(x, w)) dc_vlc.adapt();
  }

  // split of a residual into its coded part and refinement at `bits`
  static void split(i32 v, unsigned bits, i32 &coded, u32 &refinement) {
    const u32 mag = magnitude(v);
    const i32 level = (i32)(mag >> bits);
    coded = v < 0 ? -level : level;
    refinement = mag & ((1u << bits) - 1);
  }

  void write_lp_refinement(BitWriter &wr, const i32 *res, const i32 *coded, const int *positions, int count,
                           unsigned bits) {
    for (int k = 0; k < count; k++) {
      const int p = positions[k];
      i32 cp;
      u32 ref;
      split(res[p], bits, cp, ref);
      (void)cp;
      write_flex(wr, coded[p], ref, bits, res[p] < 0);
    }
  }

  void write_run_levels(BitWriter &wr, AcVlc &st, int cls, int start, const i32 *values_by_index) {
    int at[16];
    i32 level[16];
    int n = 0;
    for (int i = start; i < 16; i++)
      if (values_by_index[i]) {
        at[n] = i;
        level[n] = values_by_index[i];
        n++;
      }
    if (n) write_block(wr, st, cls, start, at, level, n);
  }

  // reads block values through an adaptive scan and counts what it finds,
  // as the reader's placements will
  static int scan_block(Scan &scan, const i32 *values, int start, i32 *by_index) {
    int n = 0;
    for (int i = 0; i < 16; i++) by_index[i] = values[scan.order[i] & 15];
    u32 occupied = 0;
    for (int i = start; i < 16; i++)
      if (by_index[i]) occupied |= 1u << i;
    for (int i = 0; i < start; i++) by_index[i] = 0;
    while (occupied) {
      const int i = __builtin_ctz(occupied);
      scan.count(i);
      occupied &= occupied - 1;
      n++;
    }
    return n;
  }

  void write_lp(BitWriter &wr, int x, int y, const i32 *const *low) {
    const int m = y * w + x;
    if (x % 16 == 0) lp_scan.reset_totals();
    const Slot &s = slot[m];
    const int comps = syn.comps;
    const int separate = syn.joint_chroma() ? 1 : comps;
    i32 res[16][16], coded[16][16];
    for (int c = 0; c < comps; c++) {
      std::memset(res[c], 0, sizeof res[c]);
      const int blocks = syn.blocks(c);
      for (int p = 0; p < blocks; p++) res[c][p] = low[c][p];
      predict_low(c, res[c], x, y, s.dc_mode, s.lp_index, true);
      const unsigned bits = lp_model.bits[c != 0];
      for (int p = 0; p < 16; p++) {
        u32 ref;
        split(res[c][p], bits, coded[c][p], ref);
      }
      coded[c][0] = 0;
    }
    const bool is420 = syn.kind == kYuv420;
    const int jstart = is420 ? 10 : 2, jcount = 16 - jstart;
    i32 joint[16] = {0};
    bool joint_coded = false;
    if (syn.joint_chroma()) {
      for (int i = 0; i < jcount; i++) {
        int c, position;
        joint_slot(i, is420, c, position);
        joint[jstart + i] = coded[c][position];
        joint_coded = joint_coded || joint[jstart + i] != 0;
      }
    }
    u32 pattern = 0;
    for (int c = 0; c < separate; c++) {
      bool any = false;
      for (int p = 1; p < 16; p++) any = any || coded[c][p] != 0;
      if (any) pattern |= 1u << c;
    }
    if (joint_coded) pattern |= 2;
    if (syn.joint()) {
      cbplp.write(wr, pattern);
    } else {
      for (int c = 0; c < separate; c++) wr.flag((pattern >> c) & 1);
    }
    i32 lap[2] = {0, 0};
    for (int c = 0; c < separate; c++) {
      if (pattern & (1u << c)) {
        i32 by_index[16];
        lap[c != 0] += scan_block(lp_scan, coded[c], 1, by_index);
        write_run_levels(wr, lp_vlc, c != 0, 1, by_index);
      }
      write_lp_refinement(wr, res[c], coded[c], kTranspose + 1, 15, lp_model.bits[c != 0]);
    }
    if (syn.joint_chroma()) {
      if (joint_coded) {
        for (int i = jstart; i < 16; i++) lap[1] += joint[i] != 0;
        write_run_levels(wr, lp_vlc, 1, jstart, joint);
      }
      static const int k420[3] = {2, 1, 3};
      static const int k422[7] = {2, 1, 3, 4, 6, 5, 7};
      const int *positions = is420 ? k420 : k422;
      const int n = is420 ? 3 : 7;
      for (int k = 0; k < n; k++)
        for (int c = 1; c < 3; c++) write_lp_refinement(wr, res[c], coded[c], positions + k, 1, lp_model.bits[1]);
    }
    lp_model.update(lap[0], lap[1], syn.kind, comps);
    for (int c = 0; c < comps; c++) {
      std::memset(R(c, m).low, 0, sizeof R(c, m).low);
      for (int p = 0; p < syn.blocks(c); p++) R(c, m).low[p] = low[c][p];
      R(c, m).lp_index = s.lp_index;
    }
    if (adapts_after(x, w)) lp_vlc.adapt();
  }

  // hp[c]: component c's blocks of 16 (positions 1..15 the coefficients);
  // `flex` receives, when not null, the flexbits to write later (frequency
  // mode); in spatial mode they are written here when `spatial_flex`.
  void write_hp(BitWriter &wr, int x, int y, const int32_t *const *hp, bool spatial_flex, unsigned trim) {
    const int m = y * w + x;
    if (x % 16 == 0) {
      hp_h.reset_totals();
      hp_v.reset_totals();
    }
    const int dir = hp_direction(m);
    const u8 bits[2] = {hp_model.bits[0], hp_model.bits[1]};
    u16 patterns[16] = {0};
    for (int c = 0; c < syn.comps; c++) {
      int count;
      const int *order = syn.block_order(c, count);
      for (int k = 0; k < count; k++) {
        const i32 *blk = hp[c] + order[k] * 16;
        bool any = false;
        for (int p = 1; p < 16; p++) any = any || (magnitude(blk[p]) >> bits[c != 0]) != 0;
        if (any) patterns[c] |= (u16)(1u << k);
      }
    }
    if (syn.joint()) {
      const int form = syn.kind == kYuv444 ? 0 : syn.kind == kYuv422 ? 1 : 2;
      cbphp.write_yuv(wr, x, y, form, patterns);
    } else {
      for (int c = 0; c < syn.comps; c++) cbphp.write_component(wr, x, y, c, patterns[c]);
    }
    Scan &scan = dir == JXRHIP_PRED_FROM_TOP ? hp_v : hp_h;
    i32 lap[2] = {0, 0};
    for (int c = 0; c < syn.comps; c++) {
      int count;
      const int *order = syn.block_order(c, count);
      for (int k = 0; k < count; k++) {
        const i32 *blk = hp[c] + order[k] * 16;
        i32 coded[16];
        u32 ref;
        for (int p = 0; p < 16; p++) split(blk[p], bits[c != 0], coded[p], ref);
        coded[0] = 0;
        if (patterns[c] & (1u << k)) {
          i32 by_index[16];
          lap[c != 0] += scan_block(scan, coded, 1, by_index);
          write_run_levels(wr, hp_vlc, c != 0, 1, by_index);
        }
        if (spatial_flex) write_flex_block(wr, blk, bits[c != 0], trim);
      }
    }
    hp_model.update(lap[0], lap[1], syn.kind, syn.comps);
    slot[m].hp_dir = dir;
    slot[m].hp_bits[0] = bits[0];
    slot[m].hp_bits[1] = bits[1];
    if (adapts_after(x, w)) {
      hp_vlc.adapt();
      cbphp.adapt();
    }
  }

  static void write_flex_block(BitWriter &wr, const i32 *blk, unsigned bits, unsigned trim) {
    const unsigned left = bits > trim ? bits - trim : 0;
    if (left == 0) return;
    for (int k = 1; k < 16; k++) {
      const int p = kTranspose[k];
      i32 coded;
      u32 ref;
      split(blk[p], bits, coded, ref);
      const u32 kept = ref >> trim;
      write_flex(wr, coded, kept, left, blk[p] < 0);
    }
  }

  void write_flex_packet(BitWriter &wr, int m, const int32_t *const *hp, unsigned trim) {
    for (int c = 0; c < syn.comps; c++) {
      int count;
      const int *order = syn.block_order(c, count);
      for (int k = 0; k < count; k++) write_flex_block(wr, hp[c] + order[k] * 16, slot[m].hp_bits[c != 0], trim);
    }
  }
};

// ---- decode orchestration -----------------------------------------------------
u32 retained_bands(u32 scale, u32 available) {
  if (scale == JXRHIP_SCALE_FULL) return available;
  if (scale == JXRHIP_SCALE_QUARTER && has_lp(available)) return JXRHIP_BANDS_NO_HIGH_PASS;
  return JXRHIP_BANDS_DC_ONLY;
}

int coefficients_per_mb(u32 bands, int blocks) {
  return bands == JXRHIP_BANDS_DC_ONLY ? 1 : bands == JXRHIP_BANDS_NO_HIGH_PASS ? blocks : blocks * 16;
}

// Byte range [begin, end) of each packet the index lists, each ending at
// the next packet in the file or at the end of the codestream
// (prepare.rs indexed_tile_packet_ranges, frequency/ranges.rs).
struct PacketRange {
  size_t begin = 0, end = 0;
  bool present = false;
};

std::vector<PacketRange> packet_ranges(const Stream &s) {
  const size_t tiles = s.cols.size() * s.rows.size();
  const size_t per = packets_per_tile(s.h);
  std::vector<PacketRange> out(tiles * per);
  const size_t base = s.dir.data_offset, end = s.size;
  if (s.dir.offsets.empty()) {
    // one spatial tile right after the headers
    out[0] = {base, end, true};
    return out;
  }
  if (s.dir.offsets.size() != tiles * per) syntax_error("tile index entry count");
  std::vector<size_t> starts;
  for (size_t i = 0; i < out.size(); i++) {
    const u64 rel = s.dir.offsets[i];
    // a zero offset for a flexbits packet says it was left out
    if (per == 4 && i % 4 == 3 && rel == 0) continue;
    if (rel >= end - base) syntax_error("tile packet offset past the codestream");
    out[i] = {base + (size_t)rel, 0, true};
    starts.push_back(base + (size_t)rel);
  }
  std::sort(starts.begin(), starts.end());
  if (s.h.image.frequency && std::adjacent_find(starts.begin(), starts.end()) != starts.end())
    syntax_error("duplicate frequency packet offsets");
  starts.erase(std::unique(starts.begin(), starts.end()), starts.end());
  for (PacketRange &r : out) {
    if (!r.present) continue;
    auto it = std::upper_bound(starts.begin(), starts.end(), r.begin);
    r.end = it == starts.end() ? end : *it;
  }
  return out;
}

void packet_start(BitReader &r) {
  const u32 start = r.bits(24);
  if (start != 1) syntax_error("tile start code " + std::to_string(start));
  r.bits(8);
}

}  // namespace

struct jxrhip_decoded {
  jxr::ArenaOut primary, alpha;
  bool has_alpha = false;
};

namespace {

std::atomic<uint64_t> g_decoded_identity{1ull << 62};

struct TileJob {
  size_t index;          // raster tile index
  u32 x0, y0, w, h;      // macroblocks
  size_t mb_before;      // selected macroblocks before this tile
};

// Decodes one stream's tiles into `out` (the primary codestream's planes, or
// a separate alpha codestream's one plane): its coefficients into `external`
// when given (at least `capacity` values), else into out.coefficients. With
// `count_only`, only the arena's coefficient count is computed.
void decode_stream(const Stream &s, u32 scale, const u8 *mask, u32 threads, jxr::ArenaOut &out,
                   int32_t *external = nullptr, uint64_t capacity = 0, uint64_t *count_only = nullptr) {
  const Headers &h = s.h;
  const PlaneHeader *planes[2] = {&h.primary, h.image.alpha_plane ? &h.alpha : nullptr};
  const int nplanes = planes[1] ? 2 : 1;
  for (int p = 0; p < nplanes; p++) {
    Syntax tmp;
    if (!syntax_of(planes[p]->internal, (u32)planes[p]->comps, tmp)) unsupported("internal colour format of a plane");
  }
  if (planes[1] && (planes[1]->internal != 0 || planes[1]->comps != 1))
    unsupported("integrated alpha plane component layout");
  if (scale != JXRHIP_SCALE_FULL && !h.image.frequency)
    unsupported("native reduced resolution requires frequency-mode packets");
  const u32 avail[2] = {bands_of_code(h.primary.bands_code), planes[1] ? bands_of_code(h.alpha.bands_code) : 0};
  if (scale == JXRHIP_SCALE_QUARTER && !has_lp(avail[0])) unsupported("quarter-resolution decode requires an LP band");
  if (planes[1] && ((has_lp(avail[1]) && !has_lp(avail[0])) || (has_hp(avail[1]) && !has_hp(avail[0])) ||
                    (has_flex(avail[1]) && !has_flex(avail[0]))))
    unsupported("integrated alpha bands exceed primary bands");
  const u32 bands[2] = {retained_bands(scale, avail[0]), retained_bands(scale, avail[1])};

  // components of the arena: primary's, then the alpha plane
  struct Comp {
    int plane, c, cols, rows;
  };
  std::vector<Comp> comps;
  for (int p = 0; p < nplanes; p++)
    for (int c = 0; c < planes[p]->syn.comps; c++)
      comps.push_back({p, c, planes[p]->syn.cols(c), planes[p]->syn.rows(c)});
  if (comps.size() > 16 + 1) unsupported("more components than the tile syntax carries");

  // selected tiles in raster order
  std::vector<TileJob> jobs;
  size_t mbs = 0;
  {
    u32 y0 = 0;
    size_t t = 0;
    for (u32 th : s.rows) {
      u32 x0 = 0;
      for (u32 tw : s.cols) {
        if (!mask || mask[t]) {
          jobs.push_back({t, x0, y0, tw, th, mbs});
          mbs += (size_t)tw * th;
        }
        x0 += tw;
        t++;
      }
      y0 += th;
    }
  }
  if (jobs.empty()) syntax_error("no tile selected");
  std::vector<int> cpm;
  size_t total = 0;
  std::vector<size_t> comp_coef;
  for (const Comp &c : comps) {
    cpm.push_back(coefficients_per_mb(bands[c.plane], c.cols * c.rows));
    comp_coef.push_back(total);
    total += mbs * (size_t)cpm.back();
  }
  if (total > UINT32_MAX) unsupported("more than 2^32 coefficients in one arena");
  if (count_only) {
    *count_only = total;
    return;
  }
  out.coefficients.clear();
  int32_t *base;
  if (external) {
    if (capacity < total) throw Failure{JXRHIP_ERR_BUFFER_TOO_SMALL, "coefficient storage below the arena's size"};
    base = external;
  } else {
    out.coefficients.resize(total);
    base = out.coefficients.data();
  }
  const size_t all_mbs = mbs * comps.size();
  out.offsets.assign(all_mbs, 0);
  out.quantizers.assign(all_mbs * 3, 0);
  out.coded_x.assign(all_mbs, 0);
  out.coded_y.assign(all_mbs, 0);
  out.bands.assign(all_mbs, 0);
  out.preds.assign(all_mbs, 0);
  out.qps.assign(all_mbs * 3, 0);
  out.planes.clear();
  for (size_t k = 0; k < comps.size(); k++) {
    jxrhip_coefficient_plane cp{};
    cp.coefficient_offset = comp_coef[k];
    cp.coefficient_count = mbs * (size_t)cpm[k];
    cp.macroblock_offset = k * mbs;
    cp.macroblock_count = mbs;
    cp.block_columns = (u32)comps[k].cols;
    cp.block_rows = (u32)comps[k].rows;
    out.planes.push_back(cp);
  }
  out.tile_w = s.cols;
  out.tile_h = s.rows;

  const std::vector<PacketRange> ranges = packet_ranges(s);
  const size_t per = packets_per_tile(s.h);
  const bool frequency = h.image.frequency;

  auto decode_tile = [&](const TileJob &job) {
    const int w = (int)job.w, th = (int)job.h;
    std::vector<std::unique_ptr<PlaneCoder>> coders;
    for (int p = 0; p < nplanes; p++) coders.emplace_back(new PlaneCoder(planes[p], bands[p], w, th));
    // component k's destination for macroblock m of this tile
    auto dest = [&](size_t k, int m) { return base + comp_coef[k] + (job.mb_before + m) * cpm[k]; };
    std::vector<size_t> first_k(nplanes);
    {
      size_t k = 0;
      for (int p = 0; p < nplanes; p++) {
        first_k[p] = k;
        k += planes[p]->syn.comps;
      }
    }
    auto hp_of = [&](int p, int m, int32_t **ptrs) {
      for (int c = 0; c < planes[p]->syn.comps; c++) ptrs[c] = dest(first_k[p] + c, m);
    };
    const unsigned trim_none = 0;
    if (!frequency) {
      const PacketRange &pr = ranges[job.index];
      if (!pr.present) syntax_error("missing tile packet");
      BitReader r(s.data + pr.begin, pr.end - pr.begin);
      packet_start(r);
      const unsigned trim = h.image.trim ? r.bits(4) : trim_none;
      for (int p = 0; p < nplanes; p++) {
        PlaneCoder &pc = *coders[p];
        read_dc_quant(r, *planes[p], pc.q);
        read_lp_quant(r, *planes[p], pc.q);
        read_hp_quant(r, *planes[p], pc.q);
      }
      for (int y = 0; y < th; y++)
        for (int x = 0; x < w; x++) {
          const int m = y * w + x;
          for (int p = 0; p < nplanes; p++) {
            PlaneCoder &pc = *coders[p];
            pc.read_lp_index(r, m);
            pc.read_hp_index(r, m);
            pc.read_dc(r, x, y);
            if (has_lp(pc.bands)) pc.read_lp(r, x, y);
            if (has_hp(pc.bands)) {
              int32_t *hp[16];
              hp_of(p, m, hp);
              pc.read_hp(r, x, y, hp, has_flex(pc.bands), !has_flex(pc.bands), trim);
            }
          }
        }
      r.align();
    } else {
      const size_t first = job.index * per;
      auto open = [&](size_t band) -> BitReader {
        const PacketRange &pr = ranges[first + band];
        BitReader r(s.data + pr.begin, pr.end - pr.begin);
        packet_start(r);
        return r;
      };
      if (!ranges[first].present) syntax_error("missing DC packet");
      BitReader dc = open(0);
      for (int p = 0; p < nplanes; p++) read_dc_quant(dc, *planes[p], coders[p]->q);
      for (int y = 0; y < th; y++)
        for (int x = 0; x < w; x++)
          for (int p = 0; p < nplanes; p++) coders[p]->read_dc(dc, x, y);
      dc.align();
      if (has_lp(bands[0])) {
        if (per < 2 || !ranges[first + 1].present) syntax_error("missing LP frequency packet");
        BitReader lp = open(1);
        for (int p = 0; p < nplanes; p++)
          if (has_lp(bands[p])) read_lp_quant(lp, *planes[p], coders[p]->q);
        for (int y = 0; y < th; y++)
          for (int x = 0; x < w; x++)
            for (int p = 0; p < nplanes; p++)
              if (has_lp(bands[p])) {
                coders[p]->read_lp_index(lp, y * w + x);
                coders[p]->read_lp(lp, x, y);
              }
        lp.align();
      }
      if (has_hp(bands[0])) {
        if (per < 3 || !ranges[first + 2].present) syntax_error("missing HP frequency packet");
        BitReader hpr = open(2);
        for (int p = 0; p < nplanes; p++)
          if (has_hp(bands[p])) read_hp_quant(hpr, *planes[p], coders[p]->q);
        for (int y = 0; y < th; y++)
          for (int x = 0; x < w; x++)
            for (int p = 0; p < nplanes; p++)
              if (has_hp(bands[p])) {
                const int m = y * w + x;
                coders[p]->read_hp_index(hpr, m);
                int32_t *hp[16];
                hp_of(p, m, hp);
                coders[p]->read_hp(hpr, x, y, hp, false, false, 0);
              }
        hpr.align();
        const bool flex_packet = has_flex(bands[0]) && per == 4 && ranges[first + 3].present;
        if (flex_packet) {
          BitReader fr = open(3);
          const unsigned trim = h.image.trim ? fr.bits(4) : trim_none;
          for (int m = 0; m < w * th; m++)
            for (int p = 0; p < nplanes; p++) {
              int32_t *hp[16];
              hp_of(p, m, hp);
              if (has_flex(bands[p]))
                coders[p]->read_flex_packet(fr, m, hp, trim);
              else if (has_hp(bands[p]))
                coders[p]->scale_hp(m, hp);
            }
          fr.align();
        } else {
          for (int m = 0; m < w * th; m++)
            for (int p = 0; p < nplanes; p++)
              if (has_hp(bands[p])) {
                int32_t *hp[16];
                hp_of(p, m, hp);
                coders[p]->scale_hp(m, hp);
              }
        }
      }
    }
    // store: DC/LP values and each macroblock's record
    for (int p = 0; p < nplanes; p++) {
      PlaneCoder &pc = *coders[p];
      const PlaneHeader &ph = *planes[p];
      for (int c = 0; c < ph.syn.comps; c++) {
        const size_t k = first_k[p] + c;
        const int blocks = comps[k].cols * comps[k].rows;
        for (int y = 0; y < th; y++)
          for (int x = 0; x < w; x++) {
            const int m = y * w + x;
            const Slot &sl = pc.slot[m];
            const Record &rc = pc.R(c, m);
            int32_t *d = dest(k, m);
            if (bands[p] == JXRHIP_BANDS_DC_ONLY)
              d[0] = rc.low[0];
            else if (bands[p] == JXRHIP_BANDS_NO_HIGH_PASS)
              for (int b = 0; b < blocks; b++) d[b] = rc.low[b];
            else
              for (int b = 0; b < blocks; b++) d[b * 16] = rc.low[b];
            const size_t mi = k * mbs + job.mb_before + m;
            out.offsets[mi] = (u32)(d - base);
            const size_t set_lp = (size_t)sl.lp_index * pc.q.comps + c, set_hp = (size_t)sl.hp_index * pc.q.comps + c;
            if (set_lp >= pc.q.lp.size() || set_hp >= pc.q.hp.size()) syntax_error("quantizer index past its table");
            const u8 qd = pc.q.dc[c], ql = pc.q.lp[set_lp], qh = pc.q.hp[set_hp];
            const u32 lb = c == 0 ? 1 : 0;
            out.quantizers[mi * 3 + 0] = jxrhip_quant_step(qd, ph.scaled, lb);
            out.quantizers[mi * 3 + 1] = jxrhip_quant_step(ql, ph.scaled, lb);
            out.quantizers[mi * 3 + 2] = jxrhip_quant_step(qh, ph.scaled, 1);
            out.qps[mi * 3 + 0] = qd;
            out.qps[mi * 3 + 1] = ql;
            out.qps[mi * 3 + 2] = qh;
            out.bands[mi] = (u8)bands[p];
            out.preds[mi] = has_hp(bands[p]) ? (u8)sl.hp_dir : (u8)JXRHIP_PRED_NONE;
            out.coded_x[mi] = job.x0 + (u32)x;
            out.coded_y[mi] = job.y0 + (u32)y;
          }
      }
    }
  };

  // tiles on host threads, each into its own windows
  u32 nthreads = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
  nthreads = std::min<u32>(std::min<u32>(nthreads, 16u), (u32)jobs.size());
  if (nthreads <= 1) {
    for (const TileJob &j : jobs) decode_tile(j);
  } else {
    std::atomic<size_t> next{0};
    std::vector<char> failed(jobs.size(), 0);
    std::vector<Failure> per_job(jobs.size());
    std::vector<std::thread> pool;
    for (u32 t = 0; t < nthreads; t++)
      pool.emplace_back([&]() {
        for (size_t i; (i = next.fetch_add(1)) < jobs.size();) {
          try {
            decode_tile(jobs[i]);
          } catch (const Failure &f) {
            per_job[i] = f;
            failed[i] = true;
          } catch (const std::bad_alloc &) {
            per_job[i] = Failure{JXRHIP_ERR_INVALID_ARGUMENT, "out of host memory"};
            failed[i] = true;
          }
        }
      });
    for (auto &th : pool) th.join();
    for (size_t i = 0; i < jobs.size(); i++)
      if (failed[i]) throw per_job[i];  // the first failing tile in raster order
  }

  jxrhip_arena &a = out.arena;
  std::memset(&a, 0, sizeof a);
  a.coefficients = base;
  a.coefficient_count = total;
  a.macroblock_count = out.offsets.size();
  a.coefficient_offsets = out.offsets.data();
  a.quantizers = out.quantizers.data();
  a.bands = out.bands.data();
  a.hp_predictions = out.preds.data();
  a.coded_x = out.coded_x.data();
  a.coded_y = out.coded_y.data();
  a.planes = out.planes.data();
  a.plane_count = (u32)out.planes.size();
  a.overlap = h.image.overlap;
  a.hard_tiles = h.image.hard ? 1 : 0;
  a.tile_column_count = (u32)out.tile_w.size();
  a.tile_column_widths = out.tile_w.data();
  a.tile_row_count = (u32)out.tile_h.size();
  a.tile_row_heights = out.tile_h.data();
  a.identity = g_decoded_identity.fetch_add(1);
}

// ---- codes of the library's enums ----------------------------------------------
u32 color_of_output_code(u32 code) {
  static const u32 map[9] = {JXRHIP_COLOR_LUMA, JXRHIP_COLOR_YUV420, JXRHIP_COLOR_YUV422,
                             JXRHIP_COLOR_YUV444, JXRHIP_COLOR_CMYK, JXRHIP_COLOR_CMYK_DIRECT,
                             JXRHIP_COLOR_NCOMPONENT, JXRHIP_COLOR_RGB, JXRHIP_COLOR_RGBE};
  return map[code];
}
u32 color_of_internal_code(u32 code) {
  static const u32 map[7] = {JXRHIP_COLOR_LUMA, JXRHIP_COLOR_YUV420, JXRHIP_COLOR_YUV422, JXRHIP_COLOR_YUV444,
                             JXRHIP_COLOR_YUVK, 0, JXRHIP_COLOR_NCOMPONENT};
  return map[code];
}
u32 depth_of_code(u32 code) {
  switch (code) {
    case 0: return JXRHIP_DEPTH_BIT1_WHITE;
    case 1: return JXRHIP_DEPTH_U8;
    case 2: return JXRHIP_DEPTH_U16;
    case 3: return JXRHIP_DEPTH_I16;
    case 4: return JXRHIP_DEPTH_F16;
    case 6: return JXRHIP_DEPTH_I32;
    case 7: return JXRHIP_DEPTH_F32;
    case 8: return JXRHIP_DEPTH_RGB555;
    case 9: return JXRHIP_DEPTH_U10;
    case 10: return JXRHIP_DEPTH_RGB565;
    default: return JXRHIP_DEPTH_BIT1_BLACK;
  }
}

// ---- writing --------------------------------------------------------------------
// The distinct QP tuples of one plane in one tile, and each macroblock's
// index into them.
struct TileSets {
  std::vector<u8> dc;           // one tuple
  std::vector<std::vector<u8>> lp, hp;
  std::vector<u8> lp_index, hp_index;
  bool hp_follows_lp = true;
};

u32 find_or_add(std::vector<std::vector<u8>> &sets, const std::vector<u8> &t) {
  for (size_t i = 0; i < sets.size(); i++)
    if (sets[i] == t) return (u32)i;
  if (sets.size() == 16) unsupported("more than sixteen quantizer sets in one tile");
  sets.push_back(t);
  return (u32)(sets.size() - 1);
}

struct WritePlane {
  PlaneHeader hdr;
  const jxr::ArenaOut *src;
  int first_plane;  // first arena plane of its components
  u32 bands;
};

// macroblock index within an arena plane of tile t's macroblock (x, y)
struct TileMap {
  std::vector<u32> cols, rows;
  std::vector<size_t> before;  // macroblocks before each tile, raster
};

std::vector<u8> write_stream(const jxrhip_encode_params &prm, std::vector<WritePlane> &wp, const TileMap &tm,
                             u32 out_color, u32 out_depth, const jxrhip_codestream_options &opt, bool alpha_only) {
  (void)alpha_only;
  const size_t ntiles = tm.cols.size() * tm.rows.size();
  const int np = (int)wp.size();
  // per tile, per plane: QP sets
  std::vector<std::vector<TileSets>> sets(ntiles, std::vector<TileSets>((size_t)np));
  for (size_t t = 0; t < ntiles; t++) {
    const u32 tw = tm.cols[t % tm.cols.size()], th = tm.rows[t / tm.cols.size()];
    for (int p = 0; p < np; p++) {
      TileSets &ts = sets[t][p];
      const int comps = wp[p].hdr.comps;
      for (u32 m = 0; m < tw * th; m++) {
        std::vector<u8> dc(comps), lp(comps), hp(comps);
        for (int c = 0; c < comps; c++) {
          const jxrhip_coefficient_plane &pl = wp[p].src->planes[wp[p].first_plane + c];
          const size_t mi = pl.macroblock_offset + tm.before[t] + m;
          dc[c] = wp[p].src->qps[mi * 3 + 0];
          lp[c] = wp[p].src->qps[mi * 3 + 1];
          hp[c] = wp[p].src->qps[mi * 3 + 2];
        }
        if (m == 0)
          ts.dc = dc;
        else if (dc != ts.dc)
          unsupported("DC quantizer that varies within a tile");
        ts.lp_index.push_back((u8)find_or_add(ts.lp, lp));
        ts.hp_index.push_back((u8)find_or_add(ts.hp, hp));
      }
      ts.hp_follows_lp = ts.hp == ts.lp && ts.hp_index == ts.lp_index;
    }
  }
  // image-level QPs where one tuple serves every tile
  for (int p = 0; p < np; p++) {
    PlaneHeader &hd = wp[p].hdr;
    bool dc_uniform = true, lp_uniform = true, hp_uniform = true;
    for (size_t t = 0; t < ntiles; t++) {
      dc_uniform = dc_uniform && sets[t][p].dc == sets[0][p].dc;
      lp_uniform = lp_uniform && sets[t][p].lp.size() == 1 && sets[t][p].lp[0] == sets[0][p].lp[0];
      hp_uniform = hp_uniform && sets[t][p].hp.size() == 1 && sets[t][p].hp[0] == sets[0][p].hp[0];
    }
    hd.has_dc = dc_uniform;
    if (dc_uniform) hd.dc = sets[0][p].dc;
    hd.has_lp = lp_uniform && has_lp(wp[p].bands);
    if (hd.has_lp) hd.lp = sets[0][p].lp[0];
    hd.has_hp = hp_uniform && has_hp(wp[p].bands);
    if (hd.has_hp) hd.hp = sets[0][p].hp[0];
  }

  // headers
  BitWriter hw;
  const u32 width = prm.width, height = prm.height;
  bool short_header = opt.short_header == 0 && width <= 65536 && height <= 65536;
  for (u32 v : tm.cols) short_header = short_header && v <= 255;
  for (u32 v : tm.rows) short_header = short_header && v <= 255;
  const bool tiled = ntiles > 1;
  const bool frequency = opt.frequency_mode != 0;
  const bool index = tiled || frequency;
  const bool trim_present = opt.trim_flexbits != JXRHIP_TRIM_NONE;
  const unsigned trim = trim_present ? opt.trim_flexbits : 0;
  if (trim_present && trim > 15) unsupported("trim flexbits above 15");
  if (tm.cols.size() > 4096 || tm.rows.size() > 4096) unsupported("more than 4096 tiles on an axis");
  hw.put64(kSignature, 64);
  hw.put(1, 4);
  hw.flag(prm.hard_tiles != 0);
  hw.put(1, 3);
  hw.flag(tiled);
  hw.flag(frequency);
  hw.put(0, 3);
  hw.flag(index);
  hw.put(prm.overlap, 2);
  hw.flag(short_header);
  hw.flag(prm.input_kind != JXRHIP_INPUT_U8);  // long word
  hw.flag(false);                                 // windowing: margins inferred
  hw.flag(trim_present);
  hw.flag(false);
  hw.flag(false);  // red/blue
  hw.flag(false);  // premultiplied
  hw.flag(np == 2);
  hw.put(out_color, 4);
  hw.put(out_depth, 4);
  const unsigned db = short_header ? 16 : 32;
  hw.put(width - 1, db);
  hw.put(height - 1, db);
  if (tiled) {
    hw.put((u32)tm.cols.size() - 1, 12);
    hw.put((u32)tm.rows.size() - 1, 12);
    const unsigned tb = short_header ? 8 : 16;
    for (size_t i = 0; i + 1 < tm.cols.size(); i++) hw.put(tm.cols[i], tb);
    for (size_t i = 0; i + 1 < tm.rows.size(); i++) hw.put(tm.rows[i], tb);
  }
  for (int p = 0; p < np; p++) {
    const PlaneHeader &hd = wp[p].hdr;
    hw.put(hd.internal, 3);
    hw.flag(hd.scaled);
    hw.put(hd.bands_code, 4);
    switch (hd.internal) {
      case 1:
        hw.put(0, 1);
        hw.put(hd.cx, 3);
        hw.put(0, 1);
        hw.put(hd.cy, 3);
        break;
      case 2:
        hw.put(0, 1);
        hw.put(hd.cx, 3);
        hw.put(0, 4);
        break;
      case 3: hw.put(0, 8); break;
      case 6:
        if (hd.comps <= 15) {
          hw.put((u32)hd.comps - 1, 4);
          hw.put(0, 4);
        } else {
          hw.put(15, 4);
          hw.put((u32)hd.comps - 16, 12);
        }
        break;
      default: break;
    }
    if (depth_has_shift(out_depth)) hw.put(hd.shift_bits, 8);
    if (out_depth == 7) {
      hw.put(hd.mantissa, 8);
      hw.put((u32)(u8)(int8_t)hd.exponent_bias, 8);
    }
    hw.flag(hd.has_dc);
    if (hd.has_dc) write_qp_set(hw, hd.dc.data(), hd.comps);
    if (has_lp(wp[p].bands)) {
      hw.flag(false);
      hw.flag(hd.has_lp);
      if (hd.has_lp) write_qp_set(hw, hd.lp.data(), hd.comps);
      if (has_hp(wp[p].bands)) {
        hw.flag(false);
        hw.flag(hd.has_hp);
        if (hd.has_hp) write_qp_set(hw, hd.hp.data(), hd.comps);
      }
    }
    hw.align();
  }

  // packets
  const size_t per = frequency ? (size_t)(4 - wp[0].hdr.bands_code) : 1;
  std::vector<std::vector<u8>> packets(ntiles * per);
  auto write_tile = [&](size_t t) {
    const u32 tw = tm.cols[t % tm.cols.size()], th = tm.rows[t / tm.cols.size()];
    std::vector<std::unique_ptr<PlaneCoder>> coders;
    for (int p = 0; p < np; p++) {
      coders.emplace_back(new PlaneCoder(&wp[p].hdr, wp[p].bands, (int)tw, (int)th));
      PlaneCoder &pc = *coders.back();
      const TileSets &ts = sets[t][p];
      const int comps = wp[p].hdr.comps;
      pc.q.comps = comps;
      pc.q.dc = ts.dc;
      pc.q.lp.clear();
      for (auto &s : ts.lp) pc.q.lp.insert(pc.q.lp.end(), s.begin(), s.end());
      pc.q.hp.clear();
      for (auto &s : ts.hp) pc.q.hp.insert(pc.q.hp.end(), s.begin(), s.end());
      // what a reader will hold: inheritance as the header and tile say
      pc.q.lp_inherits = !has_lp(wp[p].bands) || (!wp[p].hdr.has_lp && ts.lp.size() == 1 && ts.lp[0] == ts.dc);
      if (pc.q.lp_inherits) pc.q.lp = pc.q.dc;
      pc.q.hp_inherits = !has_hp(wp[p].bands) || (!wp[p].hdr.has_hp && ts.hp_follows_lp);
      if (pc.q.hp_inherits) pc.q.hp = pc.q.lp;
      for (u32 m = 0; m < tw * th; m++) {
        pc.slot[m].lp_index = pc.q.lp_inherits ? 0 : ts.lp_index[m];
        pc.slot[m].hp_index = pc.q.hp_inherits ? pc.slot[m].lp_index : ts.hp_index[m];
      }
    }
    auto write_dc_header = [&](BitWriter &w, int p) {
      if (!wp[p].hdr.has_dc) write_qp_set(w, sets[t][p].dc.data(), wp[p].hdr.comps);
    };
    auto write_lp_header = [&](BitWriter &w, int p) {
      if (!has_lp(wp[p].bands) || wp[p].hdr.has_lp) return;
      const PlaneCoder &pc = *coders[p];
      w.flag(pc.q.lp_inherits);
      if (!pc.q.lp_inherits) {
        const auto &ls = sets[t][p].lp;
        w.put((u32)ls.size() - 1, 4);
        for (auto &s : ls) write_qp_set(w, s.data(), wp[p].hdr.comps);
      }
    };
    auto write_hp_header = [&](BitWriter &w, int p) {
      if (!has_hp(wp[p].bands) || wp[p].hdr.has_hp) return;
      const PlaneCoder &pc = *coders[p];
      w.flag(pc.q.hp_inherits);
      if (!pc.q.hp_inherits) {
        const auto &hs = sets[t][p].hp;
        w.put((u32)hs.size() - 1, 4);
        for (auto &s : hs) write_qp_set(w, s.data(), wp[p].hdr.comps);
      }
    };
    // macroblock m's values of plane p
    auto low_of = [&](int p, u32 m, std::vector<std::array<i32, 16>> &lows) {
      const int comps = wp[p].hdr.comps;
      lows.assign((size_t)comps, {});
      for (int c = 0; c < comps; c++) {
        const jxrhip_coefficient_plane &pl = wp[p].src->planes[wp[p].first_plane + c];
        const size_t mi = pl.macroblock_offset + tm.before[t] + m;
        const i32 *d = wp[p].src->coefficients.data() + wp[p].src->offsets[mi];
        const int blocks = wp[p].hdr.syn.blocks(c);
        lows[c].fill(0);
        if (wp[p].bands == JXRHIP_BANDS_DC_ONLY)
          lows[c][0] = d[0];
        else if (wp[p].bands == JXRHIP_BANDS_NO_HIGH_PASS)
          for (int b = 0; b < blocks; b++) lows[c][b] = d[b];
        else
          for (int b = 0; b < blocks; b++) lows[c][b] = d[b * 16];
      }
    };
    auto hp_of = [&](int p, u32 m, const int32_t **ptrs) {
      for (int c = 0; c < wp[p].hdr.comps; c++) {
        const jxrhip_coefficient_plane &pl = wp[p].src->planes[wp[p].first_plane + c];
        const size_t mi = pl.macroblock_offset + tm.before[t] + m;
        ptrs[c] = wp[p].src->coefficients.data() + wp[p].src->offsets[mi];
      }
    };
    std::vector<std::array<i32, 16>> lows;
    auto lowp = [&](const i32 **lp) {
      for (size_t c = 0; c < lows.size(); c++) lp[c] = lows[c].data();
    };
    auto start = [](BitWriter &w) {
      w.put(1, 24);
      w.put(0, 8);
    };
    if (!frequency) {
      BitWriter w;
      start(w);
      if (trim_present) w.put(trim, 4);
      for (int p = 0; p < np; p++) {
        write_dc_header(w, p);
        write_lp_header(w, p);
        write_hp_header(w, p);
      }
      for (u32 y = 0; y < th; y++)
        for (u32 x = 0; x < tw; x++) {
          const u32 m = y * tw + x;
          for (int p = 0; p < np; p++) {
            PlaneCoder &pc = *coders[p];
            pc.write_lp_index(w, (int)m);
            pc.write_hp_index(w, (int)m);
            low_of(p, m, lows);
            const i32 *lp[17];
            lowp(lp);
            pc.write_dc(w, (int)x, (int)y, lp);
            if (has_lp(pc.bands)) pc.write_lp(w, (int)x, (int)y, lp);
            if (has_hp(pc.bands)) {
              const int32_t *hp[17];
              hp_of(p, m, hp);
              pc.write_hp(w, (int)x, (int)y, hp, has_flex(pc.bands), trim);
            }
          }
        }
      packets[t] = std::move(w.bytes());
    } else {
      BitWriter dc, lp, hpw, fx;
      start(dc);
      for (int p = 0; p < np; p++) write_dc_header(dc, p);
      for (u32 y = 0; y < th; y++)
        for (u32 x = 0; x < tw; x++)
          for (int p = 0; p < np; p++) {
            low_of(p, y * tw + x, lows);
            const i32 *l[17];
            lowp(l);
            coders[p]->write_dc(dc, (int)x, (int)y, l);
          }
      packets[t * per] = std::move(dc.bytes());
      if (per >= 2) {
        start(lp);
        for (int p = 0; p < np; p++) write_lp_header(lp, p);
        for (u32 y = 0; y < th; y++)
          for (u32 x = 0; x < tw; x++)
            for (int p = 0; p < np; p++)
              if (has_lp(wp[p].bands)) {
                const u32 m = y * tw + x;
                coders[p]->write_lp_index(lp, (int)m);
                low_of(p, m, lows);
                const i32 *l[17];
                lowp(l);
                coders[p]->write_lp(lp, (int)x, (int)y, l);
              }
        packets[t * per + 1] = std::move(lp.bytes());
      }
      if (per >= 3) {
        start(hpw);
        for (int p = 0; p < np; p++) write_hp_header(hpw, p);
        for (u32 y = 0; y < th; y++)
          for (u32 x = 0; x < tw; x++)
            for (int p = 0; p < np; p++)
              if (has_hp(wp[p].bands)) {
                const u32 m = y * tw + x;
                coders[p]->write_hp_index(hpw, (int)m);
                const int32_t *hp[17];
                hp_of(p, m, hp);
                coders[p]->write_hp(hpw, (int)x, (int)y, hp, false, 0);
              }
        packets[t * per + 2] = std::move(hpw.bytes());
      }
      if (per >= 4) {
        start(fx);
        if (trim_present) fx.put(trim, 4);
        for (u32 m = 0; m < tw * th; m++)
          for (int p = 0; p < np; p++)
            if (has_flex(wp[p].bands)) {
              const int32_t *hp[17];
              hp_of(p, m, hp);
              coders[p]->write_flex_packet(fx, (int)m, hp, trim);
            }
        packets[t * per + 3] = std::move(fx.bytes());
      }
    }
  };
  {
    const u32 nthreads = std::min<u32>(std::min<u32>(std::max(1u, std::thread::hardware_concurrency()), 16u), (u32)ntiles);
    std::atomic<size_t> next{0};
    std::vector<Failure> per_tile(ntiles);
    std::vector<char> failed(ntiles, 0);
    auto run = [&]() {
      for (size_t i; (i = next.fetch_add(1)) < ntiles;) {
        try {
          write_tile(i);
        } catch (const Failure &f) {
          per_tile[i] = f;
          failed[i] = 1;
        }
      }
    };
    if (nthreads <= 1) {
      run();
    } else {
      std::vector<std::thread> pool;
      for (u32 k = 0; k < nthreads; k++) pool.emplace_back(run);
      for (auto &th : pool) th.join();
    }
    for (size_t i = 0; i < ntiles; i++)
      if (failed[i]) throw per_tile[i];
  }

  std::vector<u8> out = std::move(hw.bytes());
  if (index) {
    out.push_back(0);
    out.push_back(1);
    u64 off = 0;
    for (size_t i = 0; i < packets.size(); i++) {
      write_vlw(out, off);
      off += packets[i].size();
    }
  }
  write_vlw(out, 4);
  out.insert(out.end(), {(u8)opt.profile_idc, (u8)opt.level_idc, 0, 1});
  for (auto &pk : packets) out.insert(out.end(), pk.begin(), pk.end());
  return out;
}

void put_le16(std::vector<u8> &o, u32 v) {
  o.push_back((u8)v);
  o.push_back((u8)(v >> 8));
}
void put_le32(std::vector<u8> &o, u32 v) {
  for (int s = 0; s < 32; s += 8) o.push_back((u8)(v >> s));
}

// Annex-A container holding `primary` and, when present, a separate alpha
// codestream (annex_a/writer.rs): the tag directory, the pixel format GUID,
// then each codestream.
std::vector<u8> annex_a(u32 width, u32 height, const u8 guid[16], const std::vector<u8> &primary,
                        const std::vector<u8> *alpha) {
  const u32 entries = alpha ? 7 : 5;
  const u32 ifd = 8, ifd_bytes = 2 + entries * 12 + 4;
  const u32 guid_at = ifd + ifd_bytes;
  u32 image_at = guid_at + 16;
  image_at += image_at & 1;
  u32 alpha_at = image_at + (u32)primary.size();
  alpha_at += alpha_at & 1;
  std::vector<u8> o = {'I', 'I', 0xBC, 1};
  put_le32(o, ifd);
  put_le16(o, entries);
  auto entry = [&](u32 tag, u32 type, u32 count, u32 value) {
    put_le16(o, tag);
    put_le16(o, type);
    put_le32(o, count);
    put_le32(o, value);
  };
  entry(0xBC01, 1, 16, guid_at);
  entry(0xBC80, 4, 1, width);
  entry(0xBC81, 4, 1, height);
  entry(0xBCC0, 4, 1, image_at);
  entry(0xBCC1, 4, 1, (u32)primary.size());
  if (alpha) {
    entry(0xBCC2, 4, 1, alpha_at);
    entry(0xBCC3, 4, 1, (u32)alpha->size());
  }
  put_le32(o, 0);
  o.insert(o.end(), guid, guid + 16);
  o.resize(image_at, 0);
  o.insert(o.end(), primary.begin(), primary.end());
  if (alpha) {
    o.resize(alpha_at, 0);
    o.insert(o.end(), alpha->begin(), alpha->end());
  }
  return o;
}

// WIC pixel-format GUID of an encoder input (Annex A Table A.6)
void input_guid(const jxrhip_encode_params &p, bool alpha_in_primary, u8 guid[16]) {
  static const u8 prefix[15] = {0x24, 0xc3, 0xdd, 0x6f, 0x03, 0x4e, 0xfe, 0x4b, 0xb1, 0x85, 0x3d, 0x77, 0x76, 0x8d, 0xc9};
  std::memcpy(guid, prefix, 15);
  const bool gray = p.input_channels <= 2;
  const bool four = p.input_channels == 4 && alpha_in_primary;
  u8 code;
  if (p.input_kind == JXRHIP_INPUT_U8)
    code = gray ? 0x08 : four ? 0x0f : 0x0d;
  else if (p.input_kind == JXRHIP_INPUT_U16)
    code = gray ? 0x0b : four ? 0x16 : 0x15;
  else
    code = gray ? 0x11 : four ? 0x19 : 0x1b;
  guid[15] = code;
}

jxrhip_status fail_with(const Failure &f) {
  jxrhip_set_error(f.message.c_str());
  return f.status;
}

}  // namespace

// ---- C ABI -------------------------------------------------------------------------
extern "C" jxrhip_status jxrhip_codestream_parse(const uint8_t *bytes, uint64_t length, jxrhip_codestream **out) {
  if (!bytes || !out) {
    jxrhip_set_error("null argument");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  try {
    std::unique_ptr<jxrhip_codestream> cs(new jxrhip_codestream());
    cs->bytes.assign(bytes, bytes + length);
    const u8 *b = cs->bytes.data();
    const size_t n = cs->bytes.size();
    if (n >= 8 && std::memcmp(b, "WMPHOTO\0", 8) == 0) {
      cs->primary = parse_stream(b, n);
    } else {
      cs->container = parse_annex_a(b, n);
      cs->annex = true;
      cs->primary = parse_stream(b + cs->container.offset, cs->container.bytes);
      if (cs->primary.h.image.width != cs->container.width || cs->primary.h.image.height != cs->container.height)
        syntax_error("Annex-A and codestream dimensions differ");
      if (cs->container.alpha) {
        cs->alpha = parse_stream(b + cs->container.alpha_offset, cs->container.alpha_bytes);
        cs->has_separate = true;
        if (cs->alpha.h.image.width != cs->container.width || cs->alpha.h.image.height != cs->container.height)
          syntax_error("separate alpha dimensions");
      }
    }
    *out = cs.release();
    return JXRHIP_OK;
  } catch (const Failure &f) {
    return fail_with(f);
  } catch (const std::bad_alloc &) {
    jxrhip_set_error("out of host memory");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
}

extern "C" void jxrhip_codestream_destroy(jxrhip_codestream *cs) { delete cs; }

extern "C" jxrhip_status jxrhip_codestream_get_info(const jxrhip_codestream *cs, jxrhip_codestream_info *o) {
  if (!cs || !o) {
    jxrhip_set_error("null argument");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  std::memset(o, 0, sizeof *o);
  const Headers &h = cs->primary.h;
  const ImageHeader &im = h.image;
  o->width = im.width;
  o->height = im.height;
  o->margin_top = im.margins[0];
  o->margin_left = im.margins[1];
  o->margin_bottom = im.margins[2];
  o->margin_right = im.margins[3];
  u32 mx = 0, my = 0;
  for (u32 v : cs->primary.cols) mx += v;
  for (u32 v : cs->primary.rows) my += v;
  o->macroblocks_x = mx;
  o->macroblocks_y = my;
  o->tile_column_count = (u32)cs->primary.cols.size();
  o->tile_row_count = (u32)cs->primary.rows.size();
  o->overlap = im.overlap;
  o->hard_tiles = im.hard;
  o->frequency_mode = im.frequency;
  o->orientation = cs->annex && cs->container.has_orientation ? cs->container.orientation : im.orientation;
  o->output_color = color_of_output_code(im.out_color);
  o->bit_depth = depth_of_code(im.out_depth);
  o->color_format = color_of_internal_code(h.primary.internal);
  o->component_count = (u32)h.primary.comps;
  o->bands = bands_of_code(h.primary.bands_code);
  o->scaled = h.primary.scaled;
  o->chroma_centering_x = h.primary.cx;
  o->chroma_centering_y = h.primary.cy;
  o->shift_bits = h.primary.shift_bits;
  o->mantissa_length = h.primary.mantissa;
  o->exponent_bias = h.primary.exponent_bias;
  const PlaneHeader *alpha = nullptr;
  if (im.alpha_plane) {
    o->alpha_mode = JXRHIP_ALPHA_INTEGRATED;
    alpha = &h.alpha;
    o->alpha_bit_depth = o->bit_depth;
  } else if (cs->has_separate) {
    o->alpha_mode = JXRHIP_ALPHA_SEPARATE;
    alpha = cs->alpha.h.image.alpha_plane ? &cs->alpha.h.alpha : &cs->alpha.h.primary;
    o->alpha_bit_depth = depth_of_code(cs->alpha.h.image.out_depth);
  }
  if (alpha) {
    o->alpha_bands = bands_of_code(alpha->bands_code);
    o->alpha_scaled = alpha->scaled;
    o->alpha_shift_bits = alpha->shift_bits;
    o->alpha_mantissa_length = alpha->mantissa;
    o->alpha_exponent_bias = alpha->exponent_bias;
  }
  o->red_blue_not_swapped = im.rb_not_swapped;
  o->premultiplied_alpha = im.premultiplied;
  o->short_header = im.short_header;
  o->long_word = im.long_word;
  o->trim_flexbits_present = im.trim;
  o->index_table_present = im.index;
  o->profile_idc = cs->primary.dir.profile;
  o->level_idc = cs->primary.dir.level;
  o->annex_a = cs->annex;
  if (cs->annex) std::memcpy(o->pixel_format_guid, cs->container.guid, 16);
  o->codestream_offset = cs->annex ? cs->container.offset : 0;
  o->codestream_bytes = cs->annex ? cs->container.bytes : cs->bytes.size();
  o->alpha_offset = cs->has_separate ? cs->container.alpha_offset : 0;
  o->alpha_bytes = cs->has_separate ? cs->container.alpha_bytes : 0;
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_codestream_tile_grid(const jxrhip_codestream *cs, uint32_t *cols, uint32_t *rows) {
  if (!cs || !cols || !rows) {
    jxrhip_set_error("null argument");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  std::copy(cs->primary.cols.begin(), cs->primary.cols.end(), cols);
  std::copy(cs->primary.rows.begin(), cs->primary.rows.end(), rows);
  return JXRHIP_OK;
}

extern "C" jxrhip_status jxrhip_codestream_decode(const jxrhip_codestream *cs, uint32_t scale, const uint8_t *tile_mask,
                                                  uint32_t threads, jxrhip_decoded **out) {
  if (!cs || !out) {
    jxrhip_set_error("null argument");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  if (scale > JXRHIP_SCALE_SIXTEENTH) {
    jxrhip_set_error("decode scale");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  try {
    std::unique_ptr<jxrhip_decoded> d(new jxrhip_decoded());
    decode_stream(cs->primary, scale, tile_mask, threads, d->primary);
    if (cs->has_separate) {
      if (cs->alpha.cols != cs->primary.cols || cs->alpha.rows != cs->primary.rows)
        unsupported("separate alpha tile grid differs from the primary's");
      decode_stream(cs->alpha, scale, tile_mask, threads, d->alpha);
      d->has_alpha = true;
    }
    *out = d.release();
    return JXRHIP_OK;
  } catch (const Failure &f) {
    return fail_with(f);
  } catch (const std::bad_alloc &) {
    jxrhip_set_error("out of host memory");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
}

extern "C" jxrhip_status jxrhip_codestream_coefficient_count(const jxrhip_codestream *cs, uint32_t scale,
                                                             const uint8_t *tile_mask, uint64_t *primary,
                                                             uint64_t *alpha) {
  if (!cs || !primary || !alpha) {
    jxrhip_set_error("null argument");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  *primary = *alpha = 0;
  if (scale > JXRHIP_SCALE_SIXTEENTH) {
    jxrhip_set_error("decode scale");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  try {
    jxr::ArenaOut scratch;
    decode_stream(cs->primary, scale, tile_mask, 1, scratch, nullptr, 0, primary);
    if (cs->has_separate) decode_stream(cs->alpha, scale, tile_mask, 1, scratch, nullptr, 0, alpha);
    return JXRHIP_OK;
  } catch (const Failure &f) {
    return fail_with(f);
  } catch (const std::bad_alloc &) {
    jxrhip_set_error("out of host memory");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
}

extern "C" jxrhip_status jxrhip_codestream_decode_into(const jxrhip_codestream *cs, uint32_t scale,
                                                       const uint8_t *tile_mask, uint32_t threads, int32_t *primary,
                                                       uint64_t primary_capacity, int32_t *alpha,
                                                       uint64_t alpha_capacity, jxrhip_decoded **out) {
  if (!cs || !out || !primary || (cs->has_separate && !alpha)) {
    jxrhip_set_error("null argument");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  if (scale > JXRHIP_SCALE_SIXTEENTH) {
    jxrhip_set_error("decode scale");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  try {
    std::unique_ptr<jxrhip_decoded> d(new jxrhip_decoded());
    decode_stream(cs->primary, scale, tile_mask, threads, d->primary, primary, primary_capacity);
    if (cs->has_separate) {
      if (cs->alpha.cols != cs->primary.cols || cs->alpha.rows != cs->primary.rows)
        unsupported("separate alpha tile grid differs from the primary's");
      decode_stream(cs->alpha, scale, tile_mask, threads, d->alpha, alpha, alpha_capacity);
      d->has_alpha = true;
    }
    *out = d.release();
    return JXRHIP_OK;
  } catch (const Failure &f) {
    return fail_with(f);
  } catch (const std::bad_alloc &) {
    jxrhip_set_error("out of host memory");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
}

extern "C" void jxrhip_decoded_destroy(jxrhip_decoded *d) { delete d; }

extern "C" const jxrhip_arena *jxrhip_decoded_arena(const jxrhip_decoded *d, uint32_t which) {
  if (!d) return nullptr;
  if (which == 0) return &d->primary.arena;
  if (which == 1 && d->has_alpha) return &d->alpha.arena;
  return nullptr;
}

extern "C" jxrhip_status jxrhip_encoded_codestream(const jxrhip_encoded *e, const jxrhip_codestream_options *options,
                                                   uint8_t **bytes, uint64_t *length) {
  if (!e || !bytes || !length) {
    jxrhip_set_error("null argument");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
  *bytes = nullptr;
  *length = 0;
  jxrhip_codestream_options opt{};
  opt.trim_flexbits = JXRHIP_TRIM_NONE;
  opt.profile_idc = 111;
  opt.level_idc = 255;
  if (options) opt = *options;
  try {
    const jxrhip_encode_params &p = e->params;
    const bool separate = e->has_alpha;
    const bool has_alpha = p.input_channels == 2 || p.input_channels == 4;
    const bool integrated = has_alpha && p.alpha_mode == JXRHIP_ALPHA_INTEGRATED;
    if (separate && !opt.annex_a) unsupported("a separate alpha plane needs an Annex-A container");
    TileMap tm;
    tm.cols = e->primary.tile_w;
    tm.rows = e->primary.tile_h;
    size_t acc = 0;
    for (u32 r : tm.rows)
      for (u32 c : tm.cols) {
        tm.before.push_back(acc);
        acc += (size_t)r * c;
      }
    const u32 bands_code = 3 - p.bands;
    auto plane = [&](u32 internal, int comps, const jxr::ArenaOut *src, int first) {
      WritePlane w;
      w.hdr.internal = internal;
      w.hdr.scaled = p.scaled != 0;
      w.hdr.bands_code = bands_code;
      w.hdr.comps = comps;
      w.hdr.shift_bits = p.input_kind == JXRHIP_INPUT_U16 ? p.shift_bits : 0;
      w.hdr.mantissa = p.mantissa_length;
      w.hdr.exponent_bias = p.exponent_bias;
      syntax_of(internal, (u32)comps, w.hdr.syn);
      w.src = src;
      w.first_plane = first;
      w.bands = p.bands;
      return w;
    };
    const u32 internal = p.color_format == JXRHIP_COLOR_LUMA ? 0 : p.color_format == JXRHIP_COLOR_YUV420 ? 1
                         : p.color_format == JXRHIP_COLOR_YUV422 ? 2 : 3;
    const int comps = internal == 0 ? 1 : 3;
    const u32 out_depth = p.input_kind == JXRHIP_INPUT_U8 ? 1 : p.input_kind == JXRHIP_INPUT_U16 ? 2 : 7;
    const u32 out_color = comps == 1 ? 0 : 7;
    std::vector<WritePlane> wp = {plane(internal, comps, &e->primary, 0)};
    if (integrated) wp.push_back(plane(0, 1, &e->primary, comps));
    std::vector<u8> primary = write_stream(p, wp, tm, out_color, out_depth, opt, false);
    std::vector<u8> file;
    if (opt.annex_a) {
      u8 guid[16];
      input_guid(p, integrated, guid);
      if (separate) {
        std::vector<WritePlane> ap = {plane(0, 1, &e->alpha, 0)};
        std::vector<u8> alpha = write_stream(p, ap, tm, 0, out_depth, opt, true);
        file = annex_a(p.width, p.height, guid, primary, &alpha);
      } else {
        file = annex_a(p.width, p.height, guid, primary, nullptr);
      }
    } else {
      file = std::move(primary);
    }
    uint8_t *buf = (uint8_t *)std::malloc(file.size() ? file.size() : 1);
    if (!buf) throw std::bad_alloc();
    std::memcpy(buf, file.data(), file.size());
    *bytes = buf;
    *length = file.size();
    return JXRHIP_OK;
  } catch (const Failure &f) {
    return fail_with(f);
  } catch (const std::bad_alloc &) {
    jxrhip_set_error("out of host memory");
    return JXRHIP_ERR_INVALID_ARGUMENT;
  }
}

extern "C" void jxrhip_free_bytes(uint8_t *bytes) { std::free(bytes); }
