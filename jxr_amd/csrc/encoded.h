// encoded.h — what the forward analysis keeps of one encoded image: the
// coefficient arenas it fills, and the facts the codestream writer needs
// beside them (the encode parameters and each macroblock's coded
// quantization parameters). Shared by analysis.cpp, which fills it, and
// codestream.cpp, which codes it.
#pragma once

#include <cstdint>
#include <memory>
#include <utility>
#include <vector>

#include "../../include/jxr_hip_encode.h"

namespace jxr {

// An allocator whose resize leaves new elements uninitialised: a decode
// writes every coefficient of its arena, so zeroing it first would only
// cost a pass over memory.
template <class T>
struct UninitAllocator : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = UninitAllocator<U>;
  };
  UninitAllocator() = default;
  template <class U>
  UninitAllocator(const UninitAllocator<U> &) {}
  template <class U>
  void construct(U *p) noexcept {
    ::new (static_cast<void *>(p)) U;
  }
  template <class U, class... Args>
  void construct(U *p, Args &&...args) {
    ::new (static_cast<void *>(p)) U(std::forward<Args>(args)...);
  }
};

struct ArenaOut {
  std::vector<int32_t, UninitAllocator<int32_t>> coefficients;
  std::vector<uint32_t> offsets, quantizers, coded_x, coded_y;
  std::vector<uint8_t> bands, preds;
  // [macroblock][3]: the DC, LP and HP quantization parameters (0..255, 0
  // lossless) whose steps `quantizers` holds
  std::vector<uint8_t> qps;
  std::vector<jxrhip_coefficient_plane> planes;
  std::vector<uint32_t> tile_w, tile_h;
  jxrhip_arena arena{};
};

}  // namespace jxr

struct jxrhip_encoded {
  jxr::ArenaOut primary;
  jxr::ArenaOut alpha;
  bool has_alpha = false;
  // the request, with its tile arrays (kept in primary.tile_w / tile_h)
  // cleared
  jxrhip_encode_params params{};
};
