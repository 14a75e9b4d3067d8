// encoded.h — what the forward analysis keeps of one encoded image: the
// coefficient arenas it fills, and the facts the codestream writer needs
// beside them (the encode parameters and each macroblock's coded
// quantization parameters). Shared by analysis.cpp, which fills it, and
// codestream.cpp, which codes it.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/jxr_hip_encode.h"

namespace jxr {

struct ArenaOut {
  std::vector<int32_t> coefficients;
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
