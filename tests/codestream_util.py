"""Codestream fixtures built bit by bit from the syntax, as the reference's
tests build theirs (crates/jxr/tests/common/mod.rs:11-46), and helpers that
view arenas as numpy arrays."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "reference_codestreams.json")))


def codestream(output: int, fields, tail: bytes) -> bytes:
    """A 16x16 short-header codestream: the image header ending in the byte
    `output` (output colour format and bit depth), the plane header fields
    [(value, bits)] padded to a byte, then `tail` (index and packets)."""
    b = bytearray(b"WMPHOTO\0") + bytes([0x11, 0x00, 0x80, output]) + (15).to_bytes(2, "big") * 2
    pos = len(b) * 8
    for value, count in fields:
        for shift in range(count - 1, -1, -1):
            if pos // 8 == len(b):
                b.append(0)
            b[pos // 8] |= ((value >> shift) & 1) << (7 - pos % 8)
            pos += 1
    return bytes(b) + bytes(tail)


def dc_only_luma_codestream(scaled: bool, macroblock) -> bytes:
    """16x16 Y-only, DC-only codestream with a uniform DC quantizer of one;
    `macroblock` is the tile's entropy-coded payload."""
    fields = [(0, 3), (1 if scaled else 0, 1), (3, 4), (1, 1), (1, 8)]
    return codestream(0x01, fields, bytes([0xFD, 0, 0, 1, 0x5A]) + bytes(macroblock))


def minimal_header() -> bytes:
    """Y-only 8-bit, all bands, uniform DC/LP/HP quantizers of one
    (crates/jxr-native/tests/header_parse.rs:5-26)."""
    return codestream(0x01, [(0, 3), (0, 1), (0, 4), (1, 1), (1, 8), (0, 1), (1, 1), (1, 8), (0, 1), (1, 1),
                             (1, 8)], b"")


def bits_to_bytes(bits: str) -> bytes:
    bits = bits + "0" * (-len(bits) % 8)
    return bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))


def arena_arrays(arena) -> dict:
    """Copies of every array of a jxrhip_arena, keyed by field."""
    a = arena.contents

    def arr(p, n, dt):
        return np.ctypeslib.as_array(p, shape=(n,)).astype(dt) if n else np.zeros(0, dt)

    n, m = a.coefficient_count, a.macroblock_count
    return {
        "coefficients": arr(a.coefficients, n, np.int64),
        "offsets": arr(a.coefficient_offsets, m, np.int64),
        "quantizers": arr(a.quantizers, 3 * m, np.int64).reshape(-1, 3),
        "bands": arr(a.bands, m, np.int64),
        "hp_predictions": arr(a.hp_predictions, m, np.int64),
        "coded_x": arr(a.coded_x, m, np.int64),
        "coded_y": arr(a.coded_y, m, np.int64),
        "planes": [(a.planes[i].coefficient_offset, a.planes[i].coefficient_count, a.planes[i].macroblock_offset,
                    a.planes[i].macroblock_count, a.planes[i].block_columns, a.planes[i].block_rows)
                   for i in range(a.plane_count)],
        "overlap": a.overlap,
        "hard_tiles": a.hard_tiles,
        "tile_columns": [a.tile_column_widths[i] for i in range(a.tile_column_count)],
        "tile_rows": [a.tile_row_heights[i] for i in range(a.tile_row_count)],
    }


def arena_differences(want: dict, got: dict, coded_steps_only: int = -1) -> list:
    """Fields where two arenas differ. With coded_steps_only set to the
    arena's band set, the quantizer steps of bands not coded are not compared
    (a decoder gives them what the packet layout leaves: the band below's
    parameters in a spatial packet, the DC's in a frequency one)."""
    diffs = []
    for k, v in want.items():
        g = got[k]
        if k == "quantizers" and coded_steps_only >= 0:
            used = 1 + (coded_steps_only >= 1) + (coded_steps_only >= 2)
            v, g = v[:, :used], g[:, :used]
        if isinstance(v, np.ndarray):
            if v.shape != g.shape or not np.array_equal(v, g):
                diffs.append(k)
        elif v != g:
            diffs.append(k)
    return diffs
