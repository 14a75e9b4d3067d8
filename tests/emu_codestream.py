"""Run by tests/test_emulated.py in a process of its own, on the CPU build
of the library (optionally under AddressSanitizer): writer/reader round
trips over the encoder's layouts, the reference's T.835 file, and damaged
codestreams (truncated, bit-flipped, bytes overwritten), each of which must
either decode or fail with a status, never read or write out of bounds, half
of them decoded into caller storage of exactly the counted size."""
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import emu_lib  # noqa: E402


def main():
    asan = "--asan" in sys.argv
    emu_lib.load(asan)
    import numpy as np
    from codestream_util import GOLDEN, arena_arrays, arena_differences
    from jxr_amd import JxrHipError, abi
    from jxr_amd.codestream import Codestream, encode_codestream
    import test_codestream as T

    failures = 0
    streams = []
    for case in T.CASES:
        px, enc = T.make(case)
        data = encode_codestream(enc, frequency=case[8], annex_a=case[9] == abi.ALPHA_SEPARATE)
        got = Codestream(data).decode(threads=3)
        diff = arena_differences(arena_arrays(enc.primary), arena_arrays(got.primary), case[6])
        if case[6] == abi.BANDS_NO_FLEXBITS:
            diff = [d for d in diff if d != "coefficients"]
        if diff:
            print("round trip", case[0], diff, flush=True)
            failures += 1
        streams.append(data)
    streams.append(bytes.fromhex(GOLDEN["scaled_three_component"]["hex"]))
    rng = random.Random(20261204)
    decoded = rejected = 0
    for data in streams:
        for trial in range(40):
            b = bytearray(data)
            kind = trial % 4
            if kind == 0:
                b = b[: rng.randrange(1, len(b))]
            elif kind == 1:
                for _ in range(rng.randrange(1, 8)):
                    i = rng.randrange(len(b))
                    b[i] ^= 1 << rng.randrange(8)
            elif kind == 2:
                i = rng.randrange(len(b))
                for k in range(i, min(len(b), i + rng.randrange(1, 16))):
                    b[k] = rng.randrange(256)
            else:
                # damage past the headers only
                start = min(len(b) - 1, 40)
                i = rng.randrange(start, len(b))
                b[i] = rng.randrange(256)
            try:
                cs = Codestream(bytes(b))
                if trial % 2:
                    cs.decode(threads=2)
                else:
                    # into caller storage of exactly the counted size (a
                    # heap block of its own, so a write past it is caught)
                    n, na = cs.coefficient_count()
                    store = np.empty(n + na, np.int32)
                    keep = (store, store.ctypes.data, store.nbytes)

                    class Storage:
                        ptr, nbytes = keep[1], keep[2]
                    cs.decode(threads=2, into=Storage)
                decoded += 1
            except JxrHipError as e:
                if e.code not in (abi.ERR_SYNTAX, abi.ERR_UNSUPPORTED, abi.ERR_OVERFLOW):
                    print("unexpected status", e.code, e.message, flush=True)
                    failures += 1
                rejected += 1
    print(f"damaged codestreams: {decoded} decoded, {rejected} refused", flush=True)
    print(f"{failures} failures", flush=True)


if __name__ == "__main__":
    main()
