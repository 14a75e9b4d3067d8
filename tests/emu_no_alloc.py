"""The allocation-free batch entry point on the CPU build of the runtime
(run as a script under AddressSanitizer by tests/test_emulated.py): the
checks of tests/test_zz_no_alloc.py, through the emulated HIP runtime."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import emu_lib  # noqa: E402


def main():
    emu_lib.load("--asan" in sys.argv)
    import test_zz_no_alloc as T
    from jxr_amd import HipDecoderSession

    s = HipDecoderSession(0)
    T.test_two_calls_in_a_row_equal_the_persistent_batch(s)
    T.test_refuses_uploads_that_are_not_resident(s)
    T.test_refuses_batches_beyond_its_limits(s)
    s.close()
    print("emulated no-alloc: 0 failures", flush=True)


if __name__ == "__main__":
    main()
