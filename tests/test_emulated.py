"""The HIP kernel and runtime sources, built for the CPU against the host
emulation in tests/emu and run under AddressSanitizer: every parity case
matches the oracle byte for byte at several tile shapes, and no access
leaves its LDS region or device buffer."""
import os
import subprocess
import sys

import emu_lib

HERE = os.path.dirname(os.path.abspath(__file__))


def test_emulated_hip_path_matches_oracle_under_asan():
    emu_lib.build(asan=True)
    env = dict(os.environ, LD_PRELOAD=emu_lib.asan_runtime(), ASAN_OPTIONS="detect_leaks=0")
    res = subprocess.run([sys.executable, os.path.join(HERE, "emu_check.py"), "--asan", "--tiles"],
                         capture_output=True, text=True, env=env, timeout=900)
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-3000:]
    assert "0 failures" in res.stdout


def test_emulated_session_lifetime_and_reuploads_under_asan():
    """Owners that outlive the session handle, and serving re-uploads of
    evicted arenas over two compute streams (tests/emu_lifetime.py)."""
    emu_lib.build(asan=True)
    env = dict(os.environ, LD_PRELOAD=emu_lib.asan_runtime(), ASAN_OPTIONS="detect_leaks=0")
    res = subprocess.run([sys.executable, os.path.join(HERE, "emu_lifetime.py"), "--asan"],
                         capture_output=True, text=True, env=env, timeout=600)
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-3000:]
    assert "0 failures" in res.stdout


def test_emulated_allocation_free_batches_under_asan():
    """jxrhip_decode_batch_static on the CPU build (tests/emu_no_alloc.py)."""
    emu_lib.build(asan=True)
    env = dict(os.environ, LD_PRELOAD=emu_lib.asan_runtime(), ASAN_OPTIONS="detect_leaks=0")
    res = subprocess.run([sys.executable, os.path.join(HERE, "emu_no_alloc.py"), "--asan"],
                         capture_output=True, text=True, env=env, timeout=600)
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-3000:]
    assert "0 failures" in res.stdout


def test_emulated_codestream_round_trips_and_damage_under_asan():
    """The codestream reader and writer on the CPU build: round trips, the
    reference's T.835 file, and damaged codestreams that must decode or be
    refused without an out-of-bounds access (tests/emu_codestream.py)."""
    emu_lib.build(asan=True)
    env = dict(os.environ, LD_PRELOAD=emu_lib.asan_runtime(), ASAN_OPTIONS="detect_leaks=0")
    res = subprocess.run([sys.executable, os.path.join(HERE, "emu_codestream.py"), "--asan"],
                         capture_output=True, text=True, env=env, timeout=600)
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-3000:]
    assert "0 failures" in res.stdout
