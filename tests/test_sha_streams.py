"""SHA-256 of the replay output: the library's own compression on the x86
SHA extensions, one message at a time and several interleaved on one core
(as the batch runtime hashes the outputs of a pair of jobs band by band),
against hashlib.

The reference hashes each replay's output with one hashlib.sha256() updated
row by row (src/portable_digital_ice/streaming.py:425, :550); every digest
here must equal hashlib's for the same bytes, whatever the piece sizes and
however the messages are interleaved.  No GPU is involved.
"""

from __future__ import annotations

import ctypes as C
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

from dice_amd.hip_backend import engine as hip_engine

# lengths around every block and padding boundary (55/56 bytes leave room
# for the length or not; 64 and 128 are whole blocks)
EDGE_LENGTHS = [0, 1, 3, 55, 56, 57, 63, 64, 65, 119, 120, 127, 128, 129, 1000, 4097]
INTERLEAVED, SINGLE = 0, 1


def _digests(messages: list[bytes], piece: int, mode: int) -> tuple[list[str], int]:
    arrays = [np.frombuffer(m, np.uint8) if m else np.zeros(1, np.uint8) for m in messages]
    pointers = (C.c_void_p * len(messages))(*[a.ctypes.data for a in arrays])
    lengths = (C.c_int64 * len(messages))(*[len(m) for m in messages])
    out = C.create_string_buffer(72 * len(messages))
    extensions = C.c_int32(-1)
    code = hip_engine.library().dice_t_sha256_pieces(
        pointers, lengths, len(messages), piece, mode, out, C.byref(extensions))
    assert code == 0, hip_engine.last_error()
    raw = out.raw
    return [raw[72 * i:72 * i + 64].decode() for i in range(len(messages))], extensions.value


def _random_messages(lengths, seed):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 256, n, dtype=np.uint8).tobytes() for n in lengths]


@pytest.mark.parametrize("mode", [INTERLEAVED, SINGLE], ids=["interleaved", "single"])
@pytest.mark.parametrize("piece", [1, 7, 64, 100, 4096, 1 << 20])
def test_edge_lengths_match_hashlib(mode, piece):
    messages = _random_messages(EDGE_LENGTHS, seed=piece)
    got, _ = _digests(messages, piece, mode)
    assert got == [hashlib.sha256(m).hexdigest() for m in messages]


def test_known_answers():
    """FIPS 180-4 example messages (one block, two blocks) and the empty one."""

    messages = [b"abc", b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq", b""]
    want = ["ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad",
            "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1",
            "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"]
    for mode in (INTERLEAVED, SINGLE):
        assert _digests(messages, 5, mode)[0] == want


def test_many_unequal_messages():
    """Groups of four messages of unequal length: each interleaved pass runs
    over the shortest message's blocks, then on over the rest."""

    rng = np.random.default_rng(3)
    lengths = [int(n) for n in rng.integers(0, 300_000, 37)] + [0, 64, 1 << 18]
    messages = _random_messages(lengths, seed=4)
    got, _ = _digests(messages, 65_537, INTERLEAVED)
    assert got == [hashlib.sha256(m).hexdigest() for m in messages]


def test_band_sized_pieces():
    """Two 8 MiB messages in 1 MiB pieces, as a pair of jobs hashes its
    bands (a full frame's band is 17 MB)."""

    messages = _random_messages([8 << 20, (8 << 20) + 6], seed=5)
    got, _ = _digests(messages, 1 << 20, INTERLEAVED)
    assert got == [hashlib.sha256(m).hexdigest() for m in messages]


def test_extensions_used_where_present():
    flags = open("/proc/cpuinfo").read() if os.path.exists("/proc/cpuinfo") else ""
    _, used = _digests([b"x"], 1, SINGLE)
    assert used == (1 if " sha_ni" in flags and " sse4_1" in flags else 0)


def test_without_extensions_in_a_fresh_process():
    """DICE_SHA_EXTENSIONS=0 routes every message through OpenSSL; the
    digests are the same."""

    script = (
        "import sys, hashlib\n"
        f"sys.path[:0] = {[p for p in sys.path if p]!r}\n"
        "from test_sha_streams import _digests, _random_messages, EDGE_LENGTHS\n"
        "m = _random_messages(EDGE_LENGTHS, seed=9)\n"
        "for mode in (0, 1):\n"
        "    got, used = _digests(m, 33, mode)\n"
        "    assert used == 0, used\n"
        "    assert got == [hashlib.sha256(x).hexdigest() for x in m], mode\n"
        "print('ok')\n"
    )
    env = dict(os.environ, DICE_SHA_EXTENSIONS="0")
    done = subprocess.run([sys.executable, "-c", script], env=env, capture_output=True,
                          text=True, timeout=120, cwd=os.path.dirname(__file__))
    assert done.returncode == 0, done.stderr[-2000:]
    assert done.stdout.strip().endswith("ok")


def test_bad_arguments_are_rejected():
    lib = hip_engine.library()
    out = C.create_string_buffer(72)
    one = (C.c_void_p * 1)(None)
    length = (C.c_int64 * 1)(5)
    assert lib.dice_t_sha256_pieces(one, length, 1, 1, 0, out, None) != 0  # null message
    length[0] = 0
    assert lib.dice_t_sha256_pieces(one, length, 0, 1, 0, out, None) != 0  # no messages
    assert lib.dice_t_sha256_pieces(one, length, 1, 0, 0, out, None) != 0  # empty pieces
    assert lib.dice_t_sha256_pieces(one, length, 1, 1, 2, out, None) != 0  # no such mode
    assert lib.dice_t_sha256_pieces(one, length, 1, 1, 0, out, None) == 0
    assert out.raw[:64].decode() == hashlib.sha256(b"").hexdigest()
