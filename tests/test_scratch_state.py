"""The radix-sort scratch bookkeeping of a pipeline (gs::ScratchState), driven
through random sequences of frames against a word-level model of the scratch.

A look-back sort reads its words [0, L) and needs them all zero: it adds digit
histograms with atomics and takes any status word with flag bits as a
published prefix. The frame clears [0, L) after the sort. A reduce-then-scan
sort leaves digit counts and flagged chunk status words in [0, R). The model
tracks every word; the state machine must never let a look-back sort start on
a non-zero word, over frames of mixed modes and sizes (the way a batch or a
context alternates between large and small views)."""

import ctypes
import pathlib
import subprocess

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
D, P, FIXED = 256, 4, 8 * 5 * 256 + 5  # digits, passes, histograms + tile counters


def lookback_words(tiles):
    return FIXED + P * tiles * D


def rts_words(tiles):
    entries = D * tiles
    return entries + 2 + -(-entries // 8192)


@pytest.fixture(scope="module")
def ss(tmp_path_factory):
    out = tmp_path_factory.mktemp("ss") / "libss.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-I", str(ROOT / "gsplat-rs_amd/csrc"),
                    str(ROOT / "tests/native/scratch_state_check.cpp"), "-o", str(out)], check=True)
    lib = ctypes.CDLL(str(out))
    lib.ss_new.restype = ctypes.c_void_p
    lib.ss_free.argtypes = [ctypes.c_void_p]
    lib.ss_before.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_ulonglong),
                              ctypes.POINTER(ctypes.c_int)]
    lib.ss_after.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_ulonglong]
    lib.ss_reset.argtypes = [ctypes.c_void_p]
    return lib


class Native:
    def __init__(self, lib):
        self.lib, self.h = lib, lib.ss_new()

    def before(self, lookback, lb):
        fill, skip = ctypes.c_ulonglong(), ctypes.c_int()
        self.lib.ss_before(self.h, int(lookback), lb, ctypes.byref(fill), ctypes.byref(skip))
        return fill.value, bool(skip.value)

    def after(self, lookback, rts):
        self.lib.ss_after(self.h, int(lookback), rts)


class OldProtocol:
    """The earlier bookkeeping: 'clean' iff the last frame's sort was a look-back one."""

    def __init__(self):
        self.clean = True

    def before(self, lookback, lb):
        return 0, self.clean

    def after(self, lookback, rts):
        self.clean = bool(lookback)


def run_frames(state, frames, size):
    """Replays (lookback, tiles) frames; returns the first frame whose look-back sort would
    start on a non-zero word, or None."""
    words = np.zeros(size, np.int8)  # 0 zero, 1 look-back state, 2 reduce-then-scan counts / flags
    for k, (lookback, tiles) in enumerate(frames):
        lb, rts = lookback_words(tiles), rts_words(tiles)
        fill, skip = state.before(lookback, lb)
        words[:fill] = 0
        if lookback:
            if not skip:  # the sort's own clear of the words its live tiles use
                words[:lb] = 0
            if words[:lb].any():
                return k
            words[:lb] = 1
            words[:lb] = 0  # the frame's later kernel clears the look-back words
        else:
            words[:rts] = 2
        state.after(lookback, rts)
    return None


def random_frames(rng, count):
    # look-back below 256 tiles, reduce-then-scan from 256 (the pipeline's default threshold)
    tiles = rng.choice([1, 3, 13, 49, 120, 177, 255, 256, 300, 489, 731, 1200], count)
    return [(int(t) < 256, int(t)) for t in tiles]


def test_lookback_sorts_always_start_clean(ss):
    rng = np.random.default_rng(2024)
    size = max(lookback_words(1200), rts_words(1200)) + 16
    for trial in range(300):
        frames = random_frames(rng, 40)
        st = Native(ss)
        assert run_frames(st, frames, size) is None, f"trial {trial}: {frames}"
        ss.ss_free(st.h)


def test_model_catches_the_earlier_protocol():
    # A large reduce-then-scan frame, a small look-back frame, then a larger
    # look-back frame: the earlier bookkeeping skipped the clear over stale flags.
    size = max(lookback_words(1200), rts_words(1200)) + 16
    frames = [(False, 489), (True, 13), (True, 177)]
    assert run_frames(OldProtocol(), frames, size) == 2


def test_fill_covers_rts_words_and_resets(ss):
    st = Native(ss)
    assert st.before(True, lookback_words(10)) == (0, True)
    st.after(False, rts_words(489))
    fill, skip = st.before(True, lookback_words(13))
    assert fill == max(rts_words(489), lookback_words(13)) and skip
    st.after(True, rts_words(13))
    assert st.before(True, lookback_words(177)) == (0, True)
    assert st.before(False, lookback_words(300)) == (0, False)
    st.after(False, rts_words(300))
    ss.ss_reset(st.h)
    assert st.before(True, lookback_words(5)) == (0, True)
    ss.ss_free(st.h)
