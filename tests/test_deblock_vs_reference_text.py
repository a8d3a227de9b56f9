"""Pins the oracle's deblocking filter bodies and sample decisions to the
reference's own expressions: each reference function is read as text from
crates/rustmpeg-video/src/vvc/filter.rs and evaluated in Python
(tests/rust_text.py), then compared with the oracle on random lines over
every filter form, both with tC large enough that no clip binds (the linear
taps alone: a wrong tap or rounding term fails even when it keeps the tap
sum) and with small tC (the clips). Skipped where the reference tree is
absent.

  long_tap_luma_line      filter.rs:477-582   orc_dbk_long_line
  short_filter_luma_line  filter.rs:586-631   orc_dbk_short_line
  chroma_strong_line      filter.rs:1010-1031 orc_dbk_chroma_line (3, 3)
  chroma_weak_line_p1q3   filter.rs:1036-1049 orc_dbk_chroma_line (1, 3)
  chroma_weak_line        filter.rs:1052-1057 orc_dbk_chroma_line (1, 1)
  luma_sample_decision    filter.rs:276-304
  chroma_sample_decision  filter.rs:907-919
"""
import types

import numpy as np
import pytest

import rust_text
from conftest import reference_text


@pytest.fixture(scope="module")
def ref():
    text = reference_text("crates/rustmpeg-video/src/vvc/filter.rs")
    names = ["long_tap_luma_line", "short_filter_luma_line", "chroma_strong_line", "chroma_weak_line_p1q3",
             "chroma_weak_line", "luma_sample_decision", "chroma_sample_decision"]
    return {n: rust_text.compile_fn(text, n) for n in names}


def _lines(rng, n, length, bd, smooth):
    """Random lines: half of them smooth ramps plus small noise (so that the
    weak filter's |delta| < 10 tC gate passes), half fully random."""
    maxv = (1 << bd) - 1
    out = []
    for k in range(n):
        if smooth and k % 2 == 0:
            base = rng.integers(0, maxv + 1)
            step = rng.integers(-8, 9)
            jump = rng.integers(-40, 41)
            v = base + step * np.arange(length) + np.where(np.arange(length) >= length // 2, jump, 0)
            v = v + rng.integers(-3, 4, length)
        else:
            v = rng.integers(0, maxv + 1, length)
        out.append(np.clip(v, 0, maxv).astype(int).tolist())
    return out


LONG_PAIRS = [(7, 7), (7, 5), (5, 7), (5, 5), (5, 3), (3, 5), (7, 3), (3, 7)]


@pytest.mark.parametrize("bd", [8, 10])
def test_long_luma_filter_equals_reference_text(po, ref, bd):
    rng = np.random.default_rng(1)
    maxv = (1 << bd) - 1
    for lp, lq in LONG_PAIRS:
        for tc in [1 << 20, 1, 7, 40]:
            for line in _lines(rng, 60, 16, bd, smooth=True):
                want = list(line)
                ref["long_tap_luma_line"](want, lp, lq, tc, maxv)
                assert po.dbk_long_line(line, lp, lq, tc, maxv) == want, (lp, lq, tc, line)


def test_long_luma_filter_impulse_responses(po, ref):
    """With no clip binding, each output of the long filter is an affine
    function of the inputs; its response to a unit step in every input
    position (scaled so no rounding hides a tap) must agree."""
    for lp, lq in LONG_PAIRS:
        for j in range(16):
            line = [512] * 16
            line[j] += 256
            want = list(line)
            ref["long_tap_luma_line"](want, lp, lq, 1 << 20, (1 << 12) - 1)
            assert po.dbk_long_line(line, lp, lq, 1 << 20, (1 << 12) - 1) == want, (lp, lq, j)


@pytest.mark.parametrize("bd", [8, 10, 12])
def test_short_luma_filters_equal_reference_text(po, ref, bd):
    rng = np.random.default_rng(2)
    maxv = (1 << bd) - 1
    for de in (1, 2):
        for dep in (0, 1):
            for deq in (0, 1):
                for tc in [1 << 16, 1, 3, 12, 50]:
                    edge = types.SimpleNamespace(de=de, dep=bool(dep), deq=bool(deq), tc=tc)
                    for line in _lines(rng, 40, 16, bd, smooth=True):
                        want = list(line)
                        ref["short_filter_luma_line"](want, edge, maxv)
                        got = po.dbk_short_line(line, de, dep, deq, tc, maxv)
                        assert got == want, (de, dep, deq, tc, line)


@pytest.mark.parametrize("bd", [8, 10])
def test_chroma_filters_equal_reference_text(po, ref, bd):
    rng = np.random.default_rng(3)
    maxv = (1 << bd) - 1
    forms = [((3, 3), "chroma_strong_line"), ((1, 3), "chroma_weak_line_p1q3"), ((1, 1), "chroma_weak_line")]
    for (lp, lq), name in forms:
        for tc in [1 << 16, 1, 4, 20]:
            for line in _lines(rng, 80, 8, bd, smooth=True):
                want = list(line)
                ref[name](want, tc, maxv)
                assert po.dbk_chroma_line(line, lp, lq, tc, maxv) == want, (name, tc, line)


def test_sample_decisions_equal_reference_text(po, ref):
    rng = np.random.default_rng(4)
    L = po.lib()
    for _ in range(20000):
        p0, p3, q0, q3 = (int(v) for v in rng.integers(0, 1024, 4))
        dpq, sp, sq, spq = (int(v) for v in rng.integers(0, 120, 4))
        spl, sql = (bool(v) for v in rng.integers(0, 2, 2))
        beta, tc = int(rng.integers(0, 400)), int(rng.integers(0, 100))
        want = ref["luma_sample_decision"](p0, p3, q0, q3, dpq, sp, sq, spq, spl, sql, beta, tc)
        assert bool(L.orc_dbk_luma_sample_decision(p0, p3, q0, q3, dpq, sp, sq, spq, spl, sql, beta, tc)) == want
        want = ref["chroma_sample_decision"](p0, p3, q0, q3, dpq, beta, tc)
        assert bool(L.orc_dbk_chroma_sample_decision(p0, p3, q0, q3, dpq, beta, tc)) == want
