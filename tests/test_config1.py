"""Intra-subset round trip (CPU, oracle) and the two-pass GPU decode.

Known answers ported: QP 4 is lossless (crates/rustmpeg-video/src/vvc/
slice.rs:5152-5175 and vvc/mod.rs:3308-3340) and QP 50 selects both MTS and
LFNST on a smooth ramp (slice.rs:5278-5323). Encoder reconstruction equals
decoder reconstruction for every QP. On the GPU, the device-residual decode
(rmh_intra_subset_decode) must reproduce the CPU decode's YUV MD5 per frame.
Byte-level bitstream parity with the reference is not pinned (no golden
access units exist in the reference); entropy coding is not restated."""
import numpy as np
import pytest

import config1 as C1
import workloads as W


def _encode(po, planes, qp):
    return po.intra_encode(*planes, qp, W.TU_DTYPE)


def test_qp4_picture_round_trip_is_exact(po):
    # slice.rs:5152-5175
    w = h = 64
    jj, ii = np.mgrid[0:h, 0:w]
    Y = ((ii * 3 + jj * 5) % 251).astype(np.uint8)
    U = np.full((32, 32), 90, np.uint8)
    V = np.full((32, 32), 160, np.uint8)
    rec, modes, tus, coeff = _encode(po, (Y, U, V), 4)
    dec = po.intra_decode(w, h, modes, tus, coeff)
    for a, b in zip(dec, (Y, U, V)):
        assert (a == b).all()
    assert (tus["flags"] == 1).all()  # every block transform-skip at QP 4


def test_three_frames_qp4_lossless_and_qp26_round_trip(po):
    for qp in (4, 26):
        for f in range(3):
            src = C1.source(f)
            rec, modes, tus, coeff = _encode(po, src, qp)
            dec = po.intra_decode(64, 64, modes, tus, coeff)
            assert C1.md5(rec) == C1.md5(dec)
            if qp == 4:
                assert C1.md5(dec) == C1.md5(src)


def test_qp50_selection_uses_mts_and_lfnst(po):
    # slice.rs:5278-5323: 200x120, Y = 16 + (x + y) % 40, flat chroma
    w, h = 200, 120
    jj, ii = np.mgrid[0:h, 0:w]
    Y = (16 + (ii + jj) % 40).astype(np.uint8)
    U = np.full((60, 100), 128, np.uint8)
    V = np.full((60, 100), 128, np.uint8)
    planes = C1.pad((Y, U, V), 224, 128)
    rec, modes, tus, coeff = _encode(po, planes, 50)
    luma = tus[0::3]
    assert ((luma["tr_hor"] != 0) | (luma["tr_ver"] != 0)).any(), "MTS never selected"
    assert (luma["lfnst_idx"] != 0).any(), "LFNST never selected"
    dec = po.intra_decode(224, 128, modes, tus, coeff)
    assert C1.md5(rec) == C1.md5(dec)


@pytest.mark.gpu
@pytest.mark.parametrize("qp", [4, 26, 50])
def test_gpu_two_pass_decode_md5_per_frame(acc, po, qp):
    for f in range(3):
        print("config1 qp=%d frame=%d" % (qp, f), flush=True)
        src = C1.source(f)
        rec, modes, tus, coeff = _encode(po, src, qp)
        cpu = po.intra_decode(64, 64, modes, tus, coeff)
        gpu = acc.intra_subset_decode(64, 64, modes, tus, coeff)
        assert C1.md5(gpu) == C1.md5(cpu), "frame %d" % f


@pytest.mark.gpu
def test_gpu_two_pass_decode_with_mts_and_lfnst(acc, po):
    w, h = 200, 120
    jj, ii = np.mgrid[0:h, 0:w]
    planes = C1.pad(((16 + (ii + jj) % 40).astype(np.uint8), np.full((60, 100), 128, np.uint8),
                     np.full((60, 100), 128, np.uint8)), 224, 128)
    rec, modes, tus, coeff = _encode(po, planes, 50)
    cpu = po.intra_decode(224, 128, modes, tus, coeff)
    gpu = acc.intra_subset_decode(224, 128, modes, tus, coeff)
    assert C1.md5(gpu) == C1.md5(cpu)
