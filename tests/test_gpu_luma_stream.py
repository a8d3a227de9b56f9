"""The library's decoder of the RD search's luma CABAC substream
(rmh_luma_stream_decode): host parse, one device residual batch, host
prediction. It must rebuild the encoder's reconstruction and splits, and agree
with the oracle's separately restated decoder (orc_decode_luma_stream)."""
import numpy as np
import pytest

from conftest import load_package

rmh = load_package()
pytestmark = pytest.mark.gpu


def _content(seed, w, h):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    img = 60 + 0.8 * xx + 0.5 * yy + rng.normal(0, 6, (h, w))
    img[h // 5:h // 2, w // 6:w // 2] += 70
    img[h // 2:, w // 2:] = 40 + 60 * ((xx[h // 2:, w // 2:] // 4 + yy[h // 2:, w // 2:] // 4) % 2)
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


@pytest.mark.parametrize("qp,w,h", [(22, 160, 96), (27, 256, 128), (32, 1280, 736), (37, 96, 64)])
def test_luma_stream_decodes_on_the_device_path(acc, po, qp, w, h):
    print("luma stream decode qp=%d %dx%d" % (qp, w, h), flush=True)
    src = _content(qp + w, w, h)
    rec, splits, _, st, data = rmh.rd_encode_luma_stream(src, qp)
    got, got_splits = acc.luma_stream_decode(data, w, h, qp)
    np.testing.assert_array_equal(got, rec)
    np.testing.assert_array_equal(got_splits, splits)
    ref = po.decode_luma_stream(data, w, h, qp)
    np.testing.assert_array_equal(got, ref[0])


def test_luma_stream_flat_picture_has_no_residual_tus(acc):
    """A flat picture codes every CU with cbf 0: the device batch is empty."""
    src = np.full((64, 96), 117, np.uint8)
    rec, splits, _, _, data = rmh.rd_encode_luma_stream(src, 30)
    got, got_splits = acc.luma_stream_decode(data, 96, 64, 30)
    np.testing.assert_array_equal(got, rec)
    np.testing.assert_array_equal(got_splits, splits)


def test_luma_stream_rejects_a_truncated_stream(acc):
    src = _content(1, 96, 64)
    _, _, _, _, data = rmh.rd_encode_luma_stream(src, 27)
    with pytest.raises(rmh.InvalidInput):
        acc.luma_stream_decode(data[:len(data) // 3], 96, 64, 27)
