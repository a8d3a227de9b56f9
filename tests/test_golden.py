"""The frozen fixtures of tests/golden/ (written by tests/golden/make_fixtures.py
from the oracle): the oracle must still produce them (CPU) and the HIP path
must produce them bit for bit (GPU, through the C ABI)."""
import ctypes
import os

import numpy as np
import pytest

from frames import dbk_frame_struct, mc_bi_frame_struct, mc_frame_struct
from workloads import BI_PU_DTYPE

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RESIDUAL = ["residual_tools_10bit", "residual_plain_8bit"]


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz")) as z:  # allow_pickle stays False
        return {k: z[k] for k in z.files}


def _mc(fx, run):
    refs = [np.ascontiguousarray(fx["ref_" + c]) for c in "yuv"]
    dst = [np.zeros_like(p) for p in refs]
    f, keep = mc_frame_struct(refs, dst, 10)
    run(f, fx["pus"])
    return dst


def _bi_pus(fx):
    """The fixture's bi PUs in the current rmh_bi_pu_desc layout. The fixture
    was written before the `refine` byte took the first of the three reserved
    bytes; the bytes are the same (zero there), so this is a relabelling."""
    pus = fx["pus"]
    if pus.dtype != BI_PU_DTYPE:
        assert pus.dtype.itemsize == BI_PU_DTYPE.itemsize
        pus = pus.view(BI_PU_DTYPE)
        assert (pus["refine"] == 0).all()
    return pus


def _mc_bi(fx, run):
    r0 = [np.ascontiguousarray(fx["ref0_" + c]) for c in "yuv"]
    r1 = [np.ascontiguousarray(fx["ref1_" + c]) for c in "yuv"]
    dst = [np.zeros_like(p) for p in r0]
    f, keep = mc_bi_frame_struct(r0, r1, dst, 10)
    run(f, _bi_pus(fx))
    return dst


def _deblock(fx, run):
    planes = [np.ascontiguousarray(fx["in_" + c]).copy() for c in "yuv"]
    f, keep = dbk_frame_struct(planes, fx["edges"], 10, [int(q) for q in fx["qp"]])
    run(f)
    return planes


def _same(got, fx):
    for c, p in zip("yuv", got):
        assert (p == fx["out_" + c]).all(), c


# ---- the oracle reproduces the frozen vectors ------------------------------

@pytest.mark.parametrize("name", RESIDUAL)
def test_oracle_residual_fixture(po, name):
    fx = load(name)
    assert (po.residual_batch(fx["tus"], fx["coeff"]) == fx["residual"]).all()


def test_oracle_mc_fixtures(po):
    fx = load("mc_uni_10bit")
    _same(_mc(fx, lambda f, pus: po.mc_batch(f, pus)), fx)
    fx = load("mc_bi_10bit")
    _same(_mc_bi(fx, lambda f, pus: po.mc_bi_batch(f, pus)), fx)


def test_oracle_deblock_fixture(po):
    fx = load("deblock_10bit")

    def run(f):
        arr = (type(f) * 1)(f)
        po.lib().orc_deblock_frames(ctypes.addressof(arr), 1)
    got = _deblock(fx, run)
    _same(got, fx)
    assert any((g != fx["in_" + c]).any() for c, g in zip("yuv", got))


# ---- the HIP path reproduces them ------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("name", RESIDUAL)
def test_gpu_residual_fixture(acc, name):
    print("golden", name, flush=True)
    fx = load(name)
    assert (acc.residuals(fx["tus"], fx["coeff"]) == fx["residual"]).all()


@pytest.mark.gpu
def test_gpu_mc_fixtures(acc):
    print("golden mc", flush=True)
    fx = load("mc_uni_10bit")
    _same(_mc(fx, lambda f, pus: acc.predict([f], pus)), fx)
    fx = load("mc_bi_10bit")
    _same(_mc_bi(fx, lambda f, pus: acc.predict_bi([f], pus)), fx)


@pytest.mark.gpu
def test_gpu_deblock_fixture(acc):
    print("golden deblock", flush=True)
    fx = load("deblock_10bit")
    _same(_deblock(fx, lambda f: acc.deblock([f])), fx)
