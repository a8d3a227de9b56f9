"""The batched multi-stream configuration, whole: 16 concurrent streams
(4 x 856x480 8-bit, 8 x 1920x1080 10-bit, 4 x 3840x2160 10-bit, 4:2:0) x 8
chained frame steps, each step one graph (MC -> residual add -> deblock over
every stream) whose output is the next step's reference, with the residual
tool mix (MTS types, LFNST, transform skip, dependent quantisation). Every
frame of every step is compared with the CPU oracle run over the same chain
(per-frame MD5 of Y, Cb, Cr). The 854-wide streams are coded at 856, the next
multiple of 8 (a conformance window crops them in a real stream)."""
import concurrent.futures
import hashlib
import time

import numpy as np
import pytest

import pipeline as PL

pytestmark = pytest.mark.gpu

STEPS = 8


def _md5(planes):
    h = hashlib.md5()
    for p in planes:
        h.update(np.ascontiguousarray(p).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module")
def streams():
    t0 = time.perf_counter()
    refs, frames = PL.generate_streams(PL.STREAMS, STEPS, tools=True, seed=500, workers=16)
    print("config 4: generated %d frames in %.1f s" % (STEPS * len(PL.STREAMS), time.perf_counter() - t0), flush=True)
    return refs, frames


@pytest.fixture(scope="module")
def oracle_md5(po, rmh, streams):
    """md5[t][s] of the oracle's output of stream s after step t."""
    refs, frames = streams
    t0 = time.perf_counter()
    out = [[None] * len(refs) for _ in range(STEPS)]

    def chain(s):
        def record(t, planes):
            out[t][s] = _md5(planes)
        PL.oracle_chain(po, rmh.abi, refs[s], [frames[t][s] for t in range(STEPS)], record)

    with concurrent.futures.ThreadPoolExecutor(16) as ex:
        list(ex.map(chain, range(len(refs))))
    print("config 4: oracle over %d chained frames in %.1f s" % (STEPS * len(refs), time.perf_counter() - t0),
          flush=True)
    return out


@pytest.mark.parametrize("out_of_place,autotune", [(False, False), (True, False), (False, True)])
def test_sixteen_streams_eight_chained_steps_match_the_oracle(acc, rmh, streams, oracle_md5, out_of_place, autotune):
    """autotune: every step's MC and recon plans tuned before capture (the
    shapes must not change a sample)."""
    refs, frames = streams
    n_tu = sum(len(f["tus"]) for fr in frames for f in fr)
    tools = sum(int(((f["tus"]["lfnst_idx"] > 0) | (f["tus"]["flags"] > 0)).sum()) for fr in frames for f in fr)
    print("config 4 (%s%s): %d TUs, %d with LFNST / TS / DQ" % ("out of place" if out_of_place else "in place",
                                                               ", autotuned" if autotune else "", n_tu, tools),
          flush=True)
    chain = PL.ChainedStreams(acc, rmh, refs, frames, out_of_place=out_of_place, autotune=autotune)
    s = rmh.hip.Stream()
    try:
        for t in range(STEPS):
            chain.graphs[t].launch(s.handle)
            s.synchronize()
            got = chain.output(t)
            for k in range(len(refs)):
                assert _md5(got[k]) == oracle_md5[t][k], "step %d stream %d %s" % (t, k, PL.STREAMS[k])
        # a replay of the whole chain from the first references gives the same pictures
        chain.reset(refs)
        chain.run(s.handle)
        s.synchronize()
        got = chain.output(STEPS - 1)
        assert [_md5(g) for g in got] == oracle_md5[STEPS - 1]
    finally:
        chain.close()
