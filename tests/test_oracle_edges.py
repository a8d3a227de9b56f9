"""The edge-map oracle (orc_edge_map) against the reference: the bS and
filter-length rules are the reference's own functions read as text and
evaluated in Python (derive_boundary_strength, vvc/filter.rs:202-226;
luma_lens / chroma_lens, vvc/decode_general.rs:4916-4935), composed by the
edge loops of apply_deblock (decode_general.rs:4977-5187) and deblock_bs
(4939-4975), over random coded pictures; plus the known answers of the
length rules (decode_general.rs:5718-5739, the reference's own test)."""
import types

import numpy as np
import pytest

import rust_text
import workloads as W
from conftest import load_package, reference_text

abi = load_package().abi


def oracle_edges(po, cells, vis_w, vis_h, ctb):
    gh, gw = cells.shape
    cells = np.ascontiguousarray(cells)
    edges = np.zeros((gh, gw, 4), np.uint8)
    f = abi.EdgeFrame(cells.ctypes.data, edges.ctypes.data, gw, gh, vis_w, vis_h, ctb, 0)
    import ctypes
    assert po.lib().orc_edge_map(ctypes.byref(f), 1) == 0
    return edges


@pytest.fixture(scope="module")
def ref():
    dg = reference_text("crates/rustmpeg-video/src/vvc/decode_general.rs")
    fl = reference_text("crates/rustmpeg-video/src/vvc/filter.rs")
    return {"luma_lens": rust_text.compile_fn(dg, "luma_lens"), "chroma_lens": rust_text.compile_fn(dg, "chroma_lens"),
            "bs": rust_text.compile_fn(fl, "derive_boundary_strength")}


def reference_edges(ref, cells, vis_w, vis_h, ctb):
    gh, gw = cells.shape
    out = np.zeros((gh, gw, 4), np.uint8)
    none = 0xffffffff

    def byte(bs, lens, ctb_row=False):
        return (bs & 3) | (((lens[0] - 1) // 2) << 2) | (((lens[1] - 1) // 2) << 4) | (0x40 if ctb_row else 0)

    def bs(p, q):  # deblock_bs: intra, cbf, then the `field` MV and POC
        mv = lambda c: types.SimpleNamespace(hor=int(c["mv_x"]), ver=int(c["mv_y"]))
        return ref["bs"](bool(p["flags"] & 1), bool(q["flags"] & 1), bool(p["flags"] & 2), bool(q["flags"] & 2),
                         True, mv(p), mv(q), int(p["ref_poc"]), int(q["ref_poc"]))

    for x in range(4, vis_w, 4):
        for y in range(0, vis_h, 4):
            l, r = cells[y // 4, x // 4 - 1], cells[y // 4, x // 4]
            if l["tu_y"] != none and r["tu_y"] != none and l["tu_y"] != r["tu_y"]:
                out[y // 4, x // 4, 0] = byte(bs(l, r), ref["luma_lens"](int(l["tu_wy"]), int(r["tu_wy"])))
    for y in range(4, vis_h, 4):
        for x in range(0, vis_w, 4):
            t, b = cells[y // 4 - 1, x // 4], cells[y // 4, x // 4]
            if t["tu_y"] != none and b["tu_y"] != none and t["tu_y"] != b["tu_y"]:
                out[y // 4, x // 4, 1] = byte(bs(t, b), ref["luma_lens"](int(t["tu_hy"]), int(b["tu_hy"])), y % ctb == 0)
    cw, ch = (vis_w + 1) // 2, (vis_h + 1) // 2
    for x in range(2, cw, 2):
        if (2 * x) % 16:
            continue
        for y in range(0, ch - 1, 2):
            l, r = cells[y // 2, x // 2 - 1], cells[y // 2, x // 2]
            if l["tu_c"] != none and r["tu_c"] != none and l["tu_c"] != r["tu_c"]:
                out[y // 2, x // 2, 2] = byte(2, ref["chroma_lens"](int(l["tu_wc"]), int(r["tu_wc"]), False))
    for y in range(2, ch, 2):
        if (2 * y) % 16:
            continue
        for x in range(0, cw - 1, 2):
            t, b = cells[y // 2 - 1, x // 2], cells[y // 2, x // 2]
            if t["tu_c"] != none and b["tu_c"] != none and t["tu_c"] != b["tu_c"]:
                out[y // 2, x // 2, 3] = byte(2, ref["chroma_lens"](int(t["tu_hc"]), int(b["tu_hc"]),
                                                                    y % (ctb // 2) == 0))
    return out


@pytest.mark.parametrize("seed,w,h,ctb", [(1, 256, 128, 128), (2, 250, 134, 64), (3, 192, 96, 32)])
def test_edge_map_oracle_equals_reference_rules(po, ref, seed, w, h, ctb):
    cells = W.edge_cells(seed, width=w, height=h, ctu=ctb)
    # perturb MVs so both MV thresholds and the POC rule decide some edges
    rng = np.random.default_rng(seed)
    cells["mv_x"] += rng.choice([0, 7, 8, -8], size=cells.shape)
    got = oracle_edges(po, cells, w, h, ctb)
    want = reference_edges(ref, cells, w, h, ctb)
    assert (got == want).all()
    assert {int(v) & 3 for v in got[..., 0].reshape(-1) if v} == {0, 1, 2} or (got[..., 0] & 3).max() == 2


def test_length_tables_known_answers(ref):
    """vvc/decode_general.rs:5718-5739 (the reference's own test of the tables)."""
    assert ref["luma_lens"](4, 32) == (1, 1) and ref["luma_lens"](32, 4) == (1, 1)
    assert ref["luma_lens"](8, 32) == (3, 7) and ref["luma_lens"](64, 16) == (7, 3)
    assert ref["chroma_lens"](8, 8, False) == (3, 3) and ref["chroma_lens"](8, 8, True) == (1, 3)
    assert ref["chroma_lens"](4, 16, False) == (1, 1)
