"""Extract the normative LFNST matrices and set table into C headers.

The LFNST kernels (H.266 Table "lowFreqTransMatrix", VTM RomLFNST.cpp) are
trained integer data with no closed form, so they are carried as data. This
script reads them from the reference source text
(crates/rustmpeg-video/src/vvc/tr_tables.rs:166 LFNST_8X8, :322 LFNST_4X4,
:478 LFNST_LUT), checks their shapes, and writes them, in this project's own
layout, into the oracle and into the HIP library:

    oracle/lfnst_data.h                      ORC_LFNST8, ORC_LFNST4
    rustmpeg-next_amd/csrc/lfnst_data.h      RMH_LFNST8, RMH_LFNST4

Both headers are committed; the GPU host never runs this script. The set
table LFNST_LUT is not extracted: both consumers derive it from its ranges
and a CPU test compares the ranges with the reference text.

Usage: python oracle/gen_lfnst_data.py [reference_root]
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


def parse_array(text, name):
    m = re.search(r"pub const %s:[^=]*=\s*\[" % re.escape(name), text)
    if not m:
        raise SystemExit("table %s not found" % name)
    i = m.end() - 1
    depth = 0
    j = i
    while True:
        c = text[j]
        if c == "[":
            depth += 1
        elif c == "]":
            depth -= 1
            if depth == 0:
                break
        j += 1
    body = text[i:j + 1]
    return [int(v) for v in re.findall(r"-?\d+", body)]


def emit(path, prefix, l8, l4):
    lines = [
        "/* Generated by oracle/gen_lfnst_data.py: normative LFNST kernels (H.266",
        " * lowFreqTransMatrix). [set 0..3][lfnst_idx-1][input coefficient 0..15][output].",
        " *",
        " * The values derive from the VVC reference software VTM",
        " * (https://vcgit.hhi.fraunhofer.de/jvet/VVCSoftware_VTM),",
        " * Copyright (c) 2010-2020, ITU/ISO/IEC, BSD 3-Clause License; the license",
        " * text is reproduced in THIRD_PARTY.md at the repository root. */",
        "#ifndef %s_LFNST_DATA_H" % prefix,
        "#define %s_LFNST_DATA_H" % prefix,
        "#include <stdint.h>",
        "",
    ]

    def dump(name, vals, dims):
        lines.append("static const int8_t %s%s = {" % (name, "".join("[%d]" % d for d in dims)))
        row = dims[-1]
        for k in range(0, len(vals), row):
            lines.append("    " + ",".join("%d" % v for v in vals[k:k + row]) + ",")
        lines.append("};")
        lines.append("")

    dump("%s_LFNST8" % prefix, l8, (4, 2, 16, 48))
    dump("%s_LFNST4" % prefix, l4, (4, 2, 16, 16))
    lines.append("#endif")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "crates/rustmpeg-video/src/vvc/tr_tables.rs")
    with open(src) as f:
        text = f.read()
    l8 = parse_array(text, "LFNST_8X8")
    l4 = parse_array(text, "LFNST_4X4")
    assert len(l8) == 4 * 2 * 16 * 48, len(l8)
    assert len(l4) == 4 * 2 * 16 * 16, len(l4)
    assert all(-128 <= v <= 127 for v in l8 + l4)
    emit(os.path.join(REPO, "oracle", "lfnst_data.h"), "ORC", l8, l4)
    emit(os.path.join(REPO, "rustmpeg-next_amd", "csrc", "lfnst_data.h"), "RMH", l8, l4)
    print("wrote LFNST tables: %d + %d entries" % (len(l8), len(l4)))


if __name__ == "__main__":
    main()
