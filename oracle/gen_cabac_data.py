"""Extract the normative CABAC context initialisation tables into C headers.

The (initValue, log2 window) pairs of every context model are normative
H.266 data (Tables 54-125) with no closed form, so they are carried as data.
This script reads them from the reference source text
(crates/rustmpeg-video/src/vvc/cabac.rs:101 CTX_INIT_I, :155 CTX_INIT_P,
:209 CTX_INIT_B), checks their sizes against NUM_CONTEXTS (cabac.rs:19) and
writes them into the oracle and into the library:

    oracle/cabac_data.h                      ORC_CTX_INIT[3][379][2]
    rustmpeg-next_amd/csrc/cabac_data.h      RMH_CTX_INIT[3][379][2]

(slice table 0 = I, 1 = P, 2 = B). Both headers are committed; the GPU host
never runs this script. The context-set offsets (cabac.rs:307-398) are small
named constants written by hand in each consumer; a CPU test compares every
one of them with the reference text.

Usage: python oracle/gen_cabac_data.py [reference_root]
"""
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CABAC_RS = "crates/rustmpeg-video/src/vvc/cabac.rs"


def parse_tables(text):
    n = int(re.search(r"pub const NUM_CONTEXTS: usize = (\d+);", text).group(1))
    out = []
    for name in ("CTX_INIT_I", "CTX_INIT_P", "CTX_INIT_B"):
        m = re.search(r"pub const %s: \[\(u8, u8\); NUM_CONTEXTS\] = \[(.*?)\];" % name, text, re.S)
        if not m:
            raise SystemExit("table %s not found" % name)
        pairs = [(int(a), int(b)) for a, b in re.findall(r"\((\d+),\s*(\d+)\)", m.group(1))]
        if len(pairs) != n:
            raise SystemExit("table %s has %d entries, expected %d" % (name, len(pairs), n))
        out.append(pairs)
    return n, out


def emit(path, prefix, n, tables):
    lines = [
        "/* Generated by oracle/gen_cabac_data.py: normative CABAC context",
        " * initialisation (H.266 Tables 54-125), [slice type I, P, B][context][initValue,",
        " * log2 window size].",
        " *",
        " * The values derive from the VVC reference software VTM",
        " * (https://vcgit.hhi.fraunhofer.de/jvet/VVCSoftware_VTM),",
        " * Copyright (c) 2010-2020, ITU/ISO/IEC, BSD 3-Clause License; the license",
        " * text is reproduced in THIRD_PARTY.md at the repository root. */",
        "#ifndef %s_CABAC_DATA_H" % prefix,
        "#define %s_CABAC_DATA_H" % prefix,
        "#include <stdint.h>",
        "",
        "#define %s_NUM_CONTEXTS %d" % (prefix, n),
        "",
        "static const uint8_t %s_CTX_INIT[3][%d][2] = {" % (prefix, n),
    ]
    for t in tables:
        lines.append("    {")
        for i in range(0, n, 8):
            lines.append("        " + " ".join("{%d, %d}," % p for p in t[i:i + 8]))
        lines.append("    },")
    lines += ["};", "", "#endif", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(root, CABAC_RS)) as f:
        n, tables = parse_tables(f.read())
    emit(os.path.join(HERE, "cabac_data.h"), "ORC", n, tables)
    emit(os.path.join(REPO, "rustmpeg-next_amd", "csrc", "cabac_data.h"), "RMH", n, tables)
    print("wrote %d contexts x 3 slice types" % n)


if __name__ == "__main__":
    main()
