"""Extract the normative fixed ALF filters into C headers.

The 64 fixed luma ALF filters and the class-to-filter mapping of the 16 fixed
filter sets (H.266 Tables 48 and 49, VTM AdaptiveLoopFilter.cpp) are trained
integer data with no closed form, so they are carried as data. This script
reads them from the reference source text
(crates/rustmpeg-video/src/vvc/filter.rs:1374 ALF_FIXED_FILTER_COEFF,
:1443 ALF_CLASS_TO_FILTER_MAPPING), checks their shapes, and writes them into
the oracle and into the HIP library:

    oracle/alf_data.h                      ORC_ALF_FIXED, ORC_ALF_CLASS_MAP
    rustmpeg-next_amd/csrc/alf_data.h      RMH_ALF_FIXED, RMH_ALF_CLASS_MAP

Both headers are committed; the GPU host never runs this script.

Usage: python oracle/gen_alf_data.py [reference_root]
"""
import os
import sys

from gen_lfnst_data import parse_array

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


def emit(path, prefix, fixed, cmap):
    lines = [
        "/* Generated by oracle/gen_alf_data.py: the normative fixed luma ALF filters",
        " * (H.266 Table 49, [filter][coefficient 0..12]) and the class-to-filter",
        " * mapping of the 16 fixed filter sets (Table 48, [set][class 0..24]).",
        " *",
        " * The values derive from the VVC reference software VTM",
        " * (https://vcgit.hhi.fraunhofer.de/jvet/VVCSoftware_VTM),",
        " * Copyright (c) 2010-2020, ITU/ISO/IEC, BSD 3-Clause License; the license",
        " * text is reproduced in THIRD_PARTY.md at the repository root. */",
        "#ifndef %s_ALF_DATA_H" % prefix,
        "#define %s_ALF_DATA_H" % prefix,
        "#include <stdint.h>",
        "",
    ]

    def dump(name, ctype, vals, dims):
        lines.append("static const %s %s%s = {" % (ctype, name, "".join("[%d]" % d for d in dims)))
        row = dims[-1]
        for k in range(0, len(vals), row):
            lines.append("    {" + ",".join("%d" % v for v in vals[k:k + row]) + "},")
        lines.append("};")
        lines.append("")

    dump("%s_ALF_FIXED" % prefix, "int8_t", fixed, (64, 13))
    dump("%s_ALF_CLASS_MAP" % prefix, "uint8_t", cmap, (16, 25))
    lines.append("#endif")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    src = os.path.join(ref, "crates/rustmpeg-video/src/vvc/filter.rs")
    with open(src) as f:
        text = f.read()
    fixed = parse_array(text, "ALF_FIXED_FILTER_COEFF")
    cmap = parse_array(text, "ALF_CLASS_TO_FILTER_MAPPING")
    assert len(fixed) == 64 * 13, len(fixed)
    assert len(cmap) == 16 * 25, len(cmap)
    assert all(-128 <= v <= 127 for v in fixed) and all(0 <= v < 64 for v in cmap)
    emit(os.path.join(REPO, "oracle", "alf_data.h"), "ORC", fixed, cmap)
    emit(os.path.join(REPO, "rustmpeg-next_amd", "csrc", "alf_data.h"), "RMH", fixed, cmap)
    print("wrote ALF tables: %d + %d entries" % (len(fixed), len(cmap)))


if __name__ == "__main__":
    main()
