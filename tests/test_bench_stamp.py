"""The kernel-source stamp that ties a committed counter summary to the code
it was measured on (CPU test): comments and layout do not change it, code
does, and the committed summaries name the sources they stamp."""

import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

SRC = """// leading comment
#include "x.hpp"
/* block
   comment */
__global__ void k(float *p) {   // trailing comment
  p[0] = 1.0f;

}
"""


def test_comments_and_layout_are_not_code():
    edited = SRC.replace("leading comment", "another note").replace("block\n   comment",
                                                                      "a new block")
    relaid = SRC.replace("  p[0]", "        p[0]").replace("\n\n", "\n\n\n")
    assert bench._code_text(edited) == bench._code_text(SRC)
    assert bench._code_text(relaid) == bench._code_text(SRC)
    assert "comment" not in bench._code_text(SRC)


def test_code_changes_are_code():
    assert bench._code_text(SRC.replace("1.0f", "2.0f")) != bench._code_text(SRC)
    assert bench._code_text(SRC.replace("p[0]", "p[1]")) != bench._code_text(SRC)


def test_summaries_stamp_known_sources():
    for summary, names in bench.SUMMARY_SOURCES.items():
        assert os.path.exists(os.path.join(ROOT, "profiles", summary)), summary
        for n in names:
            assert os.path.exists(os.path.join(ROOT, "ccpi-reconstruction_amd", "csrc", n)), n
        val, state = bench.measured_traffic(
            "cone_bp_kernel" if summary.startswith("cone") else "par_fp_kernel", summary)
        assert state in ("current", "stale") and (val is None) == (state == "stale")
