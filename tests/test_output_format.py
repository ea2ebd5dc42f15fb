"""Layout of the driver's output files without a GPU: the block log (.log, read by the
reference's gosling) and the per-block .json records, line by line against the reference
program's own files for a DMC of the H atom (tests/golden/ref_output), numbers aside
(src/properties/Properties.cpp:578-641,1020-1034, src/properties/Properties_block.cpp:73-138,
src/properties/Properties_average.cpp:449-520)."""
import json
import math
import os
import re
import shutil
import subprocess

import pytest

from tests._util import GOLDEN, ROOT

LIB = os.path.join(ROOT, "mainline_amd", "libqwalk_amd.so")
NUM = re.compile(r"[-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?|[-+]?\b(inf|nan)\b")


@pytest.mark.skipif(not os.path.exists(LIB) or shutil.which("g++") is None, reason="needs the built library and g++")
def test_log_and_json_have_the_reference_layout(tmp_path):
    exe = tmp_path / "layout"
    subprocess.run(["g++", "-std=c++17", "-O1", "-o", str(exe), os.path.join(ROOT, "tests", "output_layout_harness.cpp"),
                    "-L" + os.path.dirname(LIB), "-lqwalk_amd", "-Wl,-rpath," + os.path.dirname(LIB)], check=True)
    subprocess.run([str(exe), str(tmp_path / "qw.run")], check=True)
    sk = lambda t: [NUM.sub("#", line) for line in t.splitlines()]
    for ext, kept in [("log", "qwlog"), ("json", "json")]:
        mine = open(tmp_path / ("qw.run." + ext)).read()
        theirs = open(os.path.join(GOLDEN, "ref_output", "h_dmc." + kept)).read()
        assert sk(mine) == sk(theirs), ext
    recs = [json.loads(r.replace(":inf,", ":Infinity,")) for r in open(tmp_path / "qw.run.json").read().split("<RS>")
            if r.strip()]
    assert len(recs) == 2
    p = recs[1]["properties"]["total_energy"]
    assert p["value"][0] == pytest.approx(-0.50108566, rel=1e-5) and p["error"][0] == 0
    assert p["sigma"][0] == pytest.approx(math.sqrt(0.0019923765), rel=1e-5)
    assert recs[0]["independent points"] == float("inf")
