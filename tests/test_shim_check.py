"""Host-side check of the C++ call surface's scoring inputs
(mirage_amd/csrc/matrix_hip_check.cpp), compiled with g++ and run on the
CPU: well-formed inputs pass, and each malformed one is refused with a
message naming the query, candidate and seed."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mirage_amd", "csrc")


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("no g++")
    exe = str(tmp_path_factory.mktemp("check") / "check_driver")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
                    "-I", CSRC, os.path.join(ROOT, "tests", "cpp", "check_driver.cpp"),
                    os.path.join(CSRC, "matrix_hip_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    return dict(line.split("\t", 1) for line in out.strip().split("\n"))


def test_well_formed_inputs_pass(results):
    assert results["good"] == "ok"
    # start.Query > last.Query: skipped by the scorer (MatrixCPU.cpp:57-58), not checked
    assert results["skipped_seed"] == "ok"
    # a seed that ends exactly at the read's and the reference's last base
    assert results["edge_fits"] == "ok"


@pytest.mark.parametrize("case,words", [
    ("bad_ref_id", ["query 0", "candidate 0", "reference id 2"]),
    ("last_query_past_read", ["query 0", "candidate 0", "seed 0", "last.Query 41"]),
    ("start_ref_past_ref", ["seed 0", "outside the reference's 50 bases"]),
    ("last_ref_past_ref", ["seed 0", "40..51"]),
    ("uneven_seed_lists", ["candidate 0", "2 start but 1 last"]),
    ("uneven_cand_lists", ["query 0", "2 last seed lists"]),
    ("missing_query_lists", ["2 queries", "1 candidate lists"]),
    ("query_too_long", ["query 0", "length 5000"]),
])
def test_malformed_inputs_are_refused(results, case, words):
    msg = results[case]
    assert msg != "ok"
    for w in words:
        assert w in msg, (case, msg)
