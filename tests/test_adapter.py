"""The reference-side adapter (mirage_amd/csrc/adapter/MatrixGPUHIP.cpp),
compiled with g++ against the reference's own headers (MirAge/Analyser.h)
and linked with a recording stand-in for mirage_hip::Device
(tests/cpp/adapter_driver.cpp): it forwards the query slice and the
candidates inside the reference window with their seeds, and assembles
MatrixResult the way MatrixGPUCU.cu does.  Skipped where the reference
sources are absent."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mirage_amd", "csrc")
REF = "/root/reference/MirAge"


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("no g++")
    if not os.path.exists(os.path.join(REF, "Analyser.h")):
        pytest.skip("reference sources absent")
    exe = str(tmp_path_factory.mktemp("adapter") / "adapter_driver")
    subprocess.run(["g++", "-std=c++14", "-fpermissive", "-D__linux__", "-w",
                    "-I", REF, "-I", os.path.join(ROOT, "include"), "-I", CSRC,
                    "-I", os.path.join(CSRC, "adapter"),
                    os.path.join(ROOT, "tests", "cpp", "adapter_driver.cpp"),
                    os.path.join(CSRC, "adapter", "MatrixGPUHIP.cpp"),
                    os.path.join(REF, "MatrixResult.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    return dict(line.split("\t", 1) for line in out.strip().split("\n"))


def test_device_entry_points_forward(results):
    assert results["devices"] == "2"          # InitialiseCuda(2) opened devices 0 and 1
    assert results["rowhit"] == "4096"        # Get_LOCAL_DATA_rowhit_SIZE = max read length
    assert results["blocksize"] == "123456"   # CUDAGetMaxBlockSize(1)
    assert results["loaded"] == "2 5 31"      # 2 references, 5 bases, considerthreshold 31


def test_slice_window_and_seeds(results):
    # Queries 1..3 of five; references 2..7 (inclusive, MatrixGPUCU.cu:703).
    # Query 'b' keeps candidates 2 and 7 of {1, 2, 7, 9}, each with both of
    # its seeds (start-last pairs); 'c' keeps none of {8}; 'd' keeps 3 and 4.
    assert results["seen"] == ("b:2(101,1-131,31;102,2-132,32;)7(102,1-132,31;103,2-133,32;)|"
                               "c:|"
                               "d:3(300,3-330,33;301,4-331,34;)4(301,3-331,33;302,4-332,34;)|")


def test_results_as_the_cuda_engine_reports_them(results):
    # Slice-relative QueryIDs; reference 7 has no score and is left out;
    # query 'c' has no scored candidate and gets no MatrixResult.
    assert results["results"] == "0=0[2:0,]2=21[3:20,4:21,]"


def test_errors_reach_the_callers_retry(results):
    # Analyser::TGetHighestScoreUsingGPU catches std::exception and retries
    # with a smaller batch (Analyser.cpp:841).
    assert "device allocation" in results["error"]
    assert "not initialised" in results["bad_device"]
