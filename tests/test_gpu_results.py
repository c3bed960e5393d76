"""Result logs from the HIP path: the reads of every golden dataset
classified on the device (plain and batched), their results and candidate
output written with mh_write_results, must be the reference's own four log
files byte for byte."""
import pytest

from mirage_amd import native
from oracle import results as R
from tests import _data

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", _data.golden_names())
def test_logs_from_device_results(handle, tmp_path, name):
    fx, rb, ro, qb, qo = _data.load_golden(name)
    ix = handle.build_index(rb, ro, fx["threshold"])
    try:
        for flags in (0, native.MH_CLASSIFY_BATCHED):
            out, top, cands = ix.classify_host(qb, qo, flags=flags, want_cands=True)
            qh, rh = _data.fasta_headers(fx)
            native.write_results(tmp_path, out, top, qh, rh, cands=cands)
            for k in R.KINDS:
                got = (tmp_path / f"results_{k}.log").read_text()
                assert got == fx["logs"][k], (name, k, flags)
    finally:
        ix.free()
