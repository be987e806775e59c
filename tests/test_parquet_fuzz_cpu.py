"""Malformed Parquet input against the library's host staging (wg_pq_stage,
wg_pq_scan_runs, wg_pq_decompress), in the manner of the reference's fuzz
tests (internal/storage/parquet/page_reader_fuzz_test.go,
thrift_fuzz_test.go): tests/cpp/build/pq_fuzz stages mutated copies of valid
files — bits flipped, bytes set, zeroed or 0xFF, chunks cut short — and feeds
random bytes to the run scanner and every decompressor, with the staging code
compiled under AddressSanitizer and UBSan. Every call must return a status;
an out-of-bounds access or undefined behaviour aborts the program."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pyarrow as pa
import pyarrow.parquet as pq
import pytest

from wadjet_amd import parquet as PQ

FUZZ = Path(__file__).resolve().parent / "cpp" / "build" / "pq_fuzz"


def _table(n):
    rng = np.random.default_rng(2)
    return pa.table({"k": pa.array(rng.integers(0, 300, n), pa.int64(), mask=rng.random(n) < 0.2),
                     "v": pa.array(rng.normal(size=n), pa.float64())})


@pytest.mark.parametrize("compression,version,column", [
    ("snappy", "1.0", "k"), ("none", "1.0", "v"), ("gzip", "2.0", "k"), ("zstd", "2.0", "v"), ("lz4", "1.0", "k")])
def test_staging_survives_mutated_files(tmp_path, compression, version, column):
    assert FUZZ.exists(), f"{FUZZ} missing: run __graft_entry__.build()"
    path = str(tmp_path / "f.parquet")
    pq.write_table(_table(6000), path, compression=compression, data_page_version=version, data_page_size=2048,
                   row_group_size=2500)
    cs, chunks, n = PQ.column_chunks(path, column)
    phys, _, _ = PQ._wg_type_of(cs)
    args = [str(FUZZ), path, str(phys), str(cs.max_definition_level), "400", "7", str(tmp_path / "m.parquet"), str(n)]
    for c in chunks[:n]:
        args += [str(c.offset), str(c.bytes), str(c.num_rows), str(c.codec)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=240,
                       env={"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=1", "UBSAN_OPTIONS": "print_stacktrace=1"})
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    staged, refused = (int(x) for x in re.findall(r"\d+", r.stdout)[:2])
    assert refused > 0 and staged + refused == 400
