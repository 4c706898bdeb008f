"""The benchmark's host logic on the CPU: the one-GPU refusal, the
algorithmic byte counts its roofline divides by, the committed traffic
summary it reads, and the rocprofv3 table writer."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


@pytest.mark.parametrize("N, published", [(40, 6_540_804), (200, 829_120_004), (120, 178_675_204)])
def test_csr_bytes_match_the_published_figures(N, published):
    # 7-point N^3 grid: n = N^3 rows, nnz = 7n - 6N^2; one FP64 CSR product
    # moves 12 nnz + 20 n + 4 bytes (val, col, row pointers, x, y)
    n = N ** 3
    nnz = 7 * n - 6 * N * N
    for kind in (0, 1, 2, 3, 4, 5):
        assert bench.format_bytes(kind, nnz, n) == published


def test_format_bytes_of_the_other_engines():
    n, nnz = 8_000_000, 55_760_000
    # zero-padded diagonal and three upper diagonals, p read, q written
    assert bench.format_bytes(8, nnz, n, sym_diagonals=3) == 48 * n == 384_000_000
    # 1-byte offset code per nonzero in place of the int32 column
    assert bench.format_bytes(6, nnz, n) == bench.format_bytes(0, nnz, n) - 3 * nnz
    # 1-byte pattern id per row in place of columns and row pointers
    assert bench.format_bytes(7, nnz, n) == 8 * nnz + 1 * n + 16 * n


def test_committed_traffic_lies_between_the_format_and_csr_bytes():
    t, src = bench.traffic_per_launch("k_spmv_pq_sym", 8)
    assert src is not None and os.path.exists(src)
    n, nnz = 8_000_000, 55_760_000
    assert bench.format_bytes(8, nnz, n, sym_diagonals=3) <= t < bench.format_bytes(0, nnz, n)
    # a summary taken on another engine is not used for this one
    assert bench.traffic_per_launch("k_spmv_pq_sym", 0) == (None, None)


def test_kernel_stats_table(tmp_path):
    csv = tmp_path / "k_kernel_stats.csv"
    csv.write_text('"Name","Calls","TotalDurationNs","AverageNs","Percentage","MinNs","MaxNs","StdDev"\n'
                   '"void tfg::k_spmv_pq_sym<3, 1>(tfg::CgArgs, int)",10,851000,85100,61.5,84000,90000,1.0\n')
    out = tmp_path / "table.md"
    bench.summarize_stats(str(csv), str(out), {"workload": "w", "cmd": "rocprofv3 --kernel-trace --stats -- python3 bench.py"})
    text = out.read_text()
    assert "rocprofv3 --kernel-trace --stats -- python3 bench.py" in text
    assert "| `void tfg::k_spmv_pq_sym<3, 1>(tfg::CgArgs, int)` | 10 | 0.9 | 85.1 | 84.0 | 90.0 | 61.5 |" in text


def test_roofline_prices_the_engine_format():
    # the symmetric engine's launch at the in-solve time measured on the
    # 200^3 system (88.4 us): its own 384 MB give a fraction below 1, the
    # CSR formula's 829 MB a "fraction" above 1, reported apart
    n, nnz = 8_000_000, 55_760_000
    r = bench.spmv_roofline(8, nnz, n, 3, 0.0884)
    assert r["bytes_per_launch"] == 384_000_000
    assert r["frac"] == pytest.approx(384e6 / 88.4e-6 / 8e12)
    assert 0 < r["frac"] <= 1.0
    assert r["csr_equivalent"]["bytes_per_launch"] == 829_120_004
    assert r["csr_equivalent"]["frac"] > 1.0
    # a CSR engine is priced by the CSR formula itself
    c = bench.spmv_roofline(5, nnz, n, 0, 0.150)
    assert c["bytes_per_launch"] == c["csr_equivalent"]["bytes_per_launch"]
    assert c["frac"] <= 1.0
