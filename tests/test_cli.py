"""The reference-compatible CLI binaries (hivehashtable_amd/bin/, built with the library)."""
import os
import re
import subprocess

import pytest

from conftest import gpu_available

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "hivehashtable_amd", "bin")
NAMES = ("hive_hash_table_benchmark", "hive_hash_table_ycsb", "hive_hash_table_lookup_only_workload")


def run(name, *args, timeout=120):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built (python -m hivehashtable_amd.build)")
    return subprocess.run([exe, *args], capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("name", NAMES)
def test_cli_help_and_unknown_flag(name):
    r = run(name, "--help")
    assert r.returncode == 0 and "--table_size" in r.stdout and "--data_layout" in r.stdout
    r = run(name, "--no-such-flag")
    assert r.returncode == 1 and "Unknown argument '--no-such-flag'" in r.stderr


def test_cli_flag_sets_match_the_reference_binaries():
    assert "--ycsb_workload_type" in run("hive_hash_table_ycsb", "--help").stdout
    assert "--hash_policy" in run("hive_hash_table_benchmark", "--help").stdout
    assert "--n_operations" in run("hive_hash_table_lookup_only_workload", "--help").stdout
    # flags of one binary are unknown to another, as in the reference
    assert run("hive_hash_table_ycsb", "--insert_ratio", "0.5").returncode == 1
    assert run("hive_hash_table_benchmark", "--ycsb_workload_type", "A").returncode == 1


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_cli_runs_on_the_gpu():
    print("[case] cli binaries", flush=True)
    r = run("hive_hash_table_benchmark", "--table_size", "16", "--load_factor", "0.8", "--num_iterations", "2",
            "--seed", "3", "--distribution", "unique")
    assert r.returncode == 0, r.stderr
    assert float(re.search(r"Throughput: ([0-9.e+]+)", r.stdout).group(1)) > 0
    r = run("hive_hash_table_benchmark", "--table_size", "16", "--data_layout", "WarpSpeed", "--seed", "3")
    assert r.returncode == 1  # the competitor layout is not provided
    for w in "ABCD":
        r = run("hive_hash_table_ycsb", "--table_size", "16", "--num_iterations", "2", "--ycsb_workload_type", w,
                "--num_operations_ycsb", "20000", "--seed", "5")
        assert r.returncode == 0, (w, r.stderr)
        assert float(re.search(r"Throughput: ([0-9.e+]+)", r.stdout).group(1)) > 0
    r = run("hive_hash_table_lookup_only_workload", "--table_size", "16", "--num_iterations", "2", "--seed", "7")
    assert r.returncode == 0, r.stderr
    rows = re.findall(r"Throughput \((\d+)% exist\): ([0-9.e+]+)  \(found (\d+) of (\d+)\)", r.stdout)
    assert [int(p) for p, *_ in rows] == [100, 75, 50, 25, 0]
    for p, thr, found, n in rows:
        # lookups of prefilled keys are all found, lookups of key + N never
        assert int(found) == round(int(n) * int(p) / 100) and float(thr) > 0
