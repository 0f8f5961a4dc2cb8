"""The C++ API (include/cccl_amd/device_radix_sort.hpp: DeviceRadixSort,
DeviceSegmentedRadixSort, DeviceSegmentedSort, DeviceTopK, DoubleBuffer; and
include/cccl_amd/thrust_sort.hpp: Thrust's radix route) through the host
program tests/cpp/test_device_api, which checks every case against std::stable_sort on
the host.  build() compiles the program beside its source."""

import subprocess
from pathlib import Path

import pytest

EXE = Path(__file__).resolve().parent / "cpp" / "test_device_api"


def _run(*args, timeout=120):
    assert EXE.exists(), f"{EXE} missing: run __graft_entry__.build()"
    print(f"running {EXE.name} {' '.join(args)}", flush=True)
    r = subprocess.run([str(EXE), *args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout)
    return r


def test_cpp_api_size_queries_without_a_device():
    r = _run("--query-only")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASS temp-storage queries" in r.stdout


@pytest.mark.gpu
def test_cpp_api_on_the_gpu():
    r = _run()
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS") == 16 and "FAIL" not in r.stdout
