"""The C++ host API (d-core-cuda_amd/host/dcore.hpp, libdcore_host.so): the
reference's Graph / GraphGPU / dcore / KList / checkKmax / checkDCore /
maintenance / file functions, driven by tests/host_cpp/dcore_host_check the
way the reference's main.cu drives its own (main.cu:1820-1925).

  * files (CPU): the reference's text file and binary cache load to the same
    CSR; its .decomp file reads and writes back byte for byte; the seeded
    generator, edge batches and host-side edits hold their contracts.
  * device (GPU): dcore() on the reference's graph equals the reference's
    own .decomp; KList / KListInplace / checkKmax equal its rows and
    kmaxes; variants agree on a generated graph; maintenance of inserts and
    deletes equals checkDCore's recomputation; a bad delete throws and
    leaves the state unchanged.
"""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

BIN = os.path.join(ROOT, "tests", "host_cpp", "dcore_host_check")


def _driver():
    if not os.path.exists(BIN):  # built with the libraries (make -C d-core-cuda_amd)
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "d-core-cuda_amd"), BIN],
                       check=True)
    return BIN


def _golden_copy(tmp_path):
    for f in ("digraph", "digraph-binary-g=1", "digraph.decomp"):
        shutil.copy(os.path.join(GOLDEN, f), tmp_path / f)
    return str(tmp_path)


def _run(mode, d):
    r = subprocess.run([_driver(), mode, d], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    return r


def test_host_api_files(tmp_path):
    r = _run("files", _golden_copy(tmp_path))
    assert r.returncode == 0, r.stdout
    assert " 0 failed" in r.stdout


def test_host_api_without_device_fails_loudly(tmp_path):
    """Without a GPU the device flow stops at the context with
    DCX_ERR_NO_DEVICE (there is no CPU fallback)."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a device node is present")
    r = _run("device", _golden_copy(tmp_path))
    assert r.returncode != 0 and "DCX_ERR_NO_DEVICE" in r.stdout


@pytest.mark.gpu
def test_host_api_device(tmp_path):
    print("[case] C++ host API, device flow", flush=True)
    r = _run("device", _golden_copy(tmp_path))
    assert r.returncode == 0, r.stdout
    assert " 0 failed" in r.stdout
