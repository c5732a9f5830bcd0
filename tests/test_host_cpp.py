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


REF_CALLER = os.path.join(ROOT, "tests", "host_cpp", "ref_caller")


def _ref_caller():
    if not os.path.exists(REF_CALLER):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "d-core-cuda_amd"), REF_CALLER],
                       check=True)
    return REF_CALLER


def test_reference_signature_caller_builds():
    """tests/host_cpp/ref_caller.cpp, written against the reference's own
    global-namespace signatures (dcore(Graph&), KList(V, g_p, a_p, k),
    maintenance(g, g_p, m_p, a_p, edges, isDelete, inPlace), GraphGPU(g,
    g_p), maintenanceFast, the allocate*Memory family), compiles and links
    against dcore_ref.hpp / libdcore_host.so; without a GPU it stops loudly."""
    b = _ref_caller()
    assert os.access(b, os.X_OK)
    if not os.path.exists("/dev/kfd"):
        r = subprocess.run([b, os.path.join(GOLDEN, "digraph")], capture_output=True,
                           text=True, timeout=60)
        assert r.returncode != 0 and "DCX_ERR_NO_DEVICE" in (r.stdout + r.stderr)


@pytest.mark.gpu
def test_reference_signature_caller_on_golden(tmp_path):
    """The reference-style caller on the reference's graph: dcore(g) equals
    the reference's own .decomp, KList/checkKmax equal its rows, and random
    inserts, deletes and insert batches maintained through maintenance() and
    maintenanceFast() equal checkDCore's recomputation each time."""
    print("[case] reference-signature caller, golden digraph", flush=True)
    shutil.copy(os.path.join(GOLDEN, "digraph"), tmp_path / "graph")
    shutil.copy(os.path.join(GOLDEN, "digraph.decomp"), tmp_path / "golden.decomp")
    r = subprocess.run([_ref_caller(), str(tmp_path / "graph"), str(tmp_path / "golden")],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "compared with the golden decomposition" in r.stdout
    assert r.stdout.rstrip().endswith("found 0 errors")


@pytest.mark.gpu
def test_reference_signature_caller_on_generated(tmp_path):
    """The same caller on a seeded G(n,m) graph of 3,000 vertices and 24,000
    edges (more levels, hubs of degree ~20, lists that fill their gap)."""
    import dcx
    print("[case] reference-signature caller, generated graph", flush=True)
    s, d = dcx.gen_er(3000, 24000, 17)
    dcx.text_write(str(tmp_path / "er"), 3000, s, d)
    r = subprocess.run([_ref_caller(), str(tmp_path / "er")], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert r.stdout.rstrip().endswith("found 0 errors")
