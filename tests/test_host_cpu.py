"""The C++ host layer (GPU_Implementation5 / GPU_Partition / icy::Model /
SnapshotManager over the C-ABI) without a GPU: the library and driver are
built, export the reference's class surface, and the HDF5 raw-point reader
matches an independent reader (h5dump) of the reference's input file."""
import os
import subprocess

import numpy as np
import pytest

import _host as H

LIBDIR = os.path.join(H.ROOT, "icempm_multi_amd")
REF_H5 = "/root/reference/build/geometries_2d/2d_70k.h5"


def demangled_exports(lib):
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", lib], capture_output=True, text=True, check=True)
    return out.stdout


def test_host_library_exports_reference_surface():
    lib = os.path.join(LIBDIR, "libicempm_host.so")
    assert os.path.exists(lib) and os.path.exists(H.RUN)
    syms = demangled_exports(lib)
    for name in ["GPU_Implementation5::split_hssoa_into_partitions()", "GPU_Implementation5::allocate_arrays()",
                 "GPU_Implementation5::transfer_ponts_to_device()", "GPU_Implementation5::transfer_from_device()",
                 "GPU_Implementation5::update_constants()", "GPU_Implementation5::reset_grid()",
                 "GPU_Implementation5::reset_indenter_force_accumulator()", "GPU_Implementation5::p2g()",
                 "GPU_Implementation5::receive_halos()", "GPU_Implementation5::update_nodes(double)",
                 "GPU_Implementation5::g2p(bool, bool)", "GPU_Implementation5::receive_points()",
                 "GPU_Implementation5::record_timings(bool)", "GPU_Implementation5::synchronize()",
                 "GPU_Implementation5::initialize_and_enable_peer_access()",
                 "GPU_Partition::initialize(int, int)", "GPU_Partition::allocate(int, int)",
                 "GPU_Partition::update_nodes(double)", "GPU_Partition::normalize_timings(int)",
                 "icy::Model::Step()", "icy::Model::Prepare()", "icy::Model::Reset()",
                 "icy::Model::UnlockCycleMutex()", "icy::SnapshotManager::LoadRawPoints(",
                 "icy::SnapshotManager::SaveSnapshot(", "icy::SnapshotManager::ReadSnapshot(",
                 "icy::SimParams::ParseFile(", "icy::SimParams::ComputeHelperVariables()",
                 "HostSideSOA::RemoveDisabledAndSort(double, unsigned int)", "HostSideSOA::Allocate(unsigned int)"]:
        assert name in syms, name


def test_raw_points_hdf5_roundtrip(tmp_path):
    d = np.load(os.path.join(H.GOLDEN, "ref_2d_70k.npz"))
    flat = tmp_path / "in.bin"
    H.write_flat_raw(flat, d["x"], d["y"], d["llGrainIDs"], d["volume"][0])
    H.run("--convert-raw", flat, tmp_path / "pts.h5")
    H.run("--read-raw", tmp_path / "pts.h5", tmp_path / "back.bin")
    x, y, g, vol = H.read_flat_raw(tmp_path / "back.bin")
    assert np.array_equal(x, d["x"]) and np.array_equal(y, d["y"]) and np.array_equal(g, d["llGrainIDs"])
    assert vol == d["volume"][0]


@pytest.mark.skipif(not os.path.exists(REF_H5), reason="reference input file not present")
def test_raw_points_reader_matches_h5dump_fixture(tmp_path):
    """LoadRawPoints' reads of the reference's own input (2d_70k.h5) equal the
    fixture written from the same file by h5dump."""
    H.run("--read-raw", REF_H5, tmp_path / "r.bin")
    x, y, g, vol = H.read_flat_raw(tmp_path / "r.bin")
    d = np.load(os.path.join(H.GOLDEN, "ref_2d_70k.npz"))
    assert len(x) == 70534
    assert np.array_equal(x, d["x"]) and np.array_equal(y, d["y"]) and np.array_equal(g, d["llGrainIDs"])
    assert vol == d["volume"][0] == np.float32(2.5)


def test_driver_rejects_bad_arguments(tmp_path):
    r = subprocess.run([H.RUN, "cfg.json", "--frobnicate"], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 2
    r = subprocess.run([H.RUN, str(tmp_path / "missing.json")], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1 and "configuration file" in r.stderr


def test_soa_proxy_iterator(tmp_path):
    """ProxyPoint / SOAIterator (host_side_soa.h:18-74 of the reference): the
    reference's remove_if + sort through the iterator moves whole points and
    agrees with RemoveDisabledAndSort; FindFirstPointAtGridXIndex equals
    std::lower_bound and a linear scan."""
    exe = tmp_path / "soa_iterator_test"
    host = os.path.join(H.ROOT, "icempm_multi_amd", "host")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I" + host, os.path.join(H.ROOT, "tests", "cpp",
                    "soa_iterator_test.cpp"), "-L" + LIBDIR, "-licempm_host", "-Wl,-rpath," + LIBDIR,
                    "-Wl,-rpath,/opt/rocm/lib", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok: 0 failures" in r.stdout
