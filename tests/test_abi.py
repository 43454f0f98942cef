"""C-ABI checks that need no GPU: the library loads, exports every function
include/icephase.h declares, agrees with the oracle on host-side arithmetic,
and writes record directories with the reference's on-disk contract."""
import ctypes
import json
import re
from pathlib import Path

import numpy as np
import pytest

import ice_crystal_phase_amd as ip
from ice_crystal_phase_amd import icephase

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "icephase.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(iph_[a-z_]+)\s*\(", text)))


def test_header_symbols_are_exported():
    lib = ip.lib()
    names = declared_functions()
    assert len(names) >= 17
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(icephase.EXPORTED) == names


def test_abi_version_and_struct_sizes():
    lib = ip.lib()
    assert lib.iph_abi_version() == 3
    import ctypes
    assert lib.iph_struct_size(0) == ctypes.sizeof(icephase.Record) == 264
    assert lib.iph_struct_size(1) == ctypes.sizeof(icephase.HistStats)
    assert lib.iph_struct_size(2) == ctypes.sizeof(icephase.CdfReport)
    assert lib.iph_struct_size(3) == ctypes.sizeof(icephase.CdfPolicy)
    assert lib.iph_struct_size(4) == ctypes.sizeof(icephase.WaterState)
    assert lib.iph_struct_size(5) == ctypes.sizeof(icephase.Area)
    assert lib.iph_struct_size(6) == ctypes.sizeof(icephase.IceState) == 48
    assert lib.iph_struct_size(7) == ctypes.sizeof(icephase.AtlasSettings)
    assert lib.iph_struct_size(8) == ctypes.sizeof(icephase.AtlasNode)
    assert lib.iph_struct_size(9) == ctypes.sizeof(icephase.CoherentConfig) == 72
    assert lib.iph_struct_size(10) == ctypes.sizeof(icephase.CoherentStats)
    assert lib.iph_struct_size(11) == icephase.DIRECTION_DTYPE.itemsize == 304
    assert lib.iph_struct_size(12) == ctypes.sizeof(icephase.CoherentGate)
    assert lib.iph_struct_size(13) == 0


def test_u_edges_equal_oracle(oracle):
    for n in (1, 2, 7, 36, 450, 900, 1800, 2047):
        assert np.array_equal(ip.u_edges(n), oracle.u_edges(n))
    with pytest.raises(ip.IcePhaseError):
        ip.u_edges(0)


def test_no_device_fails_loudly(monkeypatch):
    """Without a usable HIP device the library raises; it never computes on the CPU."""
    import ctypes
    p = ctypes.c_void_p()
    rc = ip.lib().iph_ctx_create(0, ctypes.byref(p))
    if rc == 0:
        ip.lib().iph_ctx_destroy(p)
        pytest.skip("a HIP device is present")
    assert rc == -2
    assert b"device" in ip.lib().iph_last_error(None).lower()


def npy_header(path: Path) -> str:
    raw = path.read_bytes()
    assert raw[:8] == b"\x93NUMPY\x01\x00"
    hlen = raw[8] | (raw[9] << 8)
    assert (10 + hlen) % 64 == 0   # src/npy_writer.cpp:37-42 padding rule
    return raw[10:10 + hlen].decode()


def test_write_record_contract(tmp_path):
    nt, np_ = 17, 13
    rng = np.random.default_rng(1)
    theta = np.sort(rng.random((np_, nt + 1)), axis=1)
    theta[:, 0], theta[:, -1] = 0.0, 1.0
    phi = np.concatenate([[0.0], np.sort(rng.random(np_ - 1)), [1.0]])
    meta = json.dumps({"schema": "rainbow.phase_cdf.numpy.v2", "complete": True, "theta_count": nt, "phi_count": np_})
    out = tmp_path / "rec"
    ip.write_record(out, meta, phi, theta)
    assert out.is_dir() and not (tmp_path / "rec.part").exists()
    assert npy_header(out / "phi_cdf.npy").rstrip() == "{'descr': '<f8', 'fortran_order': False, 'shape': (14, ), }"
    a = np.load(out / "phi_cdf.npy", allow_pickle=False)
    t = np.load(out / "theta_given_phi_cdf.npy", allow_pickle=False)
    u = np.load(out / "u_edges.npy", allow_pickle=False)
    assert a.dtype.str == "<f8" and a.shape == (np_ + 1,) and np.array_equal(a, phi)
    assert t.shape == (np_, nt + 1) and np.array_equal(t, theta)
    assert np.array_equal(u, ip.u_edges(nt))
    for f in ("phi_cdf.npy", "theta_given_phi_cdf.npy", "u_edges.npy"):
        h = npy_header(out / f)
        assert h.startswith("{'descr': '<f8', 'fortran_order': False, 'shape': (") and h.endswith("\n")
    assert json.loads((out / "metadata.json").read_text())["theta_count"] == nt
    # refuse to overwrite, refuse a leftover .part (src/npy_writer.cpp:97-121)
    with pytest.raises(ip.IcePhaseError):
        ip.write_record(out, meta, phi, theta)
    (tmp_path / "other.part").mkdir()
    with pytest.raises(ip.IcePhaseError):
        ip.write_record(tmp_path / "other", meta, phi, theta)
    with pytest.raises(ip.IcePhaseError):
        ip.write_record(tmp_path / "x.part", meta, phi, theta)


def test_npy_special_bit_patterns_roundtrip(tmp_path):
    """test_phase_numpy_cpu.cpp:50-55: eight binary64 patterns survive bitwise."""
    bits = np.array([0, 0x8000000000000000, 1, 0x3fefffffffffffff, 0x3ff0000000000000, 0x7fefffffffffffff,
                     0x7ff0000000000000, 0x7ff8000000001234], dtype=np.uint64)
    vals = bits.view(np.float64)
    theta = np.tile(vals, (2, 1))   # 2 x 8 -> nt = 7
    phi = vals[:3].copy()
    ip.write_record(tmp_path / "bits", "{}", phi, theta)
    back = np.load(tmp_path / "bits" / "theta_given_phi_cdf.npy", allow_pickle=False)
    assert np.array_equal(back.view(np.uint64), theta.view(np.uint64))


def test_atlas_entry_points_reject_bad_arguments_without_a_device():
    L = icephase.lib()
    s = icephase.atlas_settings(0.1, 0.2, 1.31, 3, 2, 12, 24, 2)
    mass = np.zeros(12 * 24)
    nodes = (icephase.AtlasNode * 1)()
    assert L.iph_atlas_trace(None, ctypes.byref(s), 0, 1, mass.ctypes.data, nodes) == -1
    out = ctypes.c_void_p()
    assert L.iph_atlas_mass_create(None, mass.ctypes.data, 1, 1, 12, 24, ctypes.byref(out)) == -1
    assert L.iph_atlas_compose(None, mass.ctypes.data, mass.ctypes.data) == -1
    L.iph_atlas_mass_destroy(None)
    assert L.iph_ice_index(550.0, None) == -1
    count = ctypes.c_uint64()
    node = icephase.AtlasNode()
    assert L.iph_atlas_node_rays(None, 0, 0, None, 0, ctypes.byref(count), ctypes.byref(node)) == -1
    # a too small ray buffer is refused, the count is still reported
    assert L.iph_atlas_node_rays(ctypes.byref(s), 1, 1, None, 0, ctypes.byref(count), ctypes.byref(node)) == 0
    small = np.zeros(5)
    assert L.iph_atlas_node_rays(ctypes.byref(s), 1, 1, small.ctypes.data, 1, ctypes.byref(count),
                                 ctypes.byref(node)) == -1
    assert count.value > 1


def test_cell_table_switch_rejects_a_missing_batch():
    assert icephase.lib().iph_batch_set_cell_table(None, 1) == -1


def test_write_npy_matches_numpy(tmp_path):
    data = np.arange(2 * 3 * 4, dtype=np.float64).reshape(2, 3, 4) / 7
    shape = (ctypes.c_uint64 * 3)(2, 3, 4)
    path = tmp_path / "a.npy"
    assert icephase.lib().iph_write_npy(str(path).encode(), shape, 3, data.ctypes.data) == 0
    raw = path.read_bytes()
    hlen = raw[8] | (raw[9] << 8)
    assert raw[:8] == b"\x93NUMPY\x01\x00" and (10 + hlen) % 64 == 0
    back = np.load(path, allow_pickle=False)
    assert back.shape == (2, 3, 4) and back.dtype == np.float64 and np.array_equal(back, data)
    assert len(raw) == 10 + hlen + data.nbytes
    assert icephase.lib().iph_write_npy(str(path).encode(), shape, 0, data.ctypes.data) == -1


def test_closed_handles_raise_before_the_library():
    """A closed batch or atlas mass raises its own message (the library's
    message for a NULL handle is the thread's, not the context's)."""
    from types import SimpleNamespace
    b = icephase.Batch.__new__(icephase.Batch)
    b._p, b.ctx, b.records, b.policy = ctypes.c_void_p(), SimpleNamespace(ptr=None), [icephase.Record()], None
    for call in (b.run, b.timing, b.stats, b.bins, b.cdf, lambda: b.set_schedule(256, 8)):
        with pytest.raises(ip.IcePhaseError, match="batch is closed"):
            call()
    a = icephase.AtlasMass.__new__(icephase.AtlasMass)
    a._p, a.ctx, a.shape = ctypes.c_void_p(), SimpleNamespace(ptr=None), (1, 1, 2, 3)
    with pytest.raises(ip.IcePhaseError, match="atlas mass is closed"):
        a.compose(np.ones((1, 1, 2, 3)))


def test_python_handles_take_their_arguments():
    """The handle classes keep their constructors (a dataclass decorator that
    slips to another class leaves Batch without arguments)."""
    import inspect
    assert list(inspect.signature(ip.Batch).parameters)[:3] == ["ctx", "records", "policy"]
    assert list(inspect.signature(ip.Coherent).parameters)[:4] == ["ctx", "radius_mm", "wavelength_nm",
                                                                    "interior_index"]
    assert list(inspect.signature(ip.AtlasMass).parameters) == ["ctx", "mass"]
