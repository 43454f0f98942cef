"""Batch-driver bookkeeping on the CPU: config validation, plan, lock, resume,
receipts and stop-at-first-failure, after the reference's orchestration tests
(tests/test_generate_phase_dataset.py:27-158). The device engine is replaced by
a stub that returns well-formed uniform-density results (as the reference's
tests replace the generator process); everything else runs for real,
including the host-side condition set-up and the record writer of
libicephase.so. The device engine itself is covered by the GPU tests."""
from __future__ import annotations

import contextlib
import io
import json
import math
import struct
import types
from pathlib import Path

import numpy as np
import pytest

import ice_crystal_phase_amd as ip
from ice_crystal_phase_amd import dataset as ds
from ice_crystal_phase_amd.record import PhaseRecord
import record_check

CONFIG = {
    "output_root": "unused", "generator": "bin/rainbow_generate", "radius_mm": 1.0, "force_sphere": False,
    "temperature_c": 20.0, "pressure_pa": 101325.0,
    "wavelengths": {"start_nm": 700, "stop_nm": 705, "step_nm": 5},
    "inclinations": {"min_degrees": -90, "max_degrees": 90, "count": 2, "placement": "linear_angle_cell_centers"},
    "solver": {"grid": 9, "query_theta": 4, "query_phi": 8, "cdf_theta": 2, "cdf_phi": 4, "stage": "incoherent",
               "device": 0, "focal_offsets": [0, 0, 0, 0], "allow_underresolved": False,
               "maximum_coarsening_tv": 1.0, "cdf_max_l1": 1e-10, "cdf_max_lost_mass": 1e-12},
    "reserve_gib": 0,
}


def write_npy(path: Path, shape, values):
    body = repr({"descr": "<f8", "fortran_order": False, "shape": shape})
    pad = (-(10 + len(body) + 1)) % 64
    header = (body + " " * pad + "\n").encode("latin1")
    path.write_bytes(b"\x93NUMPY\x01\x00" + struct.pack("<H", len(header)) + header +
                     struct.pack("<" + "d" * len(values), *values))


class Harness:
    """A config file, fake engine files, a stub device engine and a runner."""

    def __init__(self, tmp: Path, monkeypatch):
        self.tmp = tmp
        self.out = tmp / "data"
        self.raw = json.loads(json.dumps(CONFIG))
        self.raw["output_root"] = str(self.out)
        self.cfg = tmp / "config.json"
        self.save()
        self.engine = {name: tmp / "engine" / name for name in ("libicephase.so", "dataset.py", "icephase.py")}
        for name, p in self.engine.items():
            p.parent.mkdir(exist_ok=True)
            p.write_bytes(name.encode())
        self.calls: list[list[str]] = []
        self.fail_jobs: set[str] = set()
        monkeypatch.setattr(ds, "engine_files", lambda: dict(self.engine))
        monkeypatch.setattr(ds, "open_context", lambda device: types.SimpleNamespace(close=lambda: None))
        monkeypatch.setattr(ds, "trace_conditions", self.trace)

    def save(self):
        self.cfg.write_text(json.dumps(self.raw))

    def trace(self, ctx, conds, s):
        self.calls.append([c.job for c in conds])
        nt, np_ = s["cdf_theta"], s["cdf_phi"]
        u = ip.u_edges(nt)
        phi = np.linspace(0.0, 1.0, np_ + 1)
        theta = np.tile(u, (np_, 1))
        p = np.diff(phi)[:, None] * np.diff(theta, axis=1)
        g = float((p * (1 - (u[:-1] + u[1:]))[None, :]).sum())
        out = []
        for c in conds:
            rep = ip.CdfReport()
            rep.g_stored = rep.g_source = g
            rep.cdf_mass = 1.0
            rep.status = 3 if c.job in self.fail_jobs else 0
            st = ip.HistStats()
            st.qbits, st.ray_weight_mm2, st.rays = 40, 1e-3, s["grid"] ** 2
            out.append(ds.Traced(st, phi.copy(), theta.copy(), rep))
        return out

    def invoke(self, *args):
        with contextlib.redirect_stdout(io.StringIO()) as buf:
            rc = ds.run(["--config", str(self.cfg), *args])
        self.stdout = buf.getvalue()
        return rc

    def run_some(self, limit=None, *extra):
        return self.invoke("--execute", *(["--limit", str(limit)] if limit else []), *extra)

    def jobs(self):
        return [j for call in self.calls for j in call]


@pytest.fixture
def h(tmp_path, monkeypatch):
    return Harness(tmp_path, monkeypatch)


# ------------------------------------------------------------------ axes and config
def test_default_dry_run(h):
    assert h.invoke() == 0
    assert not h.out.exists() and h.calls == []
    assert "DRY RUN" in h.stdout and "rainbow_generate" in h.stdout


def test_wave_endpoints():
    w = ds.wavelengths(dict(start_nm=380, stop_nm=830, step_nm=5))
    assert len(w) == 91 and (w[0], w[-1]) == ("380", "830")


def test_decimal_steps():
    assert ds.wavelengths(dict(start_nm="700.1", stop_nm="700.3", step_nm="0.1")) == ["700.1", "700.2", "700.3"]


@pytest.mark.parametrize("step", [0, -1, 3])
def test_wave_bad_step(step):
    with pytest.raises(ValueError):
        ds.wavelengths(dict(start_nm=700, stop_nm=705, step_nm=step))


def test_fp32_alias():
    with pytest.raises(ValueError, match="FP32"):
        ds.wavelengths(dict(start_nm="700", stop_nm="700.000001", step_nm="0.000001"))


def test_angle_centers():
    v = ds.inclinations(dict(min_degrees=-90, max_degrees=90, count=900, placement="linear_angle_cell_centers"))
    assert len(v) == 900 and (v[0], v[-1]) == ("-89.9", "89.9")
    assert v[449] == "-0.1" and v[450] == "0.1"


def test_wrong_angle_domain():
    with pytest.raises(ValueError, match="NOT"):
        ds.inclinations(dict(min_degrees=0, max_degrees=180, count=900, placement="linear_angle_cell_centers"))


def test_grid_refuses_nonnested(h):
    h.raw["solver"]["cdf_phi"] = 3
    with pytest.raises(ValueError, match="divide"):
        ds.normalized_config(h.raw)


def test_unknown_config(h):
    h.raw["blur"] = 1
    with pytest.raises(ValueError, match="config keys"):
        ds.normalized_config(h.raw)


@pytest.mark.parametrize("stage", ["diffraction", "focal", "path"])
def test_coherent_stages_and_table_ii_range(h, stage):
    """The coherent stages are accepted; focal and diffraction need a Table II
    radius (0.1..1 mm), as rainbow_generate requires."""
    h.raw["solver"]["stage"] = stage
    h.raw["radius_mm"] = 0.5
    assert ds.normalized_config(h.raw)["solver"]["stage"] == stage
    h.raw["radius_mm"] = 1.5
    if stage == "path":
        assert ds.normalized_config(h.raw)["solver"]["stage"] == "path"
    else:
        with pytest.raises(ValueError, match="Table II covers radius 0.1..1.0 mm only"):
            ds.normalized_config(h.raw)


def test_table_ii_sigma_is_the_tools():
    """Table II values at and between the tabulated radii (one rounding, as an fma)."""
    assert ds.table_sigma_degrees(ds.fp32(0.5)) == 0.22
    assert ds.table_sigma_degrees(ds.fp32(1.0)) == 0.15
    assert ds.table_sigma_degrees(ds.fp32(0.05)) is None
    x = ds.table_sigma_degrees(ds.fp32(0.45))
    assert 0.22 < x < 0.25 and abs(x - 0.235) < 1e-6


def test_source_index_range(h):
    h.raw["solver"]["grid"] = 32768
    with pytest.raises(ValueError):
        ds.normalized_config(h.raw)
    h.raw["solver"]["grid"] = 32767
    assert ds.normalized_config(h.raw)["solver"]["grid"] == 32767


def test_nonfinite_json_rejected(h):
    h.cfg.write_text('{"a": NaN}')
    with pytest.raises(ValueError, match="Nonfinite"):
        ds.load_json(h.cfg)


# ------------------------------------------------------------------ runs
def test_all_jobs_and_resume(h):
    assert h.run_some() == 0
    assert h.jobs() == ["i0000_w0000", "i0000_w0001", "i0001_w0000", "i0001_w0001"]
    assert len(h.calls) == 1, "the four conditions share one persistent launch"
    progress = json.loads((h.out / "progress.json").read_text())
    assert progress["complete"] is True and progress["completed"] == 4
    plan = json.loads((h.out / "batch_plan.json").read_text())
    assert plan["version"] == ds.VERSION and set(plan["binary_sha256"]) == set(h.engine)
    for ii in range(2):
        for ll in range(2):
            rec = h.out / f"records/i{ii:04d}/w{ll:04d}"
            receipt = json.loads((h.out / f"receipts/i{ii:04d}_w{ll:04d}.json").read_text())
            assert receipt["signature"] == plan["signature"]
            assert receipt["record"] == ds.verify_record(rec, plan, ii, ll)
            assert "Saved NumPy CDF dataset" in (h.out / f"logs/i{ii:04d}_w{ll:04d}.log").read_text()
            record_check.check_record(rec)
            with PhaseRecord(rec) as r:
                assert r.nt == 2 and r.np == 4
    h.calls.clear()
    assert h.invoke("--execute", "--resume") == 0
    assert h.calls == [], "verified records are not traced again"
    assert not (h.out / ".batch.lock").exists()


def test_batches_follow_plan_order(h):
    assert h.run_some(None, "--batch-records", "3") == 0
    assert h.calls == [["i0000_w0000", "i0000_w0001", "i0001_w0000"], ["i0001_w0001"]]


def test_limit_resume(h):
    h.run_some(1)
    assert not json.loads((h.out / "progress.json").read_text())["complete"]
    h.invoke("--execute", "--resume")
    assert h.jobs() == ["i0000_w0000", "i0000_w0001", "i0001_w0000", "i0001_w0001"]
    assert json.loads((h.out / "progress.json").read_text())["complete"]


def test_no_automatic_resume(h):
    h.run_some(1)
    with pytest.raises(FileExistsError):
        h.invoke("--execute")


def test_engine_change_rejected(h):
    h.run_some(1)
    h.engine["libicephase.so"].write_bytes(b"different library")
    with pytest.raises(ValueError, match="Resume rejected"):
        h.invoke("--execute", "--resume")


def test_configuration_change_rejected(h):
    h.run_some(1)
    h.raw["pressure_pa"] = 100000
    h.save()
    with pytest.raises(ValueError, match="Resume rejected"):
        h.invoke("--execute", "--resume")


def _edit_metadata(h, fn):
    p = h.out / "records/i0000/w0000/metadata.json"
    m = json.loads(p.read_text())
    fn(m)
    p.write_text(json.dumps(m))


def test_truncated_npy_rejected(h):
    h.run_some(1)
    p = h.out / "records/i0000/w0000/u_edges.npy"
    p.write_bytes(p.read_bytes()[:-1])
    with pytest.raises(ValueError):
        h.invoke("--execute", "--resume")


def test_metadata_changed_rejected(h):
    h.run_some(1)
    _edit_metadata(h, lambda m: m["geometry"].__setitem__("projected_area_mm2", -1))
    with pytest.raises(ValueError):
        h.invoke("--execute", "--resume")


def test_frame_transpose_rejected(h):
    h.run_some(1)
    _edit_metadata(h, lambda m: m.__setitem__("sampling_frame_columns",
                                              [list(r) for r in zip(*m["sampling_frame_columns"])]))
    with pytest.raises(ValueError):
        h.invoke("--execute", "--resume")


def test_receipt_content_mismatch_rejected(h):
    h.run_some(1)
    # a valid but different metadata file: the receipt's hash no longer matches
    _edit_metadata(h, lambda m: m.__setitem__("source_commit", "edited"))
    with pytest.raises(ValueError, match="receipt/content mismatch"):
        h.invoke("--execute", "--resume")


def test_failed_record_stops(h):
    h.fail_jobs = {"i0000_w0000"}
    with pytest.raises(RuntimeError, match="stopped at i0000_w0000"):
        h.invoke("--execute")
    assert len(h.calls) == 1
    assert not (h.out / ".batch.lock").exists()
    assert not (h.out / "receipts").exists()
    assert not (h.out / "records/i0000/w0000").exists()
    assert "CDF audit failed" in (h.out / "logs/i0000_w0000.log").read_text()


def test_failure_inside_a_batch_keeps_earlier_records(h):
    h.fail_jobs = {"i0001_w0000"}
    with pytest.raises(RuntimeError, match="stopped at i0001_w0000"):
        h.invoke("--execute")
    assert sorted(p.name for p in (h.out / "receipts").iterdir()) == ["i0000_w0000.json", "i0000_w0001.json"]
    assert not (h.out / "records/i0001/w0000").exists() and not (h.out / "records/i0001/w0001").exists()
    # the failed condition is traced again on resume, never skipped
    h.fail_jobs = set()
    h.calls.clear()
    assert h.invoke("--execute", "--resume") == 0
    assert h.jobs() == ["i0001_w0000", "i0001_w0001"]


def test_incomplete_directory_preserved(h):
    h.run_some(1)
    p = h.out / "records/i0000/w0001.part"
    p.mkdir()
    (p / "keep").write_text("do not delete")
    with pytest.raises(ValueError, match=r"\.part"):
        h.invoke("--execute", "--resume")
    assert (p / "keep").exists()


def test_lock_preserved(h):
    h.out.mkdir()
    lock = h.out / ".batch.lock"
    lock.write_text("another writer")
    with pytest.raises(RuntimeError, match="lock"):
        h.invoke("--execute")
    assert lock.read_text() == "another writer"


def test_explicit_underresolved_flag_recorded(h):
    h.invoke("--execute", "--allow-underresolved", "--limit", "1")
    m = json.loads((h.out / "records/i0000/w0000/metadata.json").read_text())
    assert m["quality"]["allow_underresolved"] is True
    plan = json.loads((h.out / "batch_plan.json").read_text())
    assert plan["conditions"]["solver"]["allow_underresolved"] is True


def test_missing_receipt_not_adopted(h):
    h.run_some(1)
    (h.out / "receipts/i0000_w0000.json").unlink()
    with pytest.raises(ValueError, match="without authorized resume/receipt"):
        h.invoke("--execute", "--resume")


def test_receipt_without_record_rejected(h):
    h.run_some(1)
    (h.out / "receipts/i0000_w0001.json").write_text("{}")
    with pytest.raises(ValueError, match="record missing"):
        h.invoke("--execute", "--resume")


def test_missing_engine_file(h):
    h.engine["libicephase.so"].unlink()
    with pytest.raises(FileNotFoundError):
        h.invoke("--execute")


def test_nonempty_directory_not_adopted(h):
    h.out.mkdir()
    (h.out / "stray.txt").write_text("x")
    with pytest.raises(ValueError, match="nonempty"):
        h.invoke("--execute")


def test_no_disk_space(h, monkeypatch):
    monkeypatch.setattr(ds.shutil, "disk_usage", lambda p: types.SimpleNamespace(free=0))
    with pytest.raises(RuntimeError, match="free space"):
        h.invoke("--execute")
    assert not (h.out / "records").exists()


def test_wrong_npy_axis(h):
    h.run_some(1)
    write_npy(h.out / "records/i0000/w0000/theta_given_phi_cdf.npy", (3, 4), [0.0] * 12)
    with pytest.raises(ValueError, match="layout"):
        h.invoke("--execute", "--resume")


# ------------------------------------------------------------------ host set-up of a condition
def test_condition_setup_matches_the_command_line_tool(h, tmp_path):
    """The host values rainbow_generate derives for a condition: FP32 radius and
    wavelength, IAPWS water at the FP32 wavelength, Table I shape, FP32
    incident direction, certified projected area, grid half extent."""
    c = ds.normalized_config(h.raw)
    cond = ds.resolve(c, 1, 1, tmp_path)
    assert cond.job == "i0001_w0001" and cond.alpha == 45.0
    assert cond.wavelength == float(np.float32(705.0)) and cond.radius == 1.0
    w = ip.water_optics(705.0, 20.0, 101325.0)
    assert cond.n_solver == float(np.float32(w.refractive_index))
    assert cond.shape == ip.drop_shape(1.0) and not cond.sphere
    a = math.radians(45.0)
    assert cond.kf == [float(np.float32(math.cos(a))), float(np.float32(-math.sin(a))), 0.0]
    area = ip.drop_projected_area(cond.shape, 1.0, cond.kf)
    assert cond.area.area_mm2 == area.area_mm2
    r = ds.device_record(cond, c["solver"])
    assert (r.shape, r.grid_w, r.grid_h, r.theta_count, r.phi_count) == (ip.SHAPE_DROP, 9, 9, 4, 8)
    assert r.max_internal_hits == 8 and r.tail_tolerance == 0.0 and r.rays == 0
    # grid half extent and reference distances (drop units): the values the
    # reference's RaindropSettings::make_config gives (compiled into oracle/_ref)
    import coherent_ref as cr
    h.raw["force_sphere"] = True
    sph = ds.resolve(ds.normalized_config(h.raw), 0, 0, tmp_path)
    assert sph.sphere and sph.shape == [0.0] * 8
    if cr.available():
        for cnd, sphere in ((cond, False), (sph, True)):
            tc = np.zeros(16, np.float32)
            assert cr.lib().refc_trace_config(cr.config(radius_mm=1.0, wavelength_nm=705.0, sphere=sphere),
                                              tc.ctypes.data) == 0
            assert ds.half_extent(cnd) == float(tc[6])
            assert ds.reference_distance(cnd) == float(tc[7]) == float(tc[8])


def test_frame_is_right_handed_and_follows_inclination():
    for alpha in (-89.9, -30.0, 0.1, 45.0, 89.9):
        a = math.radians(alpha)
        kf = [float(np.float32(math.cos(a))), float(np.float32(-math.sin(a))), 0.0]
        e, f, k = ds._frame(kf)
        m = np.array([e, f, k]).T
        assert np.allclose(m.T @ m, np.eye(3), atol=2e-15) and abs(np.linalg.det(m) - 1) < 2e-15
        assert e == [0.0, 0.0, 1.0]  # +z is orthogonal to every incident direction in the x-y plane
        assert abs(k[0] - math.cos(a)) < 2e-7 and abs(k[1] + math.sin(a)) < 2e-7


def test_pipelined_launches_write_in_plan_order_and_stop_cleanly(h):
    """With one condition per launch the next launch is traced while a record
    is written; records still appear in plan order, and a failure leaves no
    record of the launch already in flight."""
    assert h.run_some(None, "--batch-records", "1") == 0
    assert h.jobs() == ["i0000_w0000", "i0000_w0001", "i0001_w0000", "i0001_w0001"]
    stamps = [(h.out / f"receipts/{j}.json").stat().st_mtime_ns for j in h.jobs()]
    assert stamps == sorted(stamps)


def test_pipelined_failure_discards_the_launch_in_flight(h):
    h.fail_jobs = {"i0000_w0001"}
    with pytest.raises(RuntimeError, match="stopped at i0000_w0001"):
        h.invoke("--execute", "--batch-records", "1")
    assert sorted(p.name for p in (h.out / "receipts").iterdir()) == ["i0000_w0000.json"]
    assert not (h.out / "records/i0000/w0001").exists() and not (h.out / "records/i0001").exists()
    assert not (h.out / ".batch.lock").exists()
