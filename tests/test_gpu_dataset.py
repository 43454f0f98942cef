"""GPU test of the in-process batch driver: every record it writes from a
batched persistent launch is the record bin/rainbow_generate writes alone for
the same condition (NPY files byte for byte, metadata value for value), passes
the PhaseRecord integrity scan, and is verified, not traced again, on resume."""
from __future__ import annotations

import contextlib
import io
import json
import subprocess
from pathlib import Path

import pytest

from ice_crystal_phase_amd import analysis
from ice_crystal_phase_amd import dataset as ds
from ice_crystal_phase_amd.record import PhaseRecord

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
CLI = ROOT / "ice_crystal_phase_amd" / "bin" / "rainbow_generate"


def _config(out: Path, sphere: bool, stage: str = "incoherent") -> dict:
    coherent = stage != "incoherent"
    return {
        "output_root": str(out), "generator": str(CLI), "radius_mm": 0.5 if coherent else (0.75 if sphere else 1.25),
        "force_sphere": sphere, "temperature_c": 20.0, "pressure_pa": 101325.0,
        "wavelengths": ({"start_nm": 500, "stop_nm": 600, "step_nm": 100} if coherent else
                        {"start_nm": 450, "stop_nm": 650, "step_nm": 100}),
        "inclinations": {"min_degrees": -90, "max_degrees": 90, "count": 2 if (sphere or coherent) else 3,
                         "placement": "linear_angle_cell_centers"},
        "solver": {"grid": 65 if coherent else 256, "query_theta": 90 if coherent else 180,
                   "query_phi": 180 if coherent else 360, "cdf_theta": 90, "cdf_phi": 180,
                   "stage": stage, "device": 0, "focal_offsets": [0, 1, 2, 3] if coherent else [0, 0, 0, 0],
                   "allow_underresolved": False,
                   "maximum_coarsening_tv": 1.0, "cdf_max_l1": 1e-10, "cdf_max_lost_mass": 1e-12},
        "reserve_gib": 0,
    }


def _driver(cfg: Path, *args) -> str:
    print("driver:", cfg, *args, flush=True)
    with contextlib.redirect_stdout(io.StringIO()) as buf:
        assert ds.run(["--config", str(cfg), *args]) == 0
    return buf.getvalue()


@pytest.mark.parametrize("sphere,stage", [(False, "incoherent"), (True, "incoherent"), (False, "focal")],
                         ids=["drop", "sphere", "focal"])
def test_driver_records_equal_command_line_records(tmp_path, sphere, stage):
    if not CLI.exists():
        pytest.skip("bin/rainbow_generate is not built (make -C ice_crystal_phase_amd)")
    out = tmp_path / "data"
    raw = _config(out, sphere, stage)
    cfg = tmp_path / "config.json"
    cfg.write_text(json.dumps(raw))
    # batches of 4: the drop plan's 9 conditions take launches of 4, 4 and 1
    _driver(cfg, "--execute", "--batch-records", "4")
    c = ds.normalized_config(raw)
    plan = json.loads((out / "batch_plan.json").read_text())
    ni, nl = len(c["inclination_degrees"]), len(c["wavelength_nm"])
    for ii in range(ni):
        for ll in range(nl):
            rec = out / f"records/i{ii:04d}/w{ll:04d}"
            alone = tmp_path / f"alone_{ii}_{ll}"
            cmd = ds.cli_command(alone, c, ii, ll)
            print("run:", " ".join(cmd), flush=True)
            res = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
            assert res.returncode == 0, res.stderr
            for name in ds.ARRAYS:
                assert (rec / name).read_bytes() == (alone / name).read_bytes(), (rec, name)
            mine = json.loads((rec / "metadata.json").read_text())
            theirs = json.loads((alone / "metadata.json").read_text())
            assert list(mine) == list(theirs)
            assert mine == theirs, {k: (mine[k], theirs[k]) for k in mine if mine[k] != theirs.get(k)}
            with PhaseRecord(rec) as r:
                assert r.nt == 90 and r.np == 180
            if not sphere and ii == 1 and ll == 0 and stage == "incoherent":
                # the record comparison tool sees no difference between the two (at normal
                # incidence: elsewhere the FP32 query_frame_axis, written unnormalized as the
                # reference writes it, is refused by the tool's 2e-12 frame check, as there)
                report = tmp_path / f"compare_{sphere}"
                assert analysis.main(["compare", str(rec), str(alone), "--out", str(report), "--no-plots",
                                      "--allow-orientation-plan-difference"]) == 0
                cmp = json.loads((report / "comparison.json").read_text())
                assert cmp["tv"] == 0 and cmp["g_a_minus_b"] == 0
            receipt = json.loads((out / f"receipts/i{ii:04d}_w{ll:04d}.json").read_text())
            assert receipt["record"] == ds.verify_record(rec, plan, ii, ll)
    text = _driver(cfg, "--execute", "--resume")
    assert f"verified_existing={ni * nl}; pending=0" in text and "Dataset complete." in text
