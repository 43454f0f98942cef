"""CPU tests of the spectral atlas runner (spectrum.py), after the reference's
hex_atlas/tests/spectral_cli_test.py: plan refusals, the dry run, the atlas
cache with job receipts and its reuse rules, and the hand-over to the
composer. The atlas tool is replaced by a stand-in script that writes a
well-formed atlas directory, and the composer by a recorder: the device parts
(bin/ice_precompute_atlas, compose.run) have their own GPU tests."""
from __future__ import annotations

import json
import stat
import sys
from pathlib import Path

import pytest

from ice_crystal_phase_amd import spectrum as sp

PLAN = {"wavelengths_nm": [450, 550.5, 650],
        "precompute": {"radius_mm": .1, "length_mm": .2, "incident_theta_intervals": 3, "incident_phi_intervals": 2,
                       "theta": 12, "phi": 24, "area_subdivisions": 2, "batch_rays": 8, "max_work_mib": 32},
        "composition": {"inclinations_deg": [-90, 0, 35, 90], "orientation": {
            "quadrature": {"azimuth": 4, "tilt": 4, "spin": 3},
            "components": [{"kind": "horizontal_gaussian", "sigma_deg": 3, "fraction": .9},
                           {"kind": "isotropic", "fraction": .1}]}}}

STAND_IN = '''#!{python}
"""Stand-in atlas tool: a well-formed atlas directory for the given flags."""
import json, sys
from pathlib import Path
import numpy as np
a = dict(zip(sys.argv[1::2], sys.argv[2::2]))
if float(a["--wavelength-nm"]) == {fail}:
    Path(a["--out"] + ".part").mkdir()
    sys.exit(1)
out = Path(a["--out"])
out.mkdir()
it, ip_, nt, np_ = (int(a[k]) for k in ("--incident-theta-intervals", "--incident-phi-intervals", "--theta", "--phi"))
np.save(out / "mass_mm2.npy", np.zeros((it + 1, ip_ + 1, nt, np_)))
np.save(out / "u_edges.npy", np.linspace(0, 1, nt + 1))
meta = dict(schema="ice.single_particle_atlas.v1", complete=True, incident_theta_intervals=it,
            incident_phi_intervals=ip_, theta_count=nt, phi_count=np_,
            geometry=dict(circumradius_mm=float(a["--radius-mm"]), full_length_mm=float(a["--length-mm"])),
            area_quadrature=dict(subdivisions=int(a["--area-subdivisions"])),
            material=dict(model="ice_ih_warren_brandt_2008_real_only", automatic_ice_dispersion=True,
                          wavelength_nm=json.loads(a["--wavelength-nm"]), exterior_index=1.0))
(out / "metadata.json").write_text(json.dumps(meta))
'''


def _plan(tmp_path: Path, name: str = "plan.json", **changes) -> Path:
    plan = json.loads(json.dumps(PLAN))
    for key, value in changes.items():
        section, _, field = key.partition("__")
        if field:
            plan[section][field] = value
        else:
            plan[section] = value
    path = tmp_path / name
    path.write_text(json.dumps(plan))
    return path


def _stand_in(tmp_path: Path, fail: float = -1.0) -> Path:
    exe = tmp_path / "atlas_tool"
    exe.write_text(STAND_IN.format(python=sys.executable, fail=fail))
    exe.chmod(exe.stat().st_mode | stat.S_IXUSR)
    return exe


@pytest.mark.parametrize("changes,message", [
    ({"wavelengths_nm": [450, 450]}, "Unique vacuum wavelengths"),
    ({"wavelengths_nm": [150]}, "Unique vacuum wavelengths"),
    ({"wavelengths_nm": [650, 450]}, "increasing order"),
    ({"precompute__theta": 0}, "theta must be a positive integer"),
    ({"precompute__tail_tolerance": -1.0}, "finite nonnegative"),
    ({"precompute__unknown": 1}, "Missing/unknown precompute parameter"),
    ({"precompute__radius_mm": 0}, "Particle dimensions"),
    ({"precompute__batch_rays": 1 << 20}, "max_work_mib"),
    ({"composition__inclinations_deg": [0, 0]}, "Duplicate incident conditions"),
    ({"composition": {"inclinations_deg": [0]}}, "inclination angles and one orientation model"),
])
def test_plan_refusals(tmp_path, changes, message):
    with pytest.raises(ValueError, match=message):
        sp.read_plan(_plan(tmp_path, **changes))


def test_atlas_names():
    assert sp.atlas_name(550) == "lambda_550_nm"
    assert sp.atlas_name(550.5) == "lambda_550p5_nm"
    assert sp.atlas_name(0.1 + 0.2 + 400) == "lambda_400.30000000000001".replace(".", "p") + "_nm"


def test_dry_run_creates_nothing(tmp_path, capsys):
    exe = _stand_in(tmp_path)
    sp.main(["--exe", str(exe), "--plan", str(_plan(tmp_path)), "--atlas-root", str(tmp_path / "atlases"),
             "--out", str(tmp_path / "dry"), "--dry-run"])
    text = capsys.readouterr().out
    assert "PLAN: 3 wavelengths x 4 incidents" in text and "DRY RUN" in text
    assert text.count("--material ice --wavelength-nm") == 3 and "--area-subdivisions 2" in text
    assert not (tmp_path / "atlases").exists() and not (tmp_path / "dry").exists()


def test_jobs_receipts_reuse_and_composition(tmp_path, monkeypatch, capsys):
    calls = []
    monkeypatch.setattr(sp.compose, "run", lambda paths, cfg, out, output_contract, device: calls.append(
        (list(paths), json.loads(Path(cfg).read_text()), out, output_contract)))
    exe, plan, root = _stand_in(tmp_path), _plan(tmp_path), tmp_path / "atlases"
    sp.run(exe, plan, root, tmp_path / "out")
    names = ["lambda_450_nm", "lambda_550p5_nm", "lambda_650_nm"]
    for name, wave in zip(names, PLAN["wavelengths_nm"]):
        receipt = json.loads((root / (name + ".job.json")).read_text())
        assert receipt["complete"] is True and receipt["request"]["wavelength_nm"] == wave
        assert receipt["request"]["build"]["executable_sha256"] == sp.compose.sha256_file(exe)
        assert (root / (name + ".log")).exists()
    assert calls[0][0] == [root / n for n in names]
    assert calls[0][1] == PLAN["composition"] and calls[0][3] == "rainbow"
    # an existing atlas is refused without --reuse, and reused after checks with it
    with pytest.raises(FileExistsError, match="use --reuse"):
        sp.run(exe, plan, root, tmp_path / "out2")
    sp.run(exe, plan, root, tmp_path / "out2", reuse=True)
    assert capsys.readouterr().out.count("REUSE") == 3 and len(calls) == 2
    # changed metadata or a changed request is refused
    meta = root / "lambda_550p5_nm" / "metadata.json"
    meta.write_text(meta.read_text() + " ")
    with pytest.raises(ValueError, match="metadata changed"):
        sp.run(exe, plan, root, tmp_path / "out3", reuse=True)
    with pytest.raises(ValueError, match="arguments/build differ"):
        sp.run(exe, _plan(tmp_path, "other.json", precompute__max_internal_hits=9), root, tmp_path / "out4",
               reuse=True)
    (tmp_path / "out5.part").mkdir()
    with pytest.raises(ValueError, match="new CDF output directory"):
        sp.run(exe, plan, root, tmp_path / "out5", reuse=True)


def test_failed_job_keeps_part_and_stops(tmp_path, monkeypatch, capsys):
    monkeypatch.setattr(sp.compose, "run", lambda *a, **k: pytest.fail("composition after a failed job"))
    exe, root = _stand_in(tmp_path, fail=550.5), tmp_path / "atlases"
    with pytest.raises(SystemExit) as e:
        sp.main(["--exe", str(exe), "--plan", str(_plan(tmp_path)), "--atlas-root", str(root),
                 "--out", str(tmp_path / "out")])
    assert e.value.code == 2
    assert "SPECTRAL GENERATION FAILED" in capsys.readouterr().err
    assert (root / "lambda_450_nm").is_dir() and (root / "lambda_550p5_nm.part").is_dir()
    assert not (root / "lambda_650_nm").exists()
    # the unfinished job blocks a retry until someone inspects it
    with pytest.raises(FileExistsError, match="Unfinished job"):
        sp.run(exe, _plan(tmp_path), root, tmp_path / "out", reuse=True)


def test_cached_atlas_of_another_job_is_refused(tmp_path):
    exe, root = _stand_in(tmp_path), tmp_path / "atlases"
    root.mkdir()
    sp.subprocess.run([str(exe), "--out", str(root / "x"), "--wavelength-nm", "550", "--radius-mm", ".1",
                       "--length-mm", ".2", "--incident-theta-intervals", "3", "--incident-phi-intervals", "2",
                       "--theta", "12", "--phi", "24", "--area-subdivisions", "2"], check=True)
    s = PLAN["precompute"]
    assert sp.check_atlas(root / "x", 550, s)["complete"] is True
    with pytest.raises(ValueError, match="dispersive ice"):
        sp.check_atlas(root / "x", 551, s)
    with pytest.raises(ValueError, match="angular grid"):
        sp.check_atlas(root / "x", 550, dict(s, theta=6))
    with pytest.raises(ValueError, match="geometry"):
        sp.check_atlas(root / "x", 550, dict(s, area_subdivisions=3))
