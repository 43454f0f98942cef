"""PhaseRecord: the read-only record reader and its integrity scan, after the
reference's tests of the same class (tests/test_phase_numpy.py:10-90), plus
the block-seam cases of the bounded-memory scan."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from ice_crystal_phase_amd.record import PhaseRecord


def fixture(path: Path, version: str = "v2") -> None:
    path.mkdir()
    edges = np.array([0, .1, .4, 1.], dtype=np.float64)
    phi = np.array([0, .2, 1.], dtype=np.float64)
    theta = np.array([[0, .1, .4, 1.], [0, .3, .3, 1.]], dtype=np.float64)
    g = np.sum(np.diff(phi)[:, None] * np.diff(theta, axis=1) * (1 - edges[:-1] - edges[1:]))
    m = dict(schema=f"rainbow.phase_cdf.numpy.{version}", complete=True, dtype="<f8", order="C", theta_count=3,
             phi_count=2, density_measure="solid_angle_sr", cell_model="constant_density_per_spherical_cell",
             coordinates="u=(1-cos(theta))/2; v=(phi+pi)/(2*pi)", normalization="integral_p_domega_equals_one",
             input_polarization="unpolarized", cdf_axis_order=["phi_cell", "theta_edge"],
             sampling_frame_columns=np.eye(3).tolist(),
             hg=dict(g=float(g), method="first_moment_of_saved_cell_pdf", target="saved_cdf",
                     cosine_convention="dot(incident_propagation,outgoing_propagation)"))
    for name, a in (("phi_cdf.npy", phi), ("theta_given_phi_cdf.npy", theta), ("u_edges.npy", edges)):
        np.save(path / name, a, allow_pickle=False)
    (path / "metadata.json").write_text(json.dumps(m), encoding="utf8")


@pytest.fixture
def rec(tmp_path):
    p = tmp_path / "data"
    fixture(p)
    return p


def _meta(p: Path, fn) -> None:
    m = json.loads((p / "metadata.json").read_text())
    fn(m)
    (p / "metadata.json").write_text(json.dumps(m))


def _array(p: Path, name: str, fn) -> None:
    a = np.load(p / name)
    a = fn(a)
    np.save(p / name, a, allow_pickle=False)


def test_v2(rec):
    with PhaseRecord(rec) as r:
        mass, g = r.mass_and_g(block_elements=1)
        assert mass == pytest.approx(1.0, abs=1e-15)
        assert g == pytest.approx(r.metadata["hg"]["g"], abs=1e-15)
        # cells by hand: phi weights .2/.8, theta steps, mean cosines .9/.5/-.4
        want = .2 * (.1 * .9 + .3 * .5 + .6 * -.4) + .8 * (.3 * .9 + 0 * .5 + .7 * -.4)
        assert g == pytest.approx(want, abs=1e-15)
        for attr in ("sample", "sample_uniforms", "to_torch"):
            assert not hasattr(r, attr)
    with pytest.raises(RuntimeError, match="closed"):
        r.mass_and_g()


def test_v1(rec):
    def v1(m):
        m["schema"] = "rainbow.phase_cdf.numpy.v1"
        del m["hg"]
    _meta(rec, v1)
    with PhaseRecord(rec) as r:
        assert r.mass_and_g()[0] == pytest.approx(1.0)


def test_label_rejected(rec):
    _meta(rec, lambda m: m["hg"].__setitem__("g", m["hg"]["g"] + .01))
    with pytest.raises(ValueError, match="HG label"):
        PhaseRecord(rec)


def test_label_not_checked_without_validate(rec):
    _meta(rec, lambda m: m["hg"].__setitem__("g", m["hg"]["g"] + .01))
    with PhaseRecord(rec, validate=False) as r:
        with pytest.raises(ValueError, match="HG label"):
            r.validate()


def test_monotonic(rec):
    _array(rec, "theta_given_phi_cdf.npy", lambda a: (a.__setitem__((1, 2), .1), a)[1])
    with pytest.raises(ValueError, match="decreases"):
        PhaseRecord(rec)


def test_decrease_across_a_block_seam(rec):
    _array(rec, "theta_given_phi_cdf.npy", lambda a: (a.__setitem__((1, 2), .1), a)[1])
    with PhaseRecord(rec, validate=False) as r:
        # elements 5 and 6 ([1,1] = .3, [1,2] = .1) fall in different blocks of 6
        with pytest.raises(ValueError, match="decreases"):
            r.validate(block_elements=6)
        with pytest.raises(ValueError, match="decreases"):
            r.validate(block_elements=1)


@pytest.mark.parametrize("block", [1, 2, 3, 4, 5, 7, 1 << 20])
def test_row_start_is_not_compared_with_previous_row(rec, block):
    # row 0 ends at 1, row 1 starts at 0: valid for every block size
    with PhaseRecord(rec, validate=False) as r:
        r.validate(block_elements=block)


def test_nonfinite(rec):
    _array(rec, "phi_cdf.npy", lambda a: (a.__setitem__(1, np.nan), a)[1])
    with pytest.raises(ValueError):
        PhaseRecord(rec)


def test_trailing(rec):
    with (rec / "u_edges.npy").open("ab") as f:
        f.write(b"bad")
    with pytest.raises(ValueError, match="trailing"):
        PhaseRecord(rec)


def test_float32(rec):
    _array(rec, "u_edges.npy", lambda a: a.astype(np.float32))
    with pytest.raises(ValueError):
        PhaseRecord(rec)


def test_partial(rec):
    dst = rec.with_name("data.part")
    rec.rename(dst)
    with pytest.raises(ValueError, match=r"\.part"):
        PhaseRecord(dst)


def test_bad_endpoints(rec):
    _array(rec, "phi_cdf.npy", lambda a: (a.__setitem__(-1, .9), a)[1])
    with pytest.raises(ValueError, match="Marginal"):
        PhaseRecord(rec)


def test_conditional_endpoint(rec):
    _array(rec, "theta_given_phi_cdf.npy", lambda a: (a.__setitem__((0, 3), .99), a)[1])
    with pytest.raises(ValueError, match="endpoints"):
        PhaseRecord(rec)


def test_bad_frame(rec):
    _meta(rec, lambda m: m["sampling_frame_columns"][0].__setitem__(0, 2))
    with pytest.raises(ValueError, match="frame"):
        PhaseRecord(rec)


def test_left_handed_frame(rec):
    _meta(rec, lambda m: m.__setitem__("sampling_frame_columns", (-np.eye(3)).tolist()))
    with pytest.raises(ValueError, match="right-handed"):
        PhaseRecord(rec)


def test_transposed_axes_rejected(rec):
    _meta(rec, lambda m: m.__setitem__("cdf_axis_order", ["theta_edge", "phi_cell"]))
    with pytest.raises(ValueError, match="axes"):
        PhaseRecord(rec)


def test_wrong_shape_rejected(rec):
    _meta(rec, lambda m: m.__setitem__("theta_count", 2))
    with pytest.raises(ValueError, match="shape"):
        PhaseRecord(rec)


def test_bad_block_size(rec):
    with PhaseRecord(rec) as r:
        for bad in (0, -1, 1.5):
            with pytest.raises(ValueError):
                r.validate(block_elements=bad)
