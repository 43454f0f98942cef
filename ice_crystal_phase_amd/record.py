"""Read-only access to one phase-CDF record directory, with an integrity scan.

Drop-in for the reference's `PhaseRecord` (python/rainbow_dataset/record.py:11-158):
the same constructor (`validate=True` by default), the same accepted schemas and
fixed metadata values, the same checks and the same `mass_and_g` statistic. The
arrays are memory-mapped and returned exactly as stored: a record is never
rebuilt, renormalized, clamped or transposed here. A rejected record raises
ValueError and releases its mappings.
"""
from __future__ import annotations

import json
from pathlib import Path
from typing import Any

import numpy as np

SCHEMAS = ("rainbow.phase_cdf.numpy.v1", "rainbow.phase_cdf.numpy.v2")
# metadata values every record must carry (README.md:153-210 of the reference)
CONTRACT = (
    ("dtype", "<f8"),
    ("order", "C"),
    ("density_measure", "solid_angle_sr"),
    ("cell_model", "constant_density_per_spherical_cell"),
    ("coordinates", "u=(1-cos(theta))/2; v=(phi+pi)/(2*pi)"),
    ("normalization", "integral_p_domega_equals_one"),
    ("input_polarization", "unpolarized"),
)
HG_CONVENTION = (
    ("method", "first_moment_of_saved_cell_pdf"),
    ("target", "saved_cdf"),
    ("cosine_convention", "dot(incident_propagation,outgoing_propagation)"),
)
METADATA_LIMIT = 1 << 20
MASS_TOLERANCE = 1e-10
G_TOLERANCE = 5e-12


def _angular_count(value: Any) -> int:
    if type(value) is not int or value < 1 or value > 0x7FFFFFFF:
        raise ValueError("Invalid angular dimension")
    return value


def _positive_block(block_elements: Any) -> int:
    if type(block_elements) is not int or block_elements <= 0:
        raise ValueError("block_elements must be positive")
    return block_elements


class PhaseRecord:
    """One record: `metadata`, `phi_cdf` [Nphi+1], `theta_cdf` [Nphi, Ntheta+1],
    `u_edges` [Ntheta+1] (read-only memory maps), `frame` (3x3 sampling frame
    columns), `nt`, `np`. Use as a context manager or call close()."""

    def __init__(self, directory: str | Path, *, validate: bool = True) -> None:
        self.directory = Path(directory)
        self._maps: list[np.memmap] = []
        self._closed = False
        self._validated = False
        if self.directory.name.endswith(".part"):
            raise ValueError("An incomplete .part directory is not a dataset record")
        meta = self.directory / "metadata.json"
        if meta.stat().st_size > METADATA_LIMIT:
            raise ValueError("Unexpectedly large metadata")
        self.metadata: dict[str, Any] = json.loads(meta.read_text(encoding="utf-8"))
        m = self.metadata
        if m.get("schema") not in SCHEMAS or m.get("complete") is not True:
            raise ValueError("Unknown schema or incomplete record")
        for key, want in CONTRACT:
            if m.get(key) != want:
                raise ValueError(f"Unsupported {key}: {m.get(key)!r}")
        if m.get("cdf_axis_order") != ["phi_cell", "theta_edge"]:
            raise ValueError("Conditional CDF axes are not phi_cell, theta_edge")
        self.nt = _angular_count(m.get("theta_count"))
        self.np = _angular_count(m.get("phi_count"))
        try:
            self.phi_cdf = self._map("phi_cdf.npy", (self.np + 1,))
            self.theta_cdf = self._map("theta_given_phi_cdf.npy", (self.np, self.nt + 1))
            self.u_edges = self._map("u_edges.npy", (self.nt + 1,))
            self.frame = np.asarray(m["sampling_frame_columns"], dtype=np.float64)
            self._check_frame()
            if validate:
                self.validate()
        except BaseException:
            self.close()
            raise

    # -------------------------------------------------------------- lifetime
    def __enter__(self) -> "PhaseRecord":
        self._require_open()
        return self

    def __exit__(self, *_exc: Any) -> None:
        self.close()

    def close(self) -> None:
        """Release the memory maps; arrays taken from the record are invalid afterwards."""
        if self._closed:
            return
        for a in self._maps:
            a._mmap.close()
        self._maps.clear()
        self._closed = True

    def _require_open(self) -> None:
        if self._closed:
            raise RuntimeError("PhaseRecord has been closed")

    # -------------------------------------------------------------- loading
    def _map(self, name: str, shape: tuple[int, ...]) -> np.memmap:
        path = self.directory / name
        a = np.load(path, mmap_mode="r", allow_pickle=False)
        if isinstance(a, np.memmap):
            self._maps.append(a)
        ok = isinstance(a, np.memmap) and a.dtype.str == "<f8" and a.shape == shape and a.flags.c_contiguous
        if not ok:
            raise ValueError(f"Unexpected shape/dtype/order in {name}")
        if path.stat().st_size != a.offset + a.nbytes:
            raise ValueError(f"Truncated or trailing data in {name}")
        return a

    def _check_frame(self) -> None:
        f = self.frame
        ok = (f.shape == (3, 3) and bool(np.isfinite(f).all())
              and np.allclose(f.T @ f, np.eye(3), rtol=0, atol=2e-12)
              and abs(float(np.linalg.det(f)) - 1.0) <= 2e-12)
        if not ok:
            raise ValueError("Sampling frame must be a right-handed orthonormal frame")

    # -------------------------------------------------------------- checks
    def validate(self, block_elements: int = 1 << 20) -> None:
        """Bounded-memory scan: u edges strictly increasing 0 -> 1, marginal and
        conditional CDFs finite, in [0, 1], nondecreasing with exact 0/1 ends,
        and (schema v2) the stored HG g equal to the saved cells' first moment."""
        self._require_open()
        block = _positive_block(block_elements)
        u = self.u_edges
        if not (np.isfinite(u).all() and u[0] == 0 and u[-1] == 1 and np.all(u[1:] > u[:-1])):
            raise ValueError("u_edges are not strictly increasing from 0 to 1")
        p = self.phi_cdf
        if not (np.isfinite(p).all() and p[0] == 0 and p[-1] == 1 and np.all(p[1:] >= p[:-1])):
            raise ValueError("Marginal CDF is invalid")
        self._scan_conditional(block)
        if self.metadata["schema"].endswith(".v2"):
            hg = self.metadata.get("hg", {})
            if any(hg.get(k) != v for k, v in HG_CONVENTION):
                raise ValueError("Unsupported HG moment convention")
            g = hg.get("g")
            if type(g) not in (int, float) or not np.isfinite(g) or not -1 < g < 1:
                raise ValueError("Invalid HG first moment")
            mass, g_cells = self.mass_and_g(block_elements=block)
            if abs(mass - 1.0) > MASS_TOLERANCE or abs(g_cells - float(g)) > G_TOLERANCE:
                raise ValueError("HG label is inconsistent with saved CDF")
        self._validated = True

    def _scan_conditional(self, block: int) -> None:
        """Walk the conditional CDF as one flat array in blocks; a row may span
        blocks, so the first element of a block is compared with the last of
        the previous one unless it starts a row."""
        flat = self.theta_cdf.reshape(-1)
        stride = self.nt + 1
        prev = None
        for start in range(0, flat.size, block):
            a = np.asarray(flat[start:start + block])
            if not np.isfinite(a).all() or (a < 0).any() or (a > 1).any():
                raise ValueError("Conditional CDF contains nonfinite or out-of-range data")
            col = (np.arange(start, start + a.size, dtype=np.int64)) % stride
            if (a[col == 0] != 0).any() or (a[col == self.nt] != 1).any():
                raise ValueError("Conditional CDF endpoints must be 0 and 1")
            before = np.empty_like(a)
            before[1:] = a[:-1]
            before[0] = prev if prev is not None else a[0]
            inside = col != 0
            if (a[inside] < before[inside]).any():
                raise ValueError("Conditional CDF decreases; refusing to repair it")
            prev = a[-1]

    def mass_and_g(self, *, block_elements: int = 1 << 20) -> tuple[float, float]:
        """Total probability and mean cosine of the saved cells, in extended
        precision: cell mass = dPhi_j * dTheta_ij, cell mean cosine =
        1 - (u_i + u_{i+1}) (exact for a density constant in u)."""
        self._require_open()
        block = _positive_block(block_elements)
        mu = 1.0 - (self.u_edges[:-1] + self.u_edges[1:])
        mass = np.longdouble(0)
        first = np.longdouble(0)
        for j in range(self.np):
            w_phi = np.longdouble(self.phi_cdf[j + 1] - self.phi_cdf[j])
            row = self.theta_cdf[j]
            for i in range(0, self.nt, block):
                stop = min(i + block, self.nt)
                w = np.diff(row[i:stop + 1]).astype(np.longdouble) * w_phi
                mass += w.sum(dtype=np.longdouble)
                first += (w * mu[i:stop]).sum(dtype=np.longdouble)
        if not (np.isfinite(mass) and mass > 0 and np.isfinite(first)):
            raise ValueError("Invalid reconstructed probability mass")
        return float(mass), float(first / mass)
