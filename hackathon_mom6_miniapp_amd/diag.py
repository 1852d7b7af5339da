"""Output of the mini-app's diagnostics (src/common/mom6_diag.F90).

* ``stats_line`` -- the line post_data_2d/3d print in output mode DIAG_STATS
  (format ``'(A,A,A,4ES14.6)'``, mom6_diag.F90:198-201, 252-255);
* ``dump_path`` / ``write_binary`` / ``read_binary`` -- the files
  write_binary_2d/3d write in mode DIAG_FILE (mom6_diag.F90:362-431): named
  ``<dir>/<name><call_count, 6 digits>.bin``, holding the int32 sizes
  ``ni, nj`` (and ``nk``) followed by the compute domain ``isc:iec, jsc:jec``
  (``, 1:nk``) as float64 in Fortran order, with no record markers (stream
  access).

The statistics themselves are computed on the device (``Model.diagnostics``,
the ``mom6h_diag_*`` entry points).
"""
from __future__ import annotations

import os

import numpy as np

STATS_FORMAT = "  Diag: {name} min/max/mean/rms: {values}"


def es14_6(x: float) -> str:
    """Fortran ES14.6 for exponents of at most two digits."""
    s = f"{x:14.6E}"
    if len(s) != 14:
        raise ValueError(f"{x!r} needs a three-digit exponent")
    return s


def stats_line(name: str, stats) -> str:
    """The Diag line of one field: min, max, mean, rms."""
    return STATS_FORMAT.format(name=name, values="".join(es14_6(x) for x in stats))


def compute_domain(cfg, field: np.ndarray) -> np.ndarray:
    """field(isc:iec, jsc:jec[, :]) of a data-domain array."""
    _, _, _, _, isc, iec, jsc, jec = cfg.bounds
    return field[isc - 1:iec, jsc - 1:jec]


def dump_path(output_dir: str, name: str, call_count: int) -> str:
    """write (filename, '(A,A,A,I6.6,A)') trim(dir), '/', trim(name), call_count, '.bin'"""
    return f"{output_dir.rstrip()}/{name.strip()}{call_count:06d}.bin"


def write_binary(path: str, field: np.ndarray) -> None:
    """The compute-domain ``field`` (2D or 3D) in the reference's dump layout."""
    a = np.asarray(field, dtype=np.float64)
    if a.ndim not in (2, 3):
        raise ValueError("a 2D or 3D field")
    with open(path, "wb") as f:
        f.write(np.asarray(a.shape, dtype="<i4").tobytes())
        f.write(np.asfortranarray(a).tobytes(order="F"))


def read_binary(path: str, ndim: int) -> np.ndarray:
    raw = open(path, "rb").read()
    shape = tuple(int(x) for x in np.frombuffer(raw, dtype="<i4", count=ndim))
    data = np.frombuffer(raw, dtype="<f8", offset=4 * ndim)
    if data.size != int(np.prod(shape)):
        raise ValueError(f"{path}: {data.size} values for the sizes {shape}")
    return data.reshape(shape, order="F")


def dump_field(model, field_name: str, name: str, output_dir: str, call_count: int) -> str:
    """Download a model field and write its compute domain as ``name``."""
    path = dump_path(output_dir, name, call_count)
    os.makedirs(output_dir, exist_ok=True)
    write_binary(path, compute_domain(model.cfg, model.get(field_name)))
    return path
