"""Data formats on either side of the operators.

* ``preprocess``: raw detector frames (uint16 or float32, stored [A,V,U] or
  [A,U,V]) plus dark and flat fields -> line integrals [A,U,V] float32, the
  projection format of ``project``/``backproject``.  The reference requires
  this to be done by the caller (README.md:97,108; docs/USAGE.md:247-251); here
  it is one HIP kernel (``cbct_preprocess``).
* ``to_auv``: the [A,V,U] -> [A,U,V] transpose alone (docs/USAGE.md:247-251).
* ``geometry_from_table``: the per-view geometry CSV convention of
  docs/USAGE.md:341 -> the eight geometry tensors:
  ``SAD_Rs -> sod``, ``DD -> sid``, ``delta_U/V`` (mm) ``-> det_u0/v0`` through
  ``detector_centers`` (operator.py:32-45), ``delta_src_x/y/z -> src_x/y/z``,
  ``delta_Beta`` (rad) added to the nominal angles.  ``ODD = SDD - SOD`` is not
  a ``sid`` input and is ignored.
"""

import csv
import math

import torch

from . import _native

CSV_COLUMNS = ("SAD_Rs", "DD", "delta_U", "delta_V", "delta_src_x", "delta_src_y", "delta_src_z", "delta_Beta")


def preprocess(raw, dark=None, flat=None, layout="AVU", min_transmission=1e-6, log=True):
    """Line integrals [A,U,V] float32 from raw frames on the GPU.

    raw: [A,V,U] (layout "AVU") or [A,U,V] (layout "AUV"), uint16 or float32.
    dark, flat: per-pixel fields in the raw frame layout ([V,U] or [U,V]); None = 0 / 1.
    With ``log``: -log(max((raw - dark) / (flat - dark), min_transmission)); a pixel with
    flat <= dark gives -log(min_transmission).  Without ``log``: conversion/transpose only."""
    if not isinstance(raw, torch.Tensor) or raw.device.type != "cuda" or raw.dim() != 3:
        raise ValueError("raw must be a 3-D GPU tensor")
    if raw.dtype == torch.uint16:
        dtype = 0
    elif raw.dtype == torch.float32:
        dtype = 1
    else:
        raise TypeError("raw must be uint16 or float32")
    if layout not in ("AVU", "AUV"):
        raise ValueError("layout must be 'AVU' or 'AUV'")
    raw = raw.contiguous()
    A, rows, cols = raw.shape
    fields = []
    for name, f in (("dark", dark), ("flat", flat)):
        if f is not None:
            if f.shape != (rows, cols) or f.device != raw.device:
                raise ValueError(f"{name} must have shape {(rows, cols)} on the raw tensor's device")
            f = f.to(torch.float32).contiguous()
        fields.append(f)
    transpose = layout == "AVU"
    out_shape = (A, cols, rows) if transpose else (A, rows, cols)
    out = torch.empty(out_shape, dtype=torch.float32, device=raw.device)
    L = _native.lib()
    stream = torch.cuda.current_stream(raw.device).cuda_stream
    _native._check(L.cbct_preprocess(raw.data_ptr(), dtype, fields[0].data_ptr() if fields[0] is not None else None,
                                     fields[1].data_ptr() if fields[1] is not None else None, A, rows, cols,
                                     1 if transpose else 0, float(min_transmission), 1 if log else 0,
                                     out.data_ptr(), stream))
    return out


def to_auv(projections_avu):
    """[A,V,U] float32 -> [A,U,V] float32 (the layout the operators take)."""
    return preprocess(projections_avu.to(torch.float32), layout="AVU", log=False)


def read_table(path):
    """CSV with a header row -> dict column -> list of floats."""
    with open(path, newline="") as f:
        rows = list(csv.DictReader(f))
    if not rows:
        raise ValueError("the geometry table is empty")
    return {k.strip(): [float(r[k]) for r in rows] for k in rows[0]}


def geometry_from_table(table, detector_u, detector_v, pixel_size, base_angles=None, device="cpu"):
    """Eight float32 [A] geometry tensors (dict in the operators' order) from per-view columns.

    ``table``: a CSV path or a dict of column -> sequence with the columns ``CSV_COLUMNS``.
    ``base_angles``: the nominal view angles [A] (radians); default a uniform full circle
    2*pi*k/A.  Missing delta columns count as zero; SAD_Rs and DD are required."""
    if isinstance(table, str):
        table = read_table(table)
    for need in ("SAD_Rs", "DD"):
        if need not in table:
            raise ValueError(f"the geometry table needs a {need} column")
    A = len(table["SAD_Rs"])
    col = lambda k: torch.tensor([float(x) for x in table.get(k, [0.0] * A)], dtype=torch.float64)
    if any(len(table[k]) != A for k in table if k in CSV_COLUMNS):
        raise ValueError("all geometry columns must have one value per view")
    if base_angles is None:
        base = torch.arange(A, dtype=torch.float64) * (2.0 * math.pi / A)
    else:
        base = torch.as_tensor(base_angles, dtype=torch.float64).reshape(-1)
        if base.numel() != A:
            raise ValueError("base_angles must have one value per view")
    if not (pixel_size > 0 and math.isfinite(pixel_size)):
        raise ValueError("pixel_size must be a positive finite number")
    geometry = {
        "angles": base + col("delta_Beta"),
        "sid": col("DD"),
        "sod": col("SAD_Rs"),
        "det_u0": (detector_u - 1.0) / 2.0 - col("delta_U") / pixel_size,
        "det_v0": (detector_v - 1.0) / 2.0 - col("delta_V") / pixel_size,
        "src_x": col("delta_src_x"),
        "src_y": col("delta_src_y"),
        "src_z": col("delta_src_z"),
    }
    return {k: v.to(device=device, dtype=torch.float32) for k, v in geometry.items()}
