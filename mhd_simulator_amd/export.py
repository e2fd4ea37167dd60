"""Frame export of accepted field packets (RunState::export_fields, src/run/export.rs:9-60, in the
format docs/RUNS.md "Headless field output" describes).

A frame is a new directory holding one ``{physics}.bin`` per field - the 16-float display packets
(the FluidCell packet mhdp_fields_dev writes, the EM and thermal packets of the caller), cells x
fastest, then y, then z, native byte order - and ``metadata.json``.  The metadata is published
last: it is written as ``metadata.writing`` and renamed once every field file is complete, so an
incomplete export has no completed metadata.  The directory is created exclusively (earlier
exports and concurrent writers are protected) and removed again when anything fails.

    from mhd_simulator_amd.export import export_frame, read_frame
    export_frame("out/frame-00000001", grid.dims, grid.h, fields={"fluid": cells}, step=1, time_s=dt)
    meta, fields = read_frame("out/frame-00000001")
"""
from __future__ import annotations

import json
import os
import shutil
import sys

import numpy as np

PHYSICS = ("em", "fluid", "thermal")  # the reference's field order
COMPONENTS = 16                       # f32 values per cell of every packet
FORMAT_VERSION = 1


def _packet_bytes(name, packets, cells):
    """The packets' bytes on the host (a device tensor is read back synchronously)."""
    if hasattr(packets, "data_ptr"):
        packets = packets.detach().to("cpu").numpy()
    arr = np.asarray(packets)
    if arr.dtype != np.float32 or arr.size != COMPONENTS * cells:
        raise ValueError(f"{name} packets must be float32, {COMPONENTS} per cell ({COMPONENTS * cells} values)")
    return np.ascontiguousarray(arr).reshape(-1).tobytes()


def export_frame(directory, dims, cell_m, *, fields, step, time_s, origin_m=(0.0, 0.0, 0.0), metadata=None):
    """Write one frame; returns the path of its metadata.json.

    fields: {"em" | "fluid" | "thermal": packets} (numpy or device tensor, float32, 16 per cell; at
    least one).  metadata: further provenance the caller records (run, fluid and EM configurations,
    materials, geometry path, energies ...), merged into the record; it may not redefine the
    format's own keys.  Raises FileExistsError when the directory exists (it is left untouched)."""
    dims = tuple(int(n) for n in dims)
    if len(dims) != 3 or min(dims) <= 0:
        raise ValueError("dims are three positive cell counts")
    cells = dims[0] * dims[1] * dims[2]
    unknown = set(fields) - set(PHYSICS)
    if unknown or not fields:
        raise ValueError(f"fields are a non-empty subset of {PHYSICS}")
    # read everything back before touching the disk: a bad packet leaves nothing behind
    data = [(name, _packet_bytes(name, fields[name], cells)) for name in PHYSICS if name in fields]
    record = {
        "format_version": FORMAT_VERSION,
        "step": int(step),
        "time_s": float(time_s),
        "origin_m": [float(v) for v in origin_m],
        "cell_m": float(cell_m),
        "dims": list(dims),
        "byte_order": sys.byteorder,
        "cell_order": "x fastest, then y, then z",
        "files": [{"physics": name, "file": f"{name}.bin", "bytes": len(b), "f32_components_per_cell": COMPONENTS}
                  for name, b in data],
    }
    for key, value in (metadata or {}).items():
        if key in record:
            raise ValueError(f"metadata may not redefine {key!r}")
        record[key] = value
    text = json.dumps(record, indent=2)  # fails here (nothing written) on values JSON cannot hold
    parent = os.path.dirname(os.path.abspath(directory))
    os.makedirs(parent, exist_ok=True)
    try:
        os.mkdir(directory)  # exclusive creation
    except FileExistsError as exc:
        raise FileExistsError(f"Create new frame directory {directory}: {exc}") from None
    try:
        for name, b in data:
            with open(os.path.join(directory, f"{name}.bin"), "wb") as f:
                f.write(b)
        staged = os.path.join(directory, "metadata.writing")
        with open(staged, "w") as f:
            f.write(text)
        os.rename(staged, os.path.join(directory, "metadata.json"))
    except BaseException:
        shutil.rmtree(directory, ignore_errors=True)  # this export created the directory
        raise
    return os.path.join(directory, "metadata.json")


def read_frame(directory):
    """(metadata, {physics: (cells, 16) float32 array}) of a completed frame; an incomplete export
    (no metadata.json) raises FileNotFoundError."""
    with open(os.path.join(directory, "metadata.json")) as f:
        meta = json.load(f)
    dtype = np.dtype("<f4") if meta["byte_order"] == "little" else np.dtype(">f4")
    cells = int(np.prod(meta["dims"]))
    out = {}
    for entry in meta["files"]:
        arr = np.fromfile(os.path.join(directory, entry["file"]), dtype=dtype)
        if arr.size * 4 != entry["bytes"] or arr.size != COMPONENTS * cells:
            raise ValueError(f"{entry['file']} does not hold {COMPONENTS} floats per cell")
        out[entry["physics"]] = arr.astype(np.float32).reshape(cells, COMPONENTS)
    return meta, out
