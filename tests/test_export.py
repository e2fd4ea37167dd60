"""Frame export (mhd_simulator_amd/export.py; RunState::export_fields, src/run/export.rs:9-60 and its
test accepted_fields_export_with_matching_grid_and_protect_previous_output, export.rs:72-120):
round trip of packets and metadata, the exclusive directory that protects earlier output, metadata
published last, nothing left behind by a failed export."""
import json
import os
import sys

import numpy as np
import pytest

from mhd_simulator_amd.export import export_frame, read_frame

DIMS = (5, 4, 3)
CELLS = 60


def packets(seed):
    return np.random.default_rng(seed).normal(size=16 * CELLS).astype(np.float32)


def test_round_trip(tmp_path):
    fluid, em = packets(1), packets(2)
    frame = tmp_path / "run" / "frame-00000001"
    path = export_frame(str(frame), DIMS, 1e-3, fields={"fluid": fluid, "em": em}, step=1, time_s=2.5e-6,
                        origin_m=(-6e-3, -6e-3, -6e-3), metadata={"fluid": {"relative_tolerance": 1e-5}})
    assert path == str(frame / "metadata.json")
    assert sorted(os.listdir(frame)) == ["em.bin", "fluid.bin", "metadata.json"]
    meta, got = read_frame(str(frame))
    assert meta["format_version"] == 1 and meta["step"] == 1 and meta["time_s"] == 2.5e-6
    assert meta["dims"] == list(DIMS) and meta["cell_m"] == 1e-3 and meta["origin_m"] == [-6e-3] * 3
    assert meta["byte_order"] == sys.byteorder and meta["cell_order"] == "x fastest, then y, then z"
    assert [f["physics"] for f in meta["files"]] == ["em", "fluid"]  # the reference's order
    assert all(f["bytes"] == 16 * 4 * CELLS and f["f32_components_per_cell"] == 16 for f in meta["files"])
    assert meta["fluid"] == {"relative_tolerance": 1e-5}
    assert np.array_equal(got["fluid"], fluid.reshape(CELLS, 16))
    assert np.array_equal(got["em"], em.reshape(CELLS, 16))
    # the raw file is the packet bytes, cell x fastest
    assert (frame / "fluid.bin").read_bytes() == fluid.tobytes()


def test_existing_directory_is_protected(tmp_path):
    frame = tmp_path / "frame-00000001"
    export_frame(str(frame), DIMS, 1e-3, fields={"fluid": packets(1)}, step=1, time_s=0.0)
    before = {n: (frame / n).read_bytes() for n in os.listdir(frame)}
    with pytest.raises(FileExistsError, match="Create new frame directory"):
        export_frame(str(frame), DIMS, 1e-3, fields={"fluid": packets(2)}, step=2, time_s=1.0)
    assert {n: (frame / n).read_bytes() for n in os.listdir(frame)} == before


@pytest.mark.parametrize("bad", ["size", "dtype", "physics", "empty", "key", "json"])
def test_failed_export_leaves_nothing(tmp_path, bad):
    frame = tmp_path / "frame"
    kw = dict(fields={"fluid": packets(1)}, step=1, time_s=0.0)
    if bad == "size":
        kw["fields"] = {"fluid": packets(1)[:-16]}
    elif bad == "dtype":
        kw["fields"] = {"fluid": packets(1).astype(np.float64)}
    elif bad == "physics":
        kw["fields"] = {"pressure": packets(1)}
    elif bad == "empty":
        kw["fields"] = {}
    elif bad == "key":
        kw["metadata"] = {"dims": [1, 1, 1]}
    else:
        kw["metadata"] = {"bad": object()}
    with pytest.raises((ValueError, TypeError)):
        export_frame(str(frame), DIMS, 1e-3, **kw)
    assert not frame.exists()


def test_failure_while_writing_removes_the_frame(tmp_path, monkeypatch):
    """A write that fails after the field files leaves no directory (and never a metadata.json)."""
    frame = tmp_path / "frame"
    real_rename = os.rename

    def failing_rename(src, dst):
        assert os.path.exists(src) and not os.path.exists(dst)  # metadata staged, not yet published
        raise OSError("disk full")

    monkeypatch.setattr(os, "rename", failing_rename)
    with pytest.raises(OSError, match="disk full"):
        export_frame(str(frame), DIMS, 1e-3, fields={"fluid": packets(1)}, step=1, time_s=0.0)
    monkeypatch.setattr(os, "rename", real_rename)
    assert not frame.exists()


def test_incomplete_frame_is_not_read(tmp_path):
    frame = tmp_path / "frame"
    frame.mkdir()
    (frame / "fluid.bin").write_bytes(packets(1).tobytes())
    (frame / "metadata.writing").write_text(json.dumps({"dims": list(DIMS)}))
    with pytest.raises(FileNotFoundError):
        read_frame(str(frame))
