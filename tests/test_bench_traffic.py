"""bench.py reports HBM traffic from a committed counter summary only when the summary was taken on
the kernel sources in the tree: a summary of other sources gives traffic null with the reason."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def summary(tmp_path, digest, cells=1000):
    f = tmp_path / "counters.json"
    f.write_text(json.dumps({"cells": cells, "copy_scale_fetch": 2.0, "copy_scale_write": 1.0,
                             "source_digest": digest,
                             "kernels": {"apply": {"fetch_kib": 10, "write_kib": 3}}}))
    return str(f)


def test_digest_is_stable_and_hex():
    d = bench.source_digest()
    assert d == bench.source_digest() and len(d) == 64
    int(d, 16)


def test_matching_summary_gives_scaled_bytes(tmp_path):
    t, note = bench.profiled_traffic(summary(tmp_path, bench.source_digest()), "cells", 1000, "apply")
    assert note is None and t == 1024.0 * (10 * 2.0 + 3 * 1.0)


def test_stale_or_foreign_summary_gives_null(tmp_path):
    f = summary(tmp_path, "0" * 64)
    t, note = bench.profiled_traffic(f, "cells", 1000, "apply")
    assert t is None and "stale" in note
    f = summary(tmp_path, bench.source_digest())
    assert bench.profiled_traffic(f, "cells", 999, "apply")[0] is None
    assert bench.profiled_traffic(f, "cells", 1000, "smooth")[0] is None
    assert bench.profiled_traffic(str(tmp_path / "missing.json"), "cells", 1000, "apply")[0] is None
