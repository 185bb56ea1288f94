"""Checkpoint files (prandtl_checkpoint_write/_read): round trip, and the
load errors for missing, truncated, corrupted, mismatched and foreign files.
Host arrays only: these run without a GPU."""
import json
import os

import numpy as np
import pytest

from prandtl_amd import PrandtlError, _lib
from prandtl_amd.operator import checkpoint_read, checkpoint_write

MESH = dict(dim=2, order=3, nelem=(5, 4), lo=(-1.0, -1.0), hi=(1.0, 1.0))


def state(seed=3):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(4 * 20 * 16)


def paths(d, cycle):
    base = os.path.join(str(d), "Cycle%d" % cycle, "checkpoint_cycle_%d" % cycle)
    return base + ".00000000.chk", base + ".json"


def expect_io(d, cycle=7, mesh=MESH):
    with pytest.raises(PrandtlError) as e:
        checkpoint_read(d, cycle, **mesh)
    assert e.value.code == _lib.PRANDTL_ERR_IO
    return str(e.value)


def test_round_trip_bitwise(tmp_path):
    u = state()
    t = 0.1 + 1e-16 * 3          # needs all 17 digits
    checkpoint_write(tmp_path, 7, t, u, **MESH)
    v, t2 = checkpoint_read(tmp_path, 7, **MESH)
    assert np.array_equal(u, v) and t2 == t
    chk, meta = paths(tmp_path, 7)
    j = json.load(open(meta))
    assert j == {"time": t, "cycle": 7}
    assert os.path.getsize(chk) == 56 + 8 * u.size + 8


def test_missing_files(tmp_path):
    assert "cannot open" in expect_io(tmp_path)
    checkpoint_write(tmp_path, 7, 0.5, state(), **MESH)
    chk, meta = paths(tmp_path, 7)
    os.remove(chk)
    assert "cannot open" in expect_io(tmp_path)


@pytest.mark.parametrize("cut", [10, 48 + 100, -4])
def test_truncated(tmp_path, cut):
    checkpoint_write(tmp_path, 7, 0.5, state(), **MESH)
    chk, _ = paths(tmp_path, 7)
    data = open(chk, "rb").read()
    open(chk, "wb").write(data[:cut])
    assert "truncated" in expect_io(tmp_path)


def test_corrupted_value_fails_checksum(tmp_path):
    checkpoint_write(tmp_path, 7, 0.5, state(), **MESH)
    chk, _ = paths(tmp_path, 7)
    data = bytearray(open(chk, "rb").read())
    data[48 + 8 * 123 + 3] ^= 0x10
    open(chk, "wb").write(bytes(data))
    assert "checksum" in expect_io(tmp_path)


def test_trailing_data(tmp_path):
    checkpoint_write(tmp_path, 7, 0.5, state(), **MESH)
    chk, _ = paths(tmp_path, 7)
    open(chk, "ab").write(b"\0")
    assert "trailing" in expect_io(tmp_path)


def test_other_mesh_or_order(tmp_path):
    checkpoint_write(tmp_path, 7, 0.5, state(), **MESH)
    assert "differ" in expect_io(tmp_path, mesh=dict(MESH, nelem=(4, 5)))
    assert "differ" in expect_io(tmp_path, mesh=dict(MESH, order=2))


def test_meta_cycle_and_magic(tmp_path):
    checkpoint_write(tmp_path, 7, 0.5, state(), **MESH)
    chk, meta = paths(tmp_path, 7)
    open(meta, "w").write('{"time": 0.5, "cycle": 8}')
    assert "cycle" in expect_io(tmp_path)
    open(meta, "w").write('{"cycle": 7}')
    assert "no time" in expect_io(tmp_path)
    open(meta, "w").write('{"time": 0.5, "cycle": 7}')
    data = bytearray(open(chk, "rb").read())
    data[0:8] = b"NOTACHK!"
    open(chk, "wb").write(bytes(data))
    assert "not a checkpoint" in expect_io(tmp_path)


def test_same_counts_other_geometry_refused(tmp_path):
    """A checkpoint keeps a fingerprint of the geometry it was written on: a
    box with the same element counts, order and variables but other extents
    does not load it."""
    checkpoint_write(tmp_path, 7, 0.5, state(), **MESH)
    other = dict(MESH, hi=(1.0, 2.0))
    assert "another mesh" in expect_io(tmp_path, mesh=other)
    v, _ = checkpoint_read(tmp_path, 7, **MESH)
    assert np.array_equal(v, state())
