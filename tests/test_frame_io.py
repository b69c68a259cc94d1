"""Headless frame output (pbd_write_frame_arrays, host side): exact round trips
of the RAW and PLY files and the renderer's shade (particleSpriteGeom.vert:40:
mix((0.95, 0.97, 0.97), (0, 0, 1), colors4.x^(1/6)))."""
import numpy as np
import pytest

from pbdfluid import capi, frames


def _arrays(n=37, seed=0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-3, 13, (n, 4)).astype(np.float32)
    col = rng.uniform(0, 1, (n, 4)).astype(np.float32)
    return pos, col


def test_raw_round_trip(tmp_path):
    pos, col = _arrays()
    path = str(tmp_path / "f.pbdf")
    capi.write_frame_arrays(path, pos, col, frame=7, texture_width=256, fmt="raw")
    r = frames.read_raw(path)
    assert (r["version"], r["n"], r["texture_width"], r["frame"]) == (1, 37, 256, 7)
    assert np.array_equal(r["positions4"], pos) and np.array_equal(r["colors4"], col)


@pytest.mark.parametrize("fmt", ["ply", "plyb"])
def test_ply_round_trip_and_shade(tmp_path, fmt):
    pos, col = _arrays()
    col[0, 0], col[1, 0], col[2, 0] = 0.0, 1.0, 1.0 / 64.0
    path = str(tmp_path / "f.ply")
    capi.write_frame_arrays(path, pos, col, frame=3, texture_width=16, fmt=fmt)
    r = frames.read_ply(path)
    assert r["n"] == 37 and r["binary"] == (fmt == "plyb")
    assert np.array_equal(r["xyz"], pos[:, :3])  # %.9g round-trips float32
    assert tuple(r["rgb"][0]) == (242, 247, 247)  # (0.95, 0.97, 0.97) * 255
    assert tuple(r["rgb"][1]) == (0, 0, 255)
    assert tuple(r["rgb"][2]) == (121, 124, 251)  # t = (1/64)^(1/6) = 1/2
    t = np.clip(col[:, 0], 0, None) ** (1 / 6)
    ref = np.array([0.95, 0.97, 0.97]) * (1 - t[:, None]) + np.array([0.0, 0.0, 1.0]) * t[:, None]
    assert np.abs(r["rgb"].astype(int) - np.rint(ref * 255)).max() <= 1


def test_empty_frame_and_default_colour(tmp_path):
    path = str(tmp_path / "e.pbdf")
    capi.write_frame_arrays(path, np.zeros((0, 4), np.float32), fmt="raw")
    assert frames.read_raw(path)["n"] == 0
    pos, _ = _arrays(5)
    capi.write_frame_arrays(path, pos, None, fmt="raw")
    assert np.all(frames.read_raw(path)["colors4"] == 1.0)


def test_errors(tmp_path):
    pos, col = _arrays(3)
    rc = capi.lib().pbd_write_frame_arrays(str(tmp_path / "f").encode(), 9, 0, 3, 0, pos.ctypes.data, None)
    assert rc == 1 and b"unknown format" in capi.lib().pbd_last_error(None)  # PBD_ERR_INVALID
    with pytest.raises(capi.PbdError) as e:
        capi.write_frame_arrays(str(tmp_path / "no" / "such" / "dir" / "f.ply"), pos, col, fmt="ply")
    assert e.value.code == 8  # PBD_ERR_IO
