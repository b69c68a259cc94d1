"""Readers of the headless frame files (pbd_write_frame / pbd_run).

RAW: 32-byte header (u32 LE: "PBDF", version, n, texture width, frame, 0, 0, 0),
then n float4 positions4 and n float4 colors4 in buffer index order (particle i
is texel (i % width, i / width) of the reference's textures).  PLY: ASCII or
binary little-endian, vertex i = particle i with x y z and uchar red green blue.
"""
from __future__ import annotations

import numpy as np

RAW_MAGIC = 0x46444250  # "PBDF"


def read_raw(path: str) -> dict:
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(32), "<u4")
        if hdr[0] != RAW_MAGIC:
            raise ValueError(f"{path}: not a PBDF frame")
        n = int(hdr[2])
        pos = np.frombuffer(f.read(n * 16), "<f4").reshape(n, 4)
        col = np.frombuffer(f.read(n * 16), "<f4").reshape(n, 4)
    return {"version": int(hdr[1]), "n": n, "texture_width": int(hdr[3]), "frame": int(hdr[4]),
            "positions4": pos, "colors4": col}


def read_ply(path: str) -> dict:
    with open(path, "rb") as f:
        data = f.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    head = data[:end].decode().splitlines()
    n = next(int(l.split()[2]) for l in head if l.startswith("element vertex"))
    binary = any(l.startswith("format binary_little_endian") for l in head)
    if binary:
        rec = np.frombuffer(data[end:end + 15 * n], np.dtype([("xyz", "<f4", 3), ("rgb", "u1", 3)]))
        xyz, rgb = rec["xyz"].copy(), rec["rgb"].copy()
    else:
        rows = np.array([l.split() for l in data[end:].decode().splitlines() if l.strip()])
        xyz = rows[:, :3].astype(np.float32) if n else np.zeros((0, 3), np.float32)
        rgb = rows[:, 3:6].astype(np.uint8) if n else np.zeros((0, 3), np.uint8)
    return {"n": n, "binary": binary, "header": head, "xyz": xyz, "rgb": rgb}
