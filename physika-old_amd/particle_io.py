"""Particle frame and state files.

ParticleWriter mirrors Dynamics/ParticleSystem/ParticleWriter.cpp:41-81: one
binary file per frame holding an int32 particle count N followed by N records
of four float32 (x, y, z, color mapping), little-endian as the reference
writes them on x86. The reference builds the file name from its output path,
a hard-coded directory and the frame index (ParticleWriter.cpp:52); here the
name is output_path + name_prefix + index + ".txt" (the prefix that
setNamePrefix stores and the reference never uses).

save_state / load_state keep a context-free dump of a simulation (positions,
velocities, scene offsets, the step count and the parameters) in a .npz
archive of plain arrays (no pickled objects), for checkpoints and resume.
"""
from __future__ import annotations

import json
import os
from typing import Optional

import numpy as np

_REC = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("c", "<f4")])


def _host(a) -> np.ndarray:
    if hasattr(a, "detach"):  # torch tensor on any device
        a = a.detach().cpu().numpy()
    return np.asarray(a)


def write_frame(path: str, position, color=None) -> None:
    """One ParticleWriter frame: int32 N, then N x (x, y, z, c) float32."""
    p = _host(position).astype(np.float32, copy=False).reshape(-1, 3)
    n = p.shape[0]
    c = np.zeros(n, np.float32) if color is None else _host(color).astype(np.float32, copy=False).reshape(-1)
    if c.shape[0] != n:
        raise ValueError("position and color mapping differ in length (ParticleWriter.cpp:46 asserts)")
    rec = np.empty(n, _REC)
    rec["x"], rec["y"], rec["z"], rec["c"] = p[:, 0], p[:, 1], p[:, 2], c
    with open(path, "wb") as f:
        f.write(np.array([n], "<i4").tobytes())
        if n:
            f.write(rec.tobytes())


def read_frame(path: str):
    """(positions (N, 3) float32, color mapping (N,) float32) of a frame file."""
    with open(path, "rb") as f:
        head = f.read(4)
        if len(head) != 4:
            raise ValueError(f"{path}: no particle count")
        n = int(np.frombuffer(head, "<i4")[0])
        if n < 0:
            raise ValueError(f"{path}: negative particle count")
        body = f.read()
    if len(body) != n * _REC.itemsize:
        raise ValueError(f"{path}: {len(body)} bytes of records, expected {n * _REC.itemsize}")
    rec = np.frombuffer(body, _REC, count=n)
    pos = np.stack([rec["x"], rec["y"], rec["z"]], 1).astype(np.float32)
    return pos, rec["c"].astype(np.float32)


class ParticleWriter:
    """IOModule writing one frame per execute() (ParticleWriter.h:12-45):
    m_position, m_color_mapping, setNamePrefix, setOutputPath."""

    def __init__(self):
        from .fluid import ArrayField
        self.m_position = ArrayField()
        self.m_color_mapping = ArrayField()
        self.m_name_prefix = ""
        self.m_output_path = ""
        self.m_output_index = 0

    def setNamePrefix(self, prefix: str) -> None:
        self.m_name_prefix = prefix

    def setOutputPath(self, path: str) -> None:
        self.m_output_path = path

    def frame_path(self, index: Optional[int] = None) -> str:
        i = self.m_output_index if index is None else index
        return f"{self.m_output_path}{self.m_name_prefix}{i}.txt"

    def execute(self) -> bool:
        pos = self.m_position.getValue()
        col = self.m_color_mapping.getValue()
        n = 0 if pos is None else pos.shape[0]
        if col is not None and col.shape[0] != n:
            raise ValueError("position and color mapping differ in length")
        write_frame(self.frame_path(), np.zeros((0, 3), np.float32) if pos is None else pos, col)
        self.m_output_index += 1
        return True


def save_state(path: str, position, velocity, params: dict, step: int = 0, scene_offsets=None) -> None:
    """Dump a simulation state for resume: arrays plus JSON parameters."""
    arrays = {"position": _host(position).astype(np.float32).reshape(-1, 3),
              "velocity": _host(velocity).astype(np.float32).reshape(-1, 3),
              "step": np.array([step], np.int64),
              "params": np.frombuffer(json.dumps(params, sort_keys=True).encode(), np.uint8)}
    if scene_offsets is not None:
        arrays["scene_offsets"] = _host(scene_offsets).astype(np.int64).reshape(-1)
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        np.savez(f, **arrays)
    os.replace(tmp, path)


def load_state(path: str) -> dict:
    """The dict save_state wrote: position, velocity, step, params,
    scene_offsets (or None)."""
    with np.load(path, allow_pickle=False) as z:
        out = {"position": z["position"], "velocity": z["velocity"], "step": int(z["step"][0]),
               "params": json.loads(bytes(z["params"]).decode()),
               "scene_offsets": z["scene_offsets"] if "scene_offsets" in z.files else None}
    return out
