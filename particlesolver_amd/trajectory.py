"""Binary trajectory files: the writer binds ps_trajectory_* of the native
library (the format is stated beside ps_trajectory_open in include/ps.h) and
the reader parses them with numpy alone, without the library or a GPU.

    with Writer("run.pstraj", capi.PS_FIELD_POSITIONS, dims=3) as w:
        for frame in range(100):
            sys.step(dt, 4)
            w.write(sys)
    for f in read("run.pstraj"):
        f.substeps, f.positions          # (n, 3) float32, insertion order
"""
import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import capi

MAGIC = b"PSTRAJ01"
VERSION = 1
FILE_HEADER = 32
FRAME_HEADER = 16


@dataclass
class Frame:
    substeps: int
    positions: Optional[np.ndarray] = None   # (n, 3) float32
    velocities: Optional[np.ndarray] = None  # (n, 3) float32
    phases: Optional[np.ndarray] = None      # (n,) int32

    @property
    def n(self):
        for a in (self.positions, self.velocities, self.phases):
            if a is not None:
                return a.shape[0]
        return 0


class Writer:
    """ps_trajectory_open / ps_trajectory_write / ps_trajectory_close"""

    def __init__(self, path, fields=capi.PS_FIELD_POSITIONS, dims=3):
        self.lib = capi.load()
        h = ctypes.c_void_p()
        rc = self.lib.ps_trajectory_open(str(path).encode(), fields, dims, ctypes.byref(h))
        if rc:
            raise capi.PsError(rc, "ps_trajectory_open(%s)" % path)
        self.h = h

    def write(self, system):
        rc = self.lib.ps_trajectory_write(self.h, system.h)
        if rc:
            raise capi.PsError(rc, "ps_trajectory_write")

    def close(self):
        if self.h:
            rc = self.lib.ps_trajectory_close(self.h)
            self.h = None
            if rc:
                raise capi.PsError(rc, "ps_trajectory_close")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_header(buf):
    """(fields, dims) of a trajectory held in `buf` (bytes or memoryview)"""
    if len(buf) < FILE_HEADER or bytes(buf[:8]) != MAGIC:
        raise ValueError("not a trajectory file")
    version, fields, dims = np.frombuffer(buf, "<u4", 3, 8)
    if version != VERSION:
        raise ValueError("unknown trajectory version %d" % version)
    if not 0 < fields < 8 or dims not in (2, 3):
        raise ValueError("corrupt trajectory header")
    return int(fields), int(dims)


def read(path):
    """Every frame of the file at `path`, in order (a list of Frame)."""
    buf = memoryview(open(path, "rb").read())
    fields, _ = read_header(buf)
    frames, off = [], FILE_HEADER
    while off < len(buf):
        if off + FRAME_HEADER > len(buf):
            raise ValueError("truncated frame header at byte %d" % off)
        sub = int(np.frombuffer(buf, "<u8", 1, off)[0])
        n, ff = (int(v) for v in np.frombuffer(buf, "<u4", 2, off + 8))
        if ff != fields:
            raise ValueError("frame fields differ from the file's at byte %d" % off)
        off += FRAME_HEADER
        f = Frame(sub)
        for bit, name in ((capi.PS_FIELD_POSITIONS, "positions"),
                          (capi.PS_FIELD_VELOCITIES, "velocities"),
                          (capi.PS_FIELD_PHASES, "phases")):
            if not fields & bit:
                continue
            size = 4 * n if bit == capi.PS_FIELD_PHASES else 12 * n
            if off + size > len(buf):
                raise ValueError("truncated frame at byte %d" % off)
            if bit == capi.PS_FIELD_PHASES:
                a = np.frombuffer(buf, "<i4", n, off).copy()
            else:
                a = np.frombuffer(buf, "<f4", 3 * n, off).reshape(n, 3).copy()
            setattr(f, name, a)
            off += size
        frames.append(f)
    return frames
