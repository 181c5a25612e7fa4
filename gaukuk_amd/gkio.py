"""`cons_XXXXX` / `prim_XXXXX` output files.

Byte layout written by Sim::WriteData (src/utils/write_sim.cpp:14-58), native endianness:

    int32  sizeof(Real)                 (8 or 4)
    Real   t
    int32  NVar, nx, ny, nz, nGhost
    Real   xc[lenx], yc[leny], zc[lenz]  cell centres incl. ghosts
    int32  size = NVar*lenz*leny*lenx
    Real   data[NVar][lenz][leny][lenx]  incl. all ghost cells

with len* = n* + 2*nGhost (so a 2D run still carries 2*nGhost empty z planes), as parsed by
the reference's ReadGaukuk (pypkg/read_gaukuk.py:3-44).  Cell centres follow
src/sim.cpp:39-41: ``xc(i) = xmin - nGhost*dx + 0.5*dx + i*dx``.
"""
from __future__ import annotations

import numpy as np

NVAR = 5


def cell_centres(lo: float, hi: float, n: int, ng: int, dtype=np.float64) -> np.ndarray:
    """Ghost-inclusive cell centres, evaluated in the reference's operation order."""
    d = dtype((dtype(hi) - dtype(lo)) / dtype(n))
    i = np.arange(n + 2 * ng, dtype=dtype)
    base = dtype(dtype(lo) - dtype(ng) * d) + dtype(0.5) * d
    return (base + i * d).astype(dtype)


def write_output(path: str, t: float, data: np.ndarray, nx: int, ny: int, nz: int, ng: int,
                 extent) -> None:
    """Write ``data`` (NVar, lenz, leny, lenx) in the reference layout."""
    dtype = data.dtype.type
    xmin, xmax, ymin, ymax, zmin, zmax = extent
    lenx, leny, lenz = nx + 2 * ng, ny + 2 * ng, nz + 2 * ng
    if data.shape != (NVAR, lenz, leny, lenx):
        raise ValueError(f"data shape {data.shape} != {(NVAR, lenz, leny, lenx)}")
    with open(path, "wb") as f:
        np.array([np.dtype(dtype).itemsize], np.int32).tofile(f)
        np.array([t], dtype).tofile(f)
        np.array([NVAR, nx, ny, nz, ng], np.int32).tofile(f)
        cell_centres(xmin, xmax, nx, ng, dtype).tofile(f)
        cell_centres(ymin, ymax, ny, ng, dtype).tofile(f)
        cell_centres(zmin, zmax, nz, ng, dtype).tofile(f)
        np.array([data.size], np.int32).tofile(f)
        np.ascontiguousarray(data).tofile(f)


class GaukukOutput:
    """Reader with the same attributes as the reference's ReadGaukuk."""

    def __init__(self, filename: str):
        with open(filename, "rb") as f:
            size = int(np.fromfile(f, np.int32, 1)[0])
            dtype = {8: np.float64, 4: np.float32}[size]
            self.t = np.fromfile(f, dtype, 1)[0]
            self.nvar, self.nx, self.ny, self.nz, self.nGhost = (int(v) for v in np.fromfile(f, np.int32, 5))
            self.lenx = self.nx + 2 * self.nGhost
            self.leny = self.ny + 2 * self.nGhost
            self.lenz = self.nz + 2 * self.nGhost
            self.xcg = np.fromfile(f, dtype, self.lenx)
            self.ycg = np.fromfile(f, dtype, self.leny)
            self.zcg = np.fromfile(f, dtype, self.lenz)
            g = self.nGhost
            self.xc = self.xcg[g:g + self.nx]
            self.yc = self.ycg[g:g + self.ny]
            self.zc = self.zcg[g:g + self.nz]
            self.dataLen = int(np.fromfile(f, np.int32, 1)[0])
            if self.dataLen != self.nvar * self.lenx * self.leny * self.lenz:
                raise ValueError("data size does not match the header")
            self.data = np.fromfile(f, dtype, self.dataLen).reshape(self.nvar, self.lenz, self.leny, self.lenx)

    def active(self) -> np.ndarray:
        """Active cells: (NVar, ny, nx) in 2D, (NVar, nz, ny, nx) when nz > 1."""
        g = self.nGhost
        if self.nz > 1:
            return self.data[:, g:g + self.nz, g:g + self.ny, g:g + self.nx]
        return self.data[:, g, g:g + self.ny, g:g + self.nx]


def embed_active(active: np.ndarray, ng: int, nz_planes: int = 1) -> np.ndarray:
    """(NVar, ny, nx) active cells -> (NVar, lenz, leny, lenx) with zero ghosts (nz = 1)."""
    nv, ny, nx = active.shape
    lenz = nz_planes + 2 * ng
    out = np.zeros((nv, lenz, ny + 2 * ng, nx + 2 * ng), active.dtype)
    out[:, ng, ng:ng + ny, ng:ng + nx] = active
    return out
