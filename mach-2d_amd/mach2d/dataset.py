"""Versioned on-disk format of the learned guess's training set: (p' system,
converged p') pairs that the MSI solver generates during SIMPLEC runs.

No reference counterpart (the reference keeps no training data); the
format sits beside the U-Net weights format (M2DW, mach2d.unet).

File layout, little-endian, version 1:

    header (64 bytes)
        magic    4s   b"M2DT"
        version  u32  1
        nx, ny   u32  grid (fictitious volumes included)
        ndiag    u32  5 (slots S, W, P, E, N) or 9 (SW .. NE)
        count    u32  records in the file
        dtype    u32  1 = float32 (the only one written)
        flags    u32  0
        gamma    f64  residual drop the targets were solved to
        source   u32  0 = unknown, 1 = SIMPLEC p' systems (internal flow)
        (zero padding to 64 bytes)
    records, count of them, each ndiag*N + 2*N float32 values:
        a (ndiag, N)   the coefficients, diagonal-major (the reference's
                       a(nij, ndiag) seen from C), N = nx*ny
        b (N)          the right-hand side
        x (N)          the converged solution of A x = b
    metadata, count rows of 4 float64:
        seed, step, itp, ||A x - b|| / ||b|| of the stored target

Records have a fixed size, so `Dataset` maps the file (numpy.memmap) and a
training loop reads batches without loading the whole set.
"""
from __future__ import annotations

import os
import struct

import numpy as np

MAGIC = b"M2DT"
VERSION = 1
HEADER = 64
_HFMT = "<4sIIIIIIIdI"
SOURCE_UNKNOWN, SOURCE_SIMPLEC = 0, 1


def record_floats(nx, ny, ndiag):
    return (ndiag + 2) * nx * ny


class Writer:
    """Append records to a new M2DT file; the count is written on close."""

    def __init__(self, path, nx, ny, ndiag, gamma, source=SOURCE_SIMPLEC):
        if ndiag not in (5, 9):
            raise ValueError("ndiag must be 5 or 9")
        self.path, self.nx, self.ny, self.ndiag = path, nx, ny, ndiag
        self.gamma, self.source = float(gamma), int(source)
        self.count = 0
        self.meta = []
        self.f = open(path, "wb")
        self.f.write(self._header())

    def _header(self):
        h = struct.pack(_HFMT, MAGIC, VERSION, self.nx, self.ny, self.ndiag, self.count, 1, 0,
                        self.gamma, self.source)
        return h + b"\0" * (HEADER - len(h))

    def add(self, a, b, x, seed=0, step=0, itp=0, rel_residual=float("nan")):
        N = self.nx * self.ny
        a = np.asarray(a, dtype=np.float32).reshape(self.ndiag, N)
        b = np.asarray(b, dtype=np.float32).reshape(N)
        x = np.asarray(x, dtype=np.float32).reshape(N)
        self.f.write(np.ascontiguousarray(a).tobytes())
        self.f.write(np.ascontiguousarray(b).tobytes())
        self.f.write(np.ascontiguousarray(x).tobytes())
        self.meta.append((float(seed), float(step), float(itp), float(rel_residual)))
        self.count += 1

    def close(self):
        if self.f is None:
            return
        self.f.write(np.asarray(self.meta, dtype="<f8").reshape(-1, 4).tobytes())
        self.f.seek(0)
        self.f.write(self._header())
        self.f.close()
        self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Dataset:
    """Read-only view of an M2DT file: len(ds), ds[i] -> (a, b, x) float32
    arrays (a: (ndiag, N)), ds.meta (count, 4)."""

    def __init__(self, path):
        with open(path, "rb") as f:
            h = f.read(HEADER)
        if len(h) < HEADER or h[:4] != MAGIC:
            raise ValueError("%s: not an M2DT file" % path)
        (_, version, nx, ny, ndiag, count, dtype, _flags, gamma,
         source) = struct.unpack(_HFMT, h[:struct.calcsize(_HFMT)])
        if version != VERSION:
            raise ValueError("%s: M2DT version %d (this reader: %d)" % (path, version, VERSION))
        if dtype != 1 or ndiag not in (5, 9):
            raise ValueError("%s: unsupported dtype %d / ndiag %d" % (path, dtype, ndiag))
        self.nx, self.ny, self.ndiag, self.count = nx, ny, ndiag, count
        self.gamma, self.source = gamma, source
        self.N = nx * ny
        rf = record_floats(nx, ny, ndiag)
        need = HEADER + 4 * rf * count + 32 * count
        if os.path.getsize(path) != need:
            raise ValueError("%s: size %d, expected %d" % (path, os.path.getsize(path), need))
        self.records = np.memmap(path, dtype="<f4", mode="r", offset=HEADER,
                                 shape=(count, rf)) if count else np.zeros((0, rf), np.float32)
        self.meta = np.fromfile(path, dtype="<f8", count=4 * count,
                                offset=HEADER + 4 * rf * count).reshape(count, 4)

    def __len__(self):
        return self.count

    def __getitem__(self, i):
        r = self.records[i]
        k = self.ndiag * self.N
        return (r[:k].reshape(self.ndiag, self.N), r[k:k + self.N], r[k + self.N:])
