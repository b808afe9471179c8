"""A small NetCDF classic-format reader (CDF-1, CDF-2, CDF-5) written from the
format specification, used by the tests as a check on the library's writer
that shares no code with it.  Returns dimensions, global attributes and
variables (attributes, type, shape, data) in file order."""
from __future__ import annotations

import struct

import numpy as np

TYPES = {1: ">i1", 2: "S1", 3: ">i2", 4: ">i4", 5: ">f4", 6: ">f8", 7: ">u1", 8: ">u2", 9: ">u4", 10: ">i8",
         11: ">u8"}


class _R:
    def __init__(self, b: bytes, v5: bool):
        self.b, self.at, self.v5 = b, 0, v5

    def take(self, n):
        if self.at + n > len(self.b):
            raise ValueError("truncated")
        s = self.b[self.at:self.at + n]
        self.at += n
        return s

    def u32(self):
        return struct.unpack(">I", self.take(4))[0]

    def u64(self):
        return struct.unpack(">Q", self.take(8))[0]

    def nonneg(self):
        return self.u64() if self.v5 else self.u32()

    def pad(self):
        self.at += (-self.at) % 4

    def name(self):
        n = self.nonneg()
        s = self.take(n).decode()
        self.pad()
        return s

    def atts(self):
        tag, n = self.u32(), self.nonneg()
        if tag == 0:
            assert n == 0
            return {}
        assert tag == 0x0C, tag
        out = {}
        for _ in range(n):
            name = self.name()
            t, m = self.u32(), self.nonneg()
            dt = np.dtype(TYPES[t])
            raw = self.take(m * dt.itemsize)
            self.pad()
            out[name] = raw.decode() if t == 2 else np.frombuffer(raw, dt).copy()
        return out


def read(path):
    b = open(path, "rb").read()
    assert b[:3] == b"CDF" and b[3] in (1, 2, 5), b[:4]
    version = b[3]
    r = _R(b, version == 5)
    r.at = 4
    numrecs = r.nonneg()
    tag, nd = r.u32(), r.nonneg()
    dims = []
    if tag:
        assert tag == 0x0A
        for _ in range(nd):
            dims.append((r.name(), r.nonneg()))
    gatts = r.atts()
    tag, nv = r.u32(), r.nonneg()
    raw_vars = []
    if tag:
        assert tag == 0x0B
        for _ in range(nv):
            name = r.name()
            nids = r.nonneg()
            ids = [r.nonneg() for _ in range(nids)]
            atts = r.atts()
            t = r.u32()
            vsize = r.nonneg()
            begin = r.u32() if version == 1 else r.u64()
            raw_vars.append((name, ids, atts, t, vsize, begin))
    is_rec = [bool(ids) and dims[ids[0]][1] == 0 for _, ids, *_ in raw_vars]
    rec_vars = [v for v, rec in zip(raw_vars, is_rec) if rec]
    recsize = sum(v[4] for v in rec_vars)
    if len(rec_vars) == 1:
        v = rec_vars[0]
        recsize = np.dtype(TYPES[v[3]]).itemsize * int(np.prod([dims[d][1] for d in v[1][1:]], dtype=np.int64))
    variables = {}
    for (name, ids, atts, t, vsize, begin), rec in zip(raw_vars, is_rec):
        dt = np.dtype(TYPES[t])
        shape = [numrecs if (rec and q == 0) else dims[d][1] for q, d in enumerate(ids)]
        if rec:
            inner = int(np.prod(shape[1:], dtype=np.int64)) * dt.itemsize
            chunks = [b[begin + r_ * recsize: begin + r_ * recsize + inner] for r_ in range(numrecs)]
            data = np.frombuffer(b"".join(chunks), dt).reshape(shape) if numrecs else np.zeros(shape, dt)
        else:
            count = int(np.prod(shape, dtype=np.int64))
            data = np.frombuffer(b[begin:begin + count * dt.itemsize], dt).reshape(shape)
        variables[name] = dict(dims=[dims[d][0] for d in ids], atts=atts, type=t, vsize=vsize, begin=begin,
                               data=data.copy(), record=rec)
    return dict(version=version, numrecs=numrecs, dims=dims, gatts=gatts, variables=variables)
