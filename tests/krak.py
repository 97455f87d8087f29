"""KRAK v1 proof bytes (src/bin/stark_cli.rs:907-1089): parse and re-serialise.
Test helper for tamper cases and field-by-field comparisons."""

from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import List

MAGIC = 0x4B52414B


@dataclass
class Decom:
    width: int
    values: list = field(default_factory=list)
    sibling_values: list = field(default_factory=list)
    paths: list = field(default_factory=list)
    sibling_paths: list = field(default_factory=list)


@dataclass
class Proof:
    log_n: int
    a: int
    b: int
    trace_root: List[int]
    quotient_root: List[int]
    fri_roots: List[List[int]]
    last_layer: List[List[int]]
    queries: List[int]
    trace: Decom
    quotient: Decom
    fri: List[Decom]


class _R:
    def __init__(self, data: bytes):
        self.w = struct.unpack(f"<{len(data) // 4}I", data[: len(data) // 4 * 4])
        self.i = 0
        if len(data) % 4:
            raise ValueError("length not a multiple of 4")

    def u(self) -> int:
        v = self.w[self.i]
        self.i += 1
        return v

    def n(self, k: int) -> List[int]:
        v = list(self.w[self.i: self.i + k])
        if len(v) != k:
            raise ValueError("truncated")
        self.i += k
        return v


def _read_decom(r: _R, width: int) -> Decom:
    n = r.u()
    d = Decom(width)
    if width == 1:
        d.values = r.n(n)
        d.sibling_values = r.n(n)
    else:
        d.values = [r.n(4) for _ in range(n)]
        d.sibling_values = [r.n(4) for _ in range(n)]
    for lst in (d.paths, d.sibling_paths):
        for _ in range(n):
            depth = r.u()
            lst.append([r.n(8) for _ in range(depth)])
    return d


def parse(data: bytes) -> Proof:
    r = _R(data)
    if r.u() != MAGIC or r.u() != 1:
        raise ValueError("bad header")
    log_n, a, b = r.u(), r.u(), r.u()
    tr, qr = r.n(8), r.n(8)
    fri_roots = [r.n(8) for _ in range(r.u())]
    last = [r.n(4) for _ in range(r.u())]
    queries = r.n(r.u())
    trace = _read_decom(r, 1)
    quot = _read_decom(r, 4)
    fri = [_read_decom(r, 4) for _ in range(r.u())]
    if r.i != len(r.w):
        raise ValueError("trailing words")
    return Proof(log_n, a, b, tr, qr, fri_roots, last, queries, trace, quot, fri)


def _write_decom(out: list, d: Decom) -> None:
    out.append(len(d.values))
    if d.width == 1:
        out += d.values + d.sibling_values
    else:
        for v in d.values + d.sibling_values:
            out += v
    for lst in (d.paths, d.sibling_paths):
        for p in lst:
            out.append(len(p))
            for h in p:
                out += h


def serialize(p: Proof) -> bytes:
    out = [MAGIC, 1, p.log_n, p.a, p.b] + p.trace_root + p.quotient_root
    out.append(len(p.fri_roots))
    for rt in p.fri_roots:
        out += rt
    out.append(len(p.last_layer))
    for v in p.last_layer:
        out += v
    out.append(len(p.queries))
    out += p.queries
    _write_decom(out, p.trace)
    _write_decom(out, p.quotient)
    out.append(len(p.fri))
    for d in p.fri:
        _write_decom(out, d)
    return struct.pack(f"<{len(out)}I", *out)
