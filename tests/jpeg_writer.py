"""Minimal baseline JPEG writer for the decoder tests (test infrastructure).

Writes quantised coefficients that are already known (raster block order, natural order
inside a block) as an ITU-T T.81 baseline file: one interleaved scan over every
component, blocks of an MCU in the standard order (row by row inside the component's
H x V group), optional restart interval.  The Huffman tables are fixed-length codes (4
bits for the 12 DC categories, 8 bits for the 162 AC run/size symbols), which are valid
T.81 tables and keep the writer short; variable-length tables are covered by files that
Pillow's libjpeg encoder writes.
"""
from __future__ import annotations

import struct

import numpy as np

ZIGZAG = [0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5, 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,
          7, 14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31,
          39, 46, 53, 60, 61, 54, 47, 55, 62, 63]

DC_SYMBOLS = list(range(12))
AC_SYMBOLS = [0x00, 0xF0] + [(r << 4) | s for r in range(16) for s in range(1, 11)]


class _Bits:
    def __init__(self):
        self.out = bytearray()
        self.acc = 0
        self.n = 0

    def put(self, value: int, length: int) -> None:
        for i in range(length - 1, -1, -1):
            self.acc = (self.acc << 1) | ((value >> i) & 1)
            self.n += 1
            if self.n == 8:
                self.out.append(self.acc)
                if self.acc == 0xFF:
                    self.out.append(0x00)
                self.acc = 0
                self.n = 0

    def flush(self) -> None:
        if self.n:
            self.put((1 << (8 - self.n)) - 1, 8 - self.n)


def _category(v: int) -> int:
    return 0 if v == 0 else int(abs(v)).bit_length()


def _bits_of(v: int, s: int) -> int:
    return v if v >= 0 else v + (1 << s) - 1


def _segment(marker: int, payload: bytes) -> bytes:
    return struct.pack(">HH", 0xFF00 | marker, len(payload) + 2) + payload


def write_baseline(width: int, height: int, components, restart_interval: int = 0, rst_log=None) -> bytes:
    """components: list of (h, v, quant[64] natural, coefs[block_h*block_w*64] natural, block_w).

    The component block grids must be the MCU-rounded planes (as the decoder produces them).
    rst_log, a list, receives for every RSTn marker written (offset of its 0xFF byte in the
    file, number of padding bits before it).
    """
    hmax = max(c[0] for c in components)
    vmax = max(c[1] for c in components)
    out = bytearray(b"\xff\xd8")
    for t, c in enumerate(components):
        q = np.asarray(c[2]).reshape(-1)
        out += _segment(0xDB, bytes([t]) + bytes(int(q[ZIGZAG[k]]) for k in range(64)))
    sof = struct.pack(">BHHB", 8, height, width, len(components))
    for t, c in enumerate(components):
        sof += bytes([t + 1, (c[0] << 4) | c[1], t])
    out += _segment(0xC0, sof)
    dc_counts = [0] * 16
    dc_counts[3] = len(DC_SYMBOLS)
    ac_counts = [0] * 16
    ac_counts[7] = len(AC_SYMBOLS)
    out += _segment(0xC4, bytes([0x00]) + bytes(dc_counts) + bytes(DC_SYMBOLS))
    out += _segment(0xC4, bytes([0x10]) + bytes(ac_counts) + bytes(AC_SYMBOLS))
    if restart_interval:
        out += _segment(0xDD, struct.pack(">H", restart_interval))
    sos = bytes([len(components)])
    for t in range(len(components)):
        sos += bytes([t + 1, 0x00])
    out += _segment(0xDA, sos + bytes([0, 63, 0]))
    dc_code = {s: i for i, s in enumerate(DC_SYMBOLS)}
    ac_code = {s: i for i, s in enumerate(AC_SYMBOLS)}
    interleaved = hmax > 1 or vmax > 1
    mw = -(-width // (8 * hmax)) if interleaved else -(-width // 8)
    mh = -(-height // (8 * vmax)) if interleaved else -(-height // 8)
    bits = _Bits()
    pred = [0] * len(components)
    todo = restart_interval
    rst = 0
    for r in range(mh):
        for col in range(mw):
            for ci, (h, v, _, coefs, bw) in enumerate(components):
                blocks = np.asarray(coefs).reshape(-1, 64)
                for vs in range(v):
                    for hs in range(h):
                        blk = blocks[(r * v + vs) * bw + col * h + hs]
                        diff = int(blk[0]) - pred[ci]
                        pred[ci] = int(blk[0])
                        s = _category(diff)
                        bits.put(dc_code[s], 4)
                        if s:
                            bits.put(_bits_of(diff, s), s)
                        run = 0
                        for k in range(1, 64):
                            a = int(blk[ZIGZAG[k]])
                            if a == 0:
                                run += 1
                                continue
                            while run > 15:
                                bits.put(ac_code[0xF0], 8)
                                run -= 16
                            s = _category(a)
                            bits.put(ac_code[(run << 4) | s], 8)
                            bits.put(_bits_of(a, s), s)
                            run = 0
                        if run:
                            bits.put(ac_code[0x00], 8)
            if restart_interval:
                todo -= 1
                last = r == mh - 1 and col == mw - 1
                if todo == 0 and not last:
                    pad = (8 - bits.n) % 8
                    bits.flush()
                    if rst_log is not None:
                        rst_log.append((len(out) + len(bits.out), pad))
                    bits.out += bytes([0xFF, 0xD0 + rst])
                    rst = (rst + 1) % 8
                    pred = [0] * len(components)
                    todo = restart_interval
    bits.flush()
    out += bits.out + b"\xff\xd9"
    return bytes(out)


class ScanBits(_Bits):
    """Entropy-coded data written symbol by symbol with the writer's fixed-length tables:
    4-bit DC category codes and 8-bit AC run/size codes."""

    _dc = {s: i for i, s in enumerate(DC_SYMBOLS)}
    _ac = {s: i for i, s in enumerate(AC_SYMBOLS)}

    def dc(self, diff: int) -> "ScanBits":
        s = _category(diff)
        self.put(self._dc[s], 4)
        if s:
            self.put(_bits_of(diff, s), s)
        return self

    def ac(self, run: int, value: int) -> "ScanBits":
        """Run/size symbol for `value` after `run` zeros; value 0 with run 0 is EOB, run 15 ZRL."""
        s = _category(value)
        self.put(self._ac[(run << 4) | s], 8)
        if s:
            self.put(_bits_of(value, s), s)
        return self

    def raw(self, value: int, length: int) -> "ScanBits":
        self.put(value, length)
        return self

    def data(self) -> bytes:
        self.flush()
        return bytes(self.out)


def write_grey_scans(width: int, height: int, quant, scans, progressive: bool = True) -> bytes:
    """A one-component file with hand-written scans.  scans: list of (ss, se, ah, al, ScanBits)."""
    q = np.asarray(quant).reshape(-1)
    out = bytearray(b"\xff\xd8")
    out += _segment(0xDB, bytes([0]) + bytes(int(q[ZIGZAG[k]]) for k in range(64)))
    out += _segment(0xC2 if progressive else 0xC0, struct.pack(">BHHB", 8, height, width, 1) + bytes([1, 0x11, 0]))
    dc_counts = [0] * 16
    dc_counts[3] = len(DC_SYMBOLS)
    ac_counts = [0] * 16
    ac_counts[7] = len(AC_SYMBOLS)
    out += _segment(0xC4, bytes([0x00]) + bytes(dc_counts) + bytes(DC_SYMBOLS))
    out += _segment(0xC4, bytes([0x10]) + bytes(ac_counts) + bytes(AC_SYMBOLS))
    for ss, se, ah, al, bits in scans:
        out += _segment(0xDA, bytes([1, 1, 0x00, ss, se, (ah << 4) | al]))
        out += bits.data()
    out += b"\xff\xd9"
    return bytes(out)
