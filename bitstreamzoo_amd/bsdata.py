"""Reader and writer for the `.bsdata` test-data format.

Format (benchmark/bsdata.py:1-17, simulator/csim/src/bsdata_reader.h:1-12):
a UTF-8 JSON header, one NUL byte, then a binary blob.  In the header

    {"_bin": [offset, nbytes]}                 one stream, little-endian bytes
    {"_array_bin": [count, offset, stride]}    `count` streams of `stride` bytes
    {"_broadcast_pack": [count, offset, nbytes]}  array whose elements are all
                                               0 or -1, one bit each, LSB first
    JSON int / "0x..." string                  small stream value inline
    JSON list / {"idx": value}                 small array inline

A file is one case, or several under "cases" (legacy key "layers").  Each case
carries bitlength (legacy "n_vectors"), params, inputs, input_arrays,
expected, and optionally iterations / feedback (bsdata_reader.h:179-213).

Two views are offered:
  * ``read_bsdata``: Python big ints, the shape benchmark/bsdata.py:196-270
    returns (what the harness compares against);
  * ``read_cases``: numpy uint64 word arrays of ceil(bitlength/64) words, the
    shape bsdata_reader.h:262-299 decodes (what executors consume).
"""

from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "BsCase", "read_bsdata", "read_cases", "write_bsdata", "int_to_words",
    "words_to_int", "n_words_for",
]


def n_words_for(bitlength: int) -> int:
    """Words per stream for a bitlength (bsdata_reader.h:51-54): at least 1."""
    return 1 if bitlength <= 0 else (bitlength + 63) // 64


def int_to_words(value: int, n_words: int) -> np.ndarray:
    """Two's-complement little-endian words of a Python int (word_utils.h:14-64, 118-133)."""
    mask = (1 << (64 * n_words)) - 1
    v = value & mask
    return np.frombuffer(v.to_bytes(8 * n_words, "little"), dtype=np.uint64).copy()


def words_to_int(words: np.ndarray) -> int:
    return int.from_bytes(np.ascontiguousarray(words, dtype=np.uint64).tobytes(), "little")


def _bytes_to_words(buf: bytes, n_words: int) -> np.ndarray:
    """bytes_to_words (word_utils.h:91-108): copy, truncate or zero-pad to n_words."""
    out = np.zeros(n_words, dtype=np.uint64)
    nb = min(len(buf), 8 * n_words)
    if nb:
        tmp = bytearray(8 * n_words)
        tmp[:nb] = buf[:nb]
        out[:] = np.frombuffer(bytes(tmp), dtype=np.uint64)
    return out


def _json_to_words(v, n_words: int) -> np.ndarray:
    if isinstance(v, str):
        s = v.strip()
        neg = False
        if s.startswith("-"):
            neg, s = True, s[1:]
        if s[:2].lower() == "0x":
            s = s[2:]
        if s.startswith("-"):  # hex_to_words also accepts "0x-..."
            neg, s = True, s[1:]
        val = int(s, 16) if s else 0
        return int_to_words(-val if neg else val, n_words)
    if isinstance(v, bool):
        return int_to_words(int(v), n_words)
    if isinstance(v, int):
        return int_to_words(v, n_words)
    return np.zeros(n_words, dtype=np.uint64)


@dataclass
class BsCase:
    """One decoded case, streams as uint64 word arrays."""
    bitlength: int = 0
    params: dict = field(default_factory=dict)
    inputs: dict = field(default_factory=dict)          # name -> ndarray
    input_arrays: dict = field(default_factory=dict)    # name -> {idx: ndarray}
    broadcast: dict = field(default_factory=dict)       # name -> (packed bytes, count)
    expected: dict = field(default_factory=dict)        # name -> int | {idx: int}
    iterations: int = 1
    feedback: dict = field(default_factory=dict)        # out name -> in name
    name: str = ""

    @property
    def n_words(self) -> int:
        if self.bitlength > 0:
            return n_words_for(self.bitlength)
        nw = 1
        for v in self.inputs.values():
            nw = max(nw, len(v))
        return nw


def _split(path: str):
    with open(path, "rb") as f:
        raw = f.read()
    sep = raw.find(b"\0")
    if sep < 0:
        raise ValueError(f"Invalid bsdata file (no NUL separator): {path}")
    return json.loads(raw[:sep].decode("utf-8")), raw[sep + 1:]


def _case_entries(header):
    key = "cases" if "cases" in header else ("layers" if "layers" in header else None)
    return header[key] if key else [header]


def _decode_expected(v, binary: bytes):
    if isinstance(v, dict):
        if "_bin" in v:
            off, nb = v["_bin"]
            return int.from_bytes(binary[off:off + nb], "little")
        if "_broadcast_pack" in v:
            count, off, _ = v["_broadcast_pack"]
            return {i: (-1 if (off + (i >> 3) < len(binary) and (binary[off + (i >> 3)] >> (i & 7)) & 1) else 0)
                    for i in range(count)}
        if "_array_bin" in v:
            count, off, stride = v["_array_bin"]
            return {i: int.from_bytes(binary[off + i * stride: off + (i + 1) * stride], "little")
                    for i in range(count)}
        return {int(k) if str(k).lstrip("-").isdigit() else k: _decode_expected(x, binary) for k, x in v.items()}
    if isinstance(v, list):
        return {i: _decode_expected(x, binary) for i, x in enumerate(v)}
    if isinstance(v, str) and v.lower().startswith(("0x", "-0x")):
        return int(v, 16)
    return v


def read_bsdata(path: str) -> dict:
    """Python-int view (benchmark/bsdata.py:196-270)."""
    header, binary = _split(path)

    def section(s):
        return {k: _decode_expected(v, binary) for k, v in s.items()}

    def one(entry, multi):
        d = {"bitlength": entry.get("bitlength", entry.get("n_vectors", 0))}
        for key in ("params",):
            if key in entry:
                d[key] = entry[key]
        for key in ("inputs", "input_arrays", "expected"):
            if key in entry:
                d[key] = section(entry[key])
        if not multi:
            for key in ("iterations", "feedback"):
                if key in entry:
                    d[key] = entry[key]
        return d

    if "cases" in header or "layers" in header:
        return {"cases": [one(e, True) for e in _case_entries(header)]}
    return one(header, False)


def _decode_inputs_section(sec: dict, binary: bytes, nw: int, case: BsCase):
    """decode_inputs + decode_arrays + decode_broadcast_metadata (bsdata_reader.h:90-199)."""
    for k, v in sec.items():
        if isinstance(v, dict) and "_bin" in v:
            off, nb = v["_bin"]
            buf = binary[off:off + nb] if off + nb <= len(binary) else b""
            case.inputs[k] = _bytes_to_words(buf, nw)
        elif isinstance(v, dict) and "_broadcast_pack" in v:
            count, off, _ = v["_broadcast_pack"]
            nbytes = (count + 7) // 8
            packed = bytes(binary[off:off + nbytes]).ljust(nbytes, b"\0")
            case.broadcast[k] = (packed, count)
        elif isinstance(v, dict) and "_array_bin" in v:
            count, off, stride = v["_array_bin"]
            arr = {}
            for i in range(count):
                s = off + i * stride
                buf = binary[s:s + stride] if s + stride <= len(binary) else b""
                arr[i] = _bytes_to_words(buf, nw)
            case.input_arrays[k] = arr
        elif isinstance(v, list):
            case.input_arrays[k] = {i: _json_to_words(x, nw) for i, x in enumerate(v)}
        elif isinstance(v, dict):
            arr = {}
            for idx, x in v.items():
                if isinstance(x, dict) and "_bin" in x:
                    off, nb = x["_bin"]
                    buf = binary[off:off + nb] if off + nb <= len(binary) else b""
                    arr[int(idx)] = _bytes_to_words(buf, nw)
                else:
                    arr[int(idx)] = _json_to_words(x, nw)
            case.input_arrays[k] = arr
        else:
            case.inputs[k] = _json_to_words(v, nw)


def read_cases(path: str) -> list[BsCase]:
    """Word view of every case in a file (bsdata_reader.h:262-299)."""
    header, binary = _split(path)
    entries = _case_entries(header)
    multi = len(entries) > 1
    out = []
    for i, e in enumerate(entries):
        c = BsCase()
        c.bitlength = int(e.get("bitlength", e.get("n_vectors", 0)))
        nw = n_words_for(c.bitlength)
        c.params = {k: int(v) for k, v in e.get("params", {}).items()}
        if "inputs" in e:
            _decode_inputs_section(e["inputs"], binary, nw, c)
        if "input_arrays" in e:
            _decode_inputs_section(e["input_arrays"], binary, nw, c)
        if "expected" in e:
            c.expected = {k: _decode_expected(v, binary) for k, v in e["expected"].items()}
        c.iterations = int(e.get("iterations", 1))
        c.feedback = dict(e.get("feedback", {}))
        c.name = f"{path}[{i}]" if multi else path
        out.append(c)
    return out


def write_bsdata(path: str, cases: list[dict]) -> None:
    """Write cases whose streams are uint64 word arrays or Python ints.

    Each case dict: bitlength, params, inputs {name: words|int},
    input_arrays {name: {idx: words|int}}, expected {name: words|int|{idx:..}}.
    Streams go to the binary section as ceil(bitlength/8) bytes
    (benchmark/bsdata.py:58-193); ints stay inline.
    """
    parts: list[bytes] = []
    offset = 0

    def alloc(b: bytes) -> int:
        nonlocal offset
        parts.append(b)
        start = offset
        offset += len(b)
        return start

    def stream_bytes(v, nbytes: int) -> bytes:
        if isinstance(v, np.ndarray):
            raw = np.ascontiguousarray(v, dtype=np.uint64).tobytes()
            return raw[:nbytes].ljust(nbytes, b"\0")
        return (int(v) & ((1 << (8 * nbytes)) - 1)).to_bytes(nbytes, "little")

    def enc(v, bl: int):
        nbytes = max(1, math.ceil(bl / 8))
        if isinstance(v, np.ndarray):
            return {"_bin": [alloc(stream_bytes(v, nbytes)), nbytes]}
        if isinstance(v, dict):
            count = (max(v.keys()) + 1) if v else 0
            start = offset
            for i in range(count):
                alloc(stream_bytes(v.get(i, 0), nbytes))
            return {"_array_bin": [count, start, nbytes]}
        if isinstance(v, int) and (v.bit_length() > 52 or v < 0):
            return {"_bin": [alloc(stream_bytes(v, nbytes)), nbytes]}
        return v

    enc_cases = []
    for c in cases:
        bl = int(c.get("bitlength", 0))
        e = {"bitlength": bl}
        if c.get("params"):
            e["params"] = c["params"]
        for key in ("inputs", "input_arrays", "expected"):
            if key in c and c[key] is not None:
                e[key] = {k: enc(v, bl) for k, v in c[key].items()}
        if c.get("iterations", 1) > 1:
            e["iterations"] = c["iterations"]
        if c.get("feedback"):
            e["feedback"] = c["feedback"]
        enc_cases.append(e)
    header = {"format": "bsdata", "version": 1}
    if len(enc_cases) == 1:
        header.update(enc_cases[0])
    else:
        header["cases"] = enc_cases
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(json.dumps(header, separators=(",", ":")).encode("utf-8"))
        f.write(b"\0")
        for p in parts:
            f.write(p)
    os.replace(tmp, path)
