"""Benchmark-harness driver for `bsim_hip`.

`HipBackend` has the interface of the reference harness's backends
(`Backend.run` / `run_batch`, benchmark/base.py:23-30 and the CudaBackend at
benchmark/base.py:181-254): it runs the `bsim_hip` executable once per case
(JSON on stdin) or per `.bsdata` file, and turns its JSON back into Python
ints.  Registering it as backend "hip" in the reference's
benchmark/run_all.py:make_backend is a two-line change: construct
`HipBackend` where that function maps the name "hip".

Everything here is process plumbing; the program runs in libbsz_hip.so on
the GPU.
"""

from __future__ import annotations

import json
import os
import subprocess
from dataclasses import dataclass

from ._lib import CLI_PATH

def subprocess_env() -> dict:
    """The environment bsim_hip runs in: this process's, plus one glibc
    tuning knob unless the caller already set it.  Large host buffers are
    cleared by memset, which on some AMD Zen 4 hosts is slow once glibc
    switches it to `rep stosb`; raising that switch point to 2e9 bytes keeps
    every stream buffer on the vector path.  The reference harness applies
    the same knob to its C++ backends (benchmark/base.py:84-95)."""
    knob, value = "glibc.cpu.x86_rep_stosb_threshold", 2 * 10**9
    env = dict(os.environ)
    have = env.get("GLIBC_TUNABLES", "")
    if knob.rsplit(".", 1)[-1] not in have:
        env["GLIBC_TUNABLES"] = ":".join(x for x in (have, f"{knob}={value}") if x)
    return env


@dataclass
class ProgramRef:
    """The one attribute of the harness's ProgramInfo a backend reads."""
    source_path: str


def _json_value(v):
    """An input value as bsim reads it: ints that JSON cannot carry exactly
    (negative, or 2^53 and above) as hex strings."""
    if isinstance(v, bool):
        return int(v)
    if isinstance(v, int) and (v < 0 or v >= 1 << 53):
        return hex(v)
    return v


def build_input_json(inputs: dict, params: dict | None, input_arrays: dict | None, bitlength: int,
                     iterations: int = 1, feedback: dict | None = None) -> dict:
    """The stdin case bsim reads (keys of bsdata_reader.h:303-335)."""
    d = {"inputs": {k: _json_value(v) for k, v in (inputs or {}).items()},
         "params": dict(params or {}),
         "n_vectors": bitlength}
    if input_arrays:
        arrs = {}
        for name, arr in input_arrays.items():
            if isinstance(arr, dict):
                n = max(arr) + 1 if arr else 0
                arrs[name] = [_json_value(arr.get(i, 0)) for i in range(n)]
            else:
                arrs[name] = [_json_value(v) for v in arr]
        d["input_arrays"] = arrs
    if iterations > 1:
        d["iterations"] = iterations
    if feedback:
        d["feedback"] = dict(feedback)
    return d


def parse_outputs(outputs: dict) -> dict:
    """bsim's "outputs" object to Python ints: "0x…" strings are streams,
    objects are arrays keyed by index, numbers are int outputs."""
    res = {}
    for k, v in outputs.items():
        if isinstance(v, str) and v.startswith("0x"):
            res[k] = int(v, 16)
        elif isinstance(v, dict):
            res[k] = {int(i): (int(x, 16) if isinstance(x, str) and x.startswith("0x") else x) for i, x in v.items()}
        else:
            res[k] = v
    return res


def read_output_bsdata(path: str) -> list[dict]:
    """The per-case JSON objects of a `bsim_hip --output-bsdata` file with
    every {"_bin": [offset, bytes]} output replaced by its words as a Python
    int (the hex strings of the default output, without building them)."""
    with open(path, "rb") as f:
        raw = f.read()
    sep = raw.index(b"\0")
    head = json.loads(raw[:sep].decode())
    blob = memoryview(raw)[sep + 1:]

    def val(v):
        if isinstance(v, dict) and "_bin" in v:
            off, nb = v["_bin"]
            return int.from_bytes(blob[off:off + nb], "little")
        return v

    cases = head if isinstance(head, list) else [head]
    for c in cases:
        outs = c.get("outputs", {})
        for k, v in list(outs.items()):
            if isinstance(v, dict) and "_bin" not in v:
                outs[k] = {int(i): val(x) for i, x in v.items()}
            else:
                outs[k] = val(v)
    return cases


class HipBackend:
    """Runs programs through bsim_hip (the MI355X executor).  With
    binary_outputs, output streams come back through a binary .bsdata file
    (--output-bsdata) instead of hex strings in the JSON.  padding="bsim"
    gives bits above the bitlength bsim_cuda's values for word_bits
    (--padding=bsim); the default keeps them clear."""

    def __init__(self, bsim_path: str | None = None, word_bits: int = 64, reuse_mem: bool = True,
                 executor: str = "fused", timeout: int = 7200, binary_outputs: bool = False,
                 padding: str = "clean"):
        self.bsim_path = bsim_path or CLI_PATH
        self.word_bits = word_bits
        self.padding = padding
        self.reuse_mem = reuse_mem
        self.executor = executor
        self.timeout = timeout
        self.binary_outputs = binary_outputs
        if not os.path.exists(self.bsim_path):
            raise FileNotFoundError(f"bsim_hip not found at {self.bsim_path}; build it with "
                                    "`python -m bitstreamzoo_amd.build`")

    def _cmd(self, prog) -> list[str]:
        cmd = [self.bsim_path, prog.source_path, f"--word-bits={self.word_bits}", f"--executor={self.executor}"]
        if self.reuse_mem:
            cmd.append("--reuse-mem")
        if self.padding != "clean":
            cmd.append(f"--padding={self.padding}")
        return cmd

    def _check(self, proc: subprocess.CompletedProcess):
        if proc.returncode != 0:
            err = proc.stderr.strip()
            if "out of memory" in err:
                raise MemoryError(f"HIP out of memory: {err}")
            raise RuntimeError(f"bsim_hip failed: {err}")

    def run(self, prog, inputs: dict, params: dict | None = None, input_arrays: dict | None = None,
            bitlength: int = 0, iterations: int = 1, feedback: dict | None = None,
            bsdata_path: str | None = None) -> tuple[dict, int, float]:
        """(outputs as Python ints, dynamic op count, exec_ms) of one case."""
        cmd = self._cmd(prog)
        tmp = None
        if self.binary_outputs:
            import tempfile
            fd, tmp = tempfile.mkstemp(suffix=".bsdata")
            os.close(fd)
            cmd += ["--output-bsdata", tmp]
        try:
            return self._run(cmd, tmp, inputs, params, input_arrays, bitlength, iterations, feedback, bsdata_path)
        finally:
            if tmp:
                os.unlink(tmp)

    def _run(self, cmd, tmp, inputs, params, input_arrays, bitlength, iterations, feedback, bsdata_path):
        if bsdata_path:
            proc = subprocess.run(cmd + ["--input", bsdata_path], capture_output=True, text=True,
                                  timeout=self.timeout, env=subprocess_env())
        else:
            case = build_input_json(inputs, params, input_arrays, bitlength, iterations, feedback)
            proc = subprocess.run(cmd, input=json.dumps(case), capture_output=True, text=True,
                                  timeout=self.timeout, env=subprocess_env())
        self._check(proc)
        if tmp:
            out = read_output_bsdata(tmp)[0]
            return out["outputs"], out["op_count"], out.get("exec_ms", 0.0)
        out = json.loads(proc.stdout)
        if isinstance(out, list):  # a .bsdata file with several cases: the first one
            out = out[0]
        return parse_outputs(out["outputs"]), out["op_count"], out.get("exec_ms", 0.0)

    def run_batch(self, prog, bsdata_paths: list[str]) -> list[dict]:
        """The raw per-case JSON objects of every case of every file, one
        bsim_hip process per file (as the harness does for bsim_cuda)."""
        results = []
        for p in bsdata_paths:
            proc = subprocess.run(self._cmd(prog) + ["--input", p], capture_output=True, text=True,
                                  timeout=self.timeout, env=subprocess_env())
            self._check(proc)
            out = json.loads(proc.stdout)
            results.extend(out if isinstance(out, list) else [out])
        return results


def make_backend(name: str):
    """Backend names this package adds to the harness's make_backend."""
    if name == "hip":
        return HipBackend()
    if name == "hip_perop":
        return HipBackend(executor="perop")
    raise ValueError(f"unknown backend {name!r} (hip, hip_perop)")
