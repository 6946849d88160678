"""Learned schedule cost model.

A small MLP (see include/bsz_hip.h, bsz_mlp_*) predicts log(run time) of a
program under a candidate schedule from features of the fused plan that
schedule produces: the partition (kernels, nodes, stateful ops, streams read
and written), the tile configuration (rows per chunk, rows in flight,
occupancy, workgroup width) and the problem size.  Features come from
planning alone (`bsz_plan_describe`: the program is recorded once and
partitioned per schedule, no code generated; milliseconds per candidate), so
ranking a few dozen candidates costs far less than compiling and timing them.
Training and inference run on the GPU in libbsz_hip.so (BF16 MFMA products,
FP32 accumulation, Adam).  The measured autotuner (tuning.py) produces the
training data and stays the judge: the model only chooses which schedule to
build; it never touches program outputs.
"""

from __future__ import annotations

import ctypes as C
import json
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib, api
from ._lib import check, lib

FEATURES = ["log_words", "log_instances", "log_chunk_rows", "log_rows_per_step", "log_waves_per_block",
            "occupancy_req", "log_max_live", "log_max_ops", "log_kernels", "log_stateful_kernels",
            "log_stream_reads", "log_stream_writes", "log_nodes", "log_max_nodes", "log_mean_nodes",
            "log_state_words", "log_max_state", "mean_eff_rows_per_step", "mean_occupancy", "min_occupancy",
            "log_bytes", "log_valu", "log_chunks", "log_launches", "log_pops", "log_nodes_per_stream",
            "log_bytes_per_chunk", "log_valu_per_chunk"]
N_FEATURES = len(FEATURES)


def _l2(x) -> float:
    return math.log2(max(1.0, float(x)))


def plan_features(info: dict, options: dict, words_per_instance: float, n_instances: int) -> np.ndarray:
    """Feature vector of one (program, schedule) from its plan description."""
    ks = info["kernels"]
    nk = max(1, len(ks))
    nodes = [k["nodes"] for k in ks] or [0]
    reads = sum(k["inputs"] for k in ks)
    writes = sum(k["outputs"] for k in ks)
    state = [k["state_words"] for k in ks] or [0]
    stateful = sum(1 for k in ks if k["state_words"] > 0)
    cr = int(options.get("chunk_rows", info.get("chunk_rows", 16)))
    rows = max(1.0, words_per_instance / 64.0)
    chunks = n_instances * math.ceil(rows / cr)
    words = words_per_instance * n_instances
    byts = 8.0 * words * (reads + writes)
    valu = words * sum(nodes)
    occ = [k["occupancy"] for k in ks] or [4]
    f = [_l2(words_per_instance), _l2(n_instances), _l2(cr), _l2(options.get("rows_per_step", 8)),
         _l2(options.get("waves_per_block", 4)), float(options.get("occupancy", 0)),
         _l2(options.get("max_live", 96)), _l2(options.get("max_ops", 4000)), _l2(nk), _l2(stateful),
         _l2(reads), _l2(writes), _l2(sum(nodes)), _l2(max(nodes)), _l2(sum(nodes) / nk), _l2(sum(state)),
         _l2(max(state)), float(np.mean([k["rows_per_step"] for k in ks] or [1])), float(np.mean(occ)),
         float(min(occ)), _l2(byts), _l2(valu), _l2(chunks), _l2(nk + stateful + 1),
         _l2(sum(k["pops"] for k in ks)), _l2(sum(nodes) / max(1, reads + writes)), _l2(byts / max(1, chunks)),
         _l2(valu / max(1, chunks))]
    return np.array(f, dtype=np.float32)


def features_for(program: api.Program, shape: api.Case, options: dict, words_per_instance: float,
                 n_instances: int) -> np.ndarray:
    return features_batch(program, shape, [options], words_per_instance, n_instances)[0]


def features_batch(program: api.Program, shape: api.Case, options_list: list[dict], words_per_instance: float,
                   n_instances: int) -> np.ndarray:
    """Features of every schedule of `options_list`: the program is recorded
    once and partitioned per schedule (bsz_plan_describe), no code generated."""
    infos = api.describe_plans(program, shape, options_list)
    return np.stack([plan_features(info, o, words_per_instance, n_instances) for info, o in zip(infos, options_list)])


class GpuMlp:
    """The MLP of libbsz_hip.so (bsz_mlp_*): n_in -> h1 -> h2 -> 1."""

    def __init__(self, ctx: api.Context, n_in: int, h1: int = 64, h2: int = 64, seed: int = 0):
        h = C.c_void_p()
        check(lib().bsz_mlp_create(ctx.h, n_in, h1, h2, seed, C.byref(h)), ctx.h)
        self.ctx, self.h, self.n_in, self.h1, self.h2 = ctx, h, n_in, h1, h2

    def close(self):
        if self.h:
            lib().bsz_mlp_free(self.h)
            self.h = None

    def __del__(self):
        if getattr(self, "h", None):
            self.close()

    @staticmethod
    def _f(a):
        return np.ascontiguousarray(a, dtype=np.float32)

    @staticmethod
    def _p(a):
        return a.ctypes.data_as(_lib.f32p)

    @property
    def n_params(self) -> int:
        return int(lib().bsz_mlp_n_params(self.h))

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.n_params, dtype=np.float32)
        check(lib().bsz_mlp_get_params(self.h, self._p(out)), self.ctx.h)
        return out

    def set_params(self, p):
        p = self._f(p)
        assert len(p) == self.n_params
        check(lib().bsz_mlp_set_params(self.h, self._p(p)), self.ctx.h)

    def predict(self, X) -> np.ndarray:
        X = self._f(X)
        assert X.ndim == 2 and X.shape[1] == self.n_in
        out = np.zeros(len(X), dtype=np.float32)
        check(lib().bsz_mlp_predict(self.h, self._p(X), len(X), self._p(out)), self.ctx.h)
        return out

    def gradients(self, X, y):
        X, y = self._f(X), self._f(y)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = C.c_float()
        check(lib().bsz_mlp_gradients(self.h, self._p(X), self._p(y), len(X), self._p(g), C.byref(loss)),
              self.ctx.h)
        return float(loss.value), g

    def train_step(self, X, y, lr: float) -> float:
        X, y = self._f(X), self._f(y)
        loss = C.c_float()
        check(lib().bsz_mlp_train_step(self.h, self._p(X), self._p(y), len(X), lr, C.byref(loss)), self.ctx.h)
        return float(loss.value)

    def train(self, X, y, epochs: int, batch: int, lr: float, seed: int = 0) -> float:
        X, y = self._f(X), self._f(y)
        loss = C.c_float()
        check(lib().bsz_mlp_train(self.h, self._p(X), self._p(y), len(X), epochs, batch, lr, seed, C.byref(loss)),
              self.ctx.h)
        return float(loss.value)


@dataclass
class CostModel:
    """Standardized features -> an ensemble of GpuMlp -> standardized log(ms).

    Members differ only in their initialisation and minibatch order; the
    prediction is their mean, which generalizes better to programs larger
    than the training ones than one network trained longer."""
    ctx: api.Context
    hidden: int = 32
    members: int = 5
    seed: int = 0
    mu: np.ndarray = field(default=None)
    sd: np.ndarray = field(default=None)
    y_mu: float = 0.0
    y_sd: float = 1.0
    nets: list = field(default_factory=list)
    train_seconds: float = 0.0

    def fit(self, X: np.ndarray, ms: np.ndarray, epochs: int = 80, batch: int = 64, lr: float = 2e-3) -> float:
        t0 = time.perf_counter()
        X = np.asarray(X, dtype=np.float64)
        y = np.log(np.asarray(ms, dtype=np.float64))
        self.mu, self.sd = X.mean(axis=0), X.std(axis=0) + 1e-6
        self.y_mu, self.y_sd = float(y.mean()), float(y.std() + 1e-6)
        Z, yz = (X - self.mu) / self.sd, (y - self.y_mu) / self.y_sd
        for n in self.nets:
            n.close()
        self.nets, losses = [], []
        for k in range(self.members):
            net = GpuMlp(self.ctx, X.shape[1], self.hidden, self.hidden, self.seed * 101 + k)
            losses.append(net.train(Z, yz, epochs, batch, lr, self.seed * 101 + k))
            self.nets.append(net)
        self.train_seconds = time.perf_counter() - t0
        return float(np.mean(losses))

    def predict_ms(self, X: np.ndarray) -> np.ndarray:
        Z = (np.asarray(X, dtype=np.float64) - self.mu) / self.sd
        z = np.mean([n.predict(Z).astype(np.float64) for n in self.nets], axis=0)
        return np.exp(z * self.y_sd + self.y_mu)

    def state(self) -> dict:
        return {"mu": self.mu.tolist(), "sd": self.sd.tolist(), "y_mu": self.y_mu, "y_sd": self.y_sd,
                "hidden": self.hidden, "params": [n.get_params().tolist() for n in self.nets]}

    def save(self, path: str):
        with open(path, "w") as f:
            json.dump(self.state(), f)

    @classmethod
    def load(cls, ctx: api.Context, path: str) -> "CostModel":
        with open(path) as f:
            s = json.load(f)
        m = cls(ctx, hidden=s["hidden"], members=len(s["params"]))
        m.mu, m.sd = np.array(s["mu"]), np.array(s["sd"])
        m.y_mu, m.y_sd = s["y_mu"], s["y_sd"]
        for p in s["params"]:
            net = GpuMlp(ctx, len(m.mu), m.hidden, m.hidden)
            net.set_params(np.array(p, dtype=np.float32))
            m.nets.append(net)
        return m
