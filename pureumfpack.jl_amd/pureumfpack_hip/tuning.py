"""Gemm tile autotuner and its learned cost model.

The grouped MFMA gemm (Schur complements and panel updates) has several tile shapes
(mf_gemm_tile_info).  A tile policy maps (use, size bucket) -> shape, per precision, and is
read when a plan is created (mf_set_gemm_tiles).  Every shape accumulates each element of C in
the same order, so a policy changes speed only, never a result bit.

* Measured autotuning (`Autotuner.exhaustive`): time every shape on representative problems of
  every bucket with mf_time_gemm and keep the fastest.
* Learned (`Autotuner.with_model`): time a fraction of the (problem, shape) pairs, train the
  small MLP (mf_mlp_*: forward, backward and Adam run on the GPU) on log-time, predict the
  rest and pick the predicted fastest; measured timing decides: each pick is timed
  against the default shape and loses to it when slower.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from ._lib import MF_OK, c_f64p

c_i32p = _lib.c_i32p
c_f32p = _lib.c_f32p

PREC = {"fp64": 0, "fp32": 1, "bf16": 2}


def _check(rc, what):
    if rc != MF_OK:
        lib = _lib.load()
        raise RuntimeError(f"{what}: {lib.mf_last_error().decode()} {lib.mf_mlp_last_error().decode()}")


def tile_info():
    """(tile_m, tile_n, bucket_hi) of the library's gemm tile shapes and size buckets."""
    lib = _lib.load()
    nt, nb = ctypes.c_int32(), ctypes.c_int32()
    tm = np.zeros(16, np.int32); tn = np.zeros(16, np.int32); hi = np.zeros(16, np.int32)
    _check(lib.mf_gemm_tile_info(ctypes.byref(nt), tm.ctypes.data_as(c_i32p), tn.ctypes.data_as(c_i32p),
                                 ctypes.byref(nb), hi.ctypes.data_as(c_i32p)), "mf_gemm_tile_info")
    return tm[:nt.value].copy(), tn[:nt.value].copy(), hi[:nb.value - 1].copy()


def bucket_of(m, n, hi):
    s = math.sqrt(m * n)
    b = 0
    while b < len(hi) and s >= hi[b]:
        b += 1
    return b


def set_policy(precision, policy):
    """Install a policy: int array (2, nbuckets) = tile index per (use, bucket)."""
    lib = _lib.load()
    pol = np.ascontiguousarray(policy, np.int32).ravel()
    _check(lib.mf_set_gemm_tiles(PREC[precision], pol.ctypes.data_as(c_i32p)), "mf_set_gemm_tiles")


def get_policy(precision):
    lib = _lib.load()
    _, _, hi = tile_info()
    pol = np.zeros(2 * (len(hi) + 1), np.int32)
    _check(lib.mf_get_gemm_tiles(PREC[precision], pol.ctypes.data_as(c_i32p)), "mf_get_gemm_tiles")
    return pol.reshape(2, -1)


def time_gemm(precision, m, n, k, nprob, tile, reps=10):
    """Seconds per launch of nprob independent m x n x k problems with the given tile shape."""
    lib = _lib.load()
    sec = ctypes.c_double()
    _check(lib.mf_time_gemm(PREC[precision], int(m), int(n), int(k), int(nprob), int(tile), int(reps),
                            ctypes.byref(sec)), "mf_time_gemm")
    return sec.value


class MLP:
    """The GPU network (mf_mlp): d_in -> hidden -> hidden -> 1, ReLU, FP32."""

    def __init__(self, d_in, hidden=32, seed=0):
        self.lib = _lib.load()
        self.d_in, self.hidden = int(d_in), int(hidden)
        h = ctypes.c_void_p()
        _check(self.lib.mf_mlp_create(self.d_in, self.hidden, int(seed), ctypes.byref(h)), "mf_mlp_create")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            self.lib.mf_mlp_destroy(h)
            self.h = None

    @property
    def num_params(self):
        return int(self.lib.mf_mlp_num_params(self.h))

    def get_params(self):
        p = np.zeros(self.num_params, np.float32)
        _check(self.lib.mf_mlp_get_params(self.h, p.ctypes.data_as(c_f32p)), "mf_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params
        _check(self.lib.mf_mlp_set_params(self.h, p.ctypes.data_as(c_f32p)), "mf_mlp_set_params")

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        y = np.zeros(X.shape[0], np.float32)
        _check(self.lib.mf_mlp_forward(self.h, X.ctypes.data_as(c_f32p), X.shape[0], y.ctypes.data_as(c_f32p)),
               "mf_mlp_forward")
        return y

    def gradient(self, X, t):
        X = np.ascontiguousarray(X, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.num_params, np.float32)
        _check(self.lib.mf_mlp_gradient(self.h, X.ctypes.data_as(c_f32p), t.ctypes.data_as(c_f32p), X.shape[0],
                                        loss.ctypes.data_as(c_f32p), g.ctypes.data_as(c_f32p)), "mf_mlp_gradient")
        return float(loss[0]), g

    def train(self, X, t, epochs, lr):
        X = np.ascontiguousarray(X, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = np.zeros(1, np.float32)
        _check(self.lib.mf_mlp_train(self.h, X.ctypes.data_as(c_f32p), t.ctypes.data_as(c_f32p), X.shape[0],
                                     int(epochs), float(lr), loss.ctypes.data_as(c_f32p)), "mf_mlp_train")
        return float(loss[0])


def features(m, n, k, nprob, fp64, tile, tile_m, tile_n):
    """Feature vector of one (problem, tile) pair: problem size, workgroup count, fill of the
    edge tiles, precision and the tile shape (one-hot)."""
    tm, tn = tile_m[tile], tile_n[tile]
    wg = math.ceil(m / tm) * math.ceil(n / tn) * nprob
    fill = (m * n) / (math.ceil(m / tm) * tm * math.ceil(n / tn) * tn)
    f = [math.log2(m), math.log2(n), math.log2(k), math.log2(nprob), math.log2(wg), fill,
         math.log2(2.0 * m * n * k * nprob), float(fp64)]
    f += [1.0 if t == tile else 0.0 for t in range(len(tile_m))]
    return f


class CostModel:
    """log(seconds) of a gemm launch from (m, n, k, nprob, precision, tile), learned by the GPU MLP."""

    def __init__(self, tile_m, tile_n, hidden=32, seed=0):
        self.tile_m, self.tile_n = tile_m, tile_n
        self.d_in = 8 + len(tile_m)
        self.net = MLP(self.d_in, hidden, seed)
        self.mu = None
        self.sd = None

    def _x(self, rows):
        X = np.array([features(*r, self.tile_m, self.tile_n) for r in rows], np.float64)
        if self.mu is None:
            self.mu = X.mean(axis=0)
            self.sd = X.std(axis=0)
            self.sd[self.sd < 1e-6] = 1.0
        return ((X - self.mu) / self.sd).astype(np.float32)

    def fit(self, rows, seconds, epochs=1500, lr=3e-3):
        """rows: (m, n, k, nprob, fp64, tile) tuples; seconds: measured times."""
        self.mu = None
        X = self._x(rows)
        y = np.log(np.asarray(seconds, np.float64))
        self.ymu, self.ysd = float(y.mean()), float(y.std() or 1.0)
        t = ((y - self.ymu) / self.ysd).astype(np.float32)
        return self.net.train(X, t, epochs, lr)

    def predict(self, rows):
        X = self._x(rows)
        return np.exp(self.net.forward(X).astype(np.float64) * self.ysd + self.ymu)


class Autotuner:
    """Tile policy for one precision from measured and/or predicted gemm times.

    Each (use, bucket) is represented by a few problems: use 0 (panel updates) m = n = s,
    k = 64; use 1 (Schur complements) m = n = s, k = s / 2; s the bucket's typical size, with a
    single problem and with a level's worth of independent problems."""

    def __init__(self, precision="fp32", reps=10):
        self.precision = precision
        self.fp64 = precision == "fp64"
        self.reps = reps
        self.tile_m, self.tile_n, self.hi = tile_info()
        self.nt = len(self.tile_m)
        self.nb = len(self.hi) + 1
        edges = [32] + list(self.hi) + [2 * self.hi[-1]]
        self.shapes = {}   # (use, bucket) -> [(m, n, k, nprob)]
        for b in range(self.nb):
            s = int(round(math.sqrt(edges[b] * edges[b + 1])))
            many = 16 if s < 256 else (4 if s < 1024 else 1)
            for use in (0, 1):
                k = 64 if use == 0 else max(16, s // 2)
                self.shapes[(use, b)] = [(s, s, k, 1)] + ([(s, s, k, many)] if many > 1 else [])
        self.timings = 0
        self._cache = {}

    def measure(self, shape, tile):
        key = (shape, tile)
        if key not in self._cache:
            self._cache[key] = time_gemm(self.precision, *shape, tile, self.reps)
            self.timings += 1
        return self._cache[key]

    def pairs(self):
        return [(ub, sh, t) for ub, shs in self.shapes.items() for sh in shs for t in range(self.nt)]

    def exhaustive(self):
        """Time every shape on every representative problem; the fastest total per bucket wins."""
        pol = np.zeros((2, self.nb), np.int32)
        for (use, b), shs in self.shapes.items():
            tot = [sum(self.measure(sh, t) for sh in shs) for t in range(self.nt)]
            pol[use, b] = int(np.argmin(tot))
        return pol

    def with_model(self, frac=0.4, seed=0, epochs=1500, lr=3e-3, hidden=32):
        """Train the cost model on a random fraction of the pairs, predict the rest, verify each
        pick against the default shape by measurement.  Returns (policy, model)."""
        rng = np.random.default_rng(seed)
        allp = self.pairs()
        idx = rng.permutation(len(allp))[:max(self.nt * 2, int(frac * len(allp)))]
        rows, secs = [], []
        for i in idx:
            _, sh, t = allp[i]
            rows.append((*sh, self.fp64, t))
            secs.append(self.measure(sh, t))
        model = CostModel(self.tile_m, self.tile_n, hidden=hidden, seed=seed)
        self.train_loss = model.fit(rows, secs, epochs=epochs, lr=lr)
        pol = np.zeros((2, self.nb), np.int32)
        for (use, b), shs in self.shapes.items():
            pred = [sum(model.predict([(*sh, self.fp64, t)])[0] for sh in shs) for t in range(self.nt)]
            pick = int(np.argmin(pred))
            if pick != 0:
                # the measured pick must beat the measured default
                mp = sum(self.measure(sh, pick) for sh in shs)
                md = sum(self.measure(sh, 0) for sh in shs)
                if mp >= md:
                    pick = 0
            pol[use, b] = pick
        return pol, model

    def policy_time(self, pol, reps=None):
        """Measured total time of the representative problems under a policy (fresh timings)."""
        r = reps or self.reps
        return sum(time_gemm(self.precision, *sh, int(pol[use, b]), r)
                   for (use, b), shs in self.shapes.items() for sh in shs)


def time_gemm_list(precision, mnk, tile, reps=10):
    """Seconds per launch of the independent problems mnk = [(m, n, k), ...] with one tile shape."""
    lib = _lib.load()
    a = np.ascontiguousarray(np.asarray(mnk, np.int32).reshape(-1, 3))
    sec = ctypes.c_double()
    _check(lib.mf_time_gemm_list(PREC[precision], a.shape[0], a.ctypes.data_as(c_i32p), int(tile), int(reps),
                                 ctypes.byref(sec)), "mf_time_gemm_list")
    return sec.value


def schur_launches(S, small_nf=64):
    """The Schur-complement gemm launches a plan on analysis S issues, by size bucket: for every
    etree level, the (m, n, k) = (nu, nu, np) of its large fronts (nf > small_nf, nu > 0),
    split like the plan into fronts with nf >= 1024 and the rest and grouped by bucket.
    Returns {bucket: [launch, ...]}, a launch being a tuple of (m, n, k)."""
    _, _, hi = tile_info()
    npv = np.diff(np.asarray(S.sstart))
    nu = np.diff(np.asarray(S.upd_ptr))
    nf = npv + nu
    cp, ci = np.asarray(S.child_ptr), np.asarray(S.child_idx)
    level = np.zeros(S.nsuper, np.int64)
    for s in range(S.nsuper):
        for c in ci[cp[s]:cp[s + 1]]:
            level[s] = max(level[s], level[c] + 1)
    out = {}
    for lv in range(int(level.max()) + 1 if S.nsuper else 0):
        idx = np.nonzero((level == lv) & (nf > small_nf) & (nu > 0))[0]
        for big in (True, False):
            sel = [s for s in idx if (nf[s] >= 1024) == big]
            by_b = {}
            for s in sel:
                by_b.setdefault(bucket_of(int(nu[s]), int(nu[s]), hi), []).append((int(nu[s]), int(nu[s]), int(npv[s])))
            for b, probs in by_b.items():
                out.setdefault(b, []).append(tuple(probs))
    return out


def launch_features(probs, fp64, tile, tile_m, tile_n):
    """Feature vector of one (launch, tile) pair: the launch's size (flops, problems, largest and
    mean dimensions), the workgroups and the fill of the edge tiles under the shape, the
    precision and the shape (one-hot)."""
    tm, tn = tile_m[tile], tile_n[tile]
    m = np.array([p[0] for p in probs], float)
    n = np.array([p[1] for p in probs], float)
    k = np.array([p[2] for p in probs], float)
    wg = float((np.ceil(m / tm) * np.ceil(n / tn)).sum())
    fill = float((m * n).sum() / (np.ceil(m / tm) * tm * np.ceil(n / tn) * tn).sum())
    f = [math.log2(2.0 * (m * n * k).sum()), math.log2(len(probs)), math.log2(m.max()), math.log2(m.mean()),
         math.log2(k.mean()), math.log2(wg), fill, float(fp64)]
    f += [1.0 if t == tile else 0.0 for t in range(len(tile_m))]
    return f


class PlanTuner:
    """Tile policy for the Schur complements (use 1) from the plan's own gemm launches.

    `exhaustive` times every launch of every bucket under every tile shape; `with_model` times a
    fraction of the (launch, shape) pairs, trains the GPU MLP on log-time, predicts the rest,
    and keeps a predicted pick only when measurement confirms it beats the default shape.  Panel
    updates (use 0) keep the default."""

    def __init__(self, S, precision="fp32", reps=5, small_nf=64):
        self.precision = precision
        self.fp64 = precision == "fp64"
        self.reps = reps
        self.tile_m, self.tile_n, self.hi = tile_info()
        self.nt = len(self.tile_m)
        self.nb = len(self.hi) + 1
        self.launches = schur_launches(S, small_nf)
        self.timings = 0
        self._cache = {}

    def measure(self, probs, tile):
        key = (probs, tile)
        if key not in self._cache:
            self._cache[key] = time_gemm_list(self.precision, probs, tile, self.reps)
            self.timings += 1
        return self._cache[key]

    def bucket_time(self, b, tile):
        return sum(self.measure(p, tile) for p in self.launches.get(b, []))

    def pairs(self):
        return [(b, p, t) for b, ls in self.launches.items() for p in ls for t in range(self.nt)]

    def _base(self):
        pol = np.zeros((2, self.nb), np.int32)
        pol[0] = get_policy(self.precision)[0]   # panel updates keep the installed shapes
        return pol

    def exhaustive(self):
        pol = self._base()
        for b in self.launches:
            pol[1, b] = int(np.argmin([self.bucket_time(b, t) for t in range(self.nt)]))
        return pol

    def with_model(self, frac=0.35, seed=0, epochs=1500, lr=3e-3, hidden=32):
        rng = np.random.default_rng(seed)
        allp = self.pairs()
        take = rng.permutation(len(allp))[:max(self.nt * 3, int(frac * len(allp)))]
        rows = [launch_features(allp[i][1], self.fp64, allp[i][2], self.tile_m, self.tile_n) for i in take]
        secs = [self.measure(allp[i][1], allp[i][2]) for i in take]
        X = np.array(rows, np.float64)
        self.mu, self.sd = X.mean(axis=0), X.std(axis=0)
        self.sd[self.sd < 1e-6] = 1.0
        y = np.log(np.asarray(secs, np.float64))
        ymu, ysd = float(y.mean()), float(y.std() or 1.0)
        net = MLP(X.shape[1], hidden, seed)
        self.train_loss = net.train(((X - self.mu) / self.sd).astype(np.float32),
                                    ((y - ymu) / ysd).astype(np.float32), epochs, lr)

        def predict(probs, t):
            f = (np.array([launch_features(probs, self.fp64, t, self.tile_m, self.tile_n)]) - self.mu) / self.sd
            return float(np.exp(net.forward(f.astype(np.float32))[0] * ysd + ymu))

        pol = self._base()
        self.model_picks = {}
        for b, ls in self.launches.items():
            pred = [sum(predict(p, t) for p in ls) for t in range(self.nt)]
            pick = int(np.argmin(pred))
            self.model_picks[b] = pick
            if pick != 0 and self.bucket_time(b, pick) >= self.bucket_time(b, 0):
                pick = 0   # measurement overrules a wrong prediction
            pol[1, b] = pick
        return pol

    def policy_time(self, pol):
        """Measured time of all Schur launches under a policy (cached timings)."""
        return sum(self.bucket_time(b, int(pol[1, b])) for b in self.launches)
