"""Learned k-th-neighbour radius: features, the MFMA MLP, training and seeding.

The network (12 -> 64 -> 64 -> 64 -> 1, ReLU, bf16 weights on MFMA with fp32
accumulation) predicts log d2_k of a self query from tree-density features
(fasttree_amd/csrc/ft_learned.hip).  It is trained on the GPU by the library's
own forward/backward kernels on exact kNN distances of seeded snapshots, and
its prediction plus a calibrated margin seeds the packet traversal's bound.  A
query that finds fewer than k neighbours inside the seed is re-run unbounded,
so results are identical with or without the model.

Train: python -m fasttree_amd.learned train --out fasttree_amd/models/<name>.npz
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

from . import (ALGO_PACKET, COORD_F32, COORD_F64, COORD_U64, SFC_PEANO_HILBERT, BvhScratch,
               Config, DeviceArray, TreeSoA, _check, _ptr, _sh, lib, particles,
               self_knn_query_subset)
from . import synth

NF = 16
NPARAM = 9488
NNORM = 2 * NF + 2
MODELS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models")

_bound = False


def _bind():
    global _bound
    if _bound:
        return
    L = lib()
    vp, i32 = C.c_void_p, C.c_int
    L.ft_mlp_create.argtypes = [C.POINTER(C.c_void_p)]
    L.ft_mlp_destroy.argtypes = [vp]
    L.ft_mlp_set.argtypes = [vp, vp, vp, C.c_double]
    L.ft_mlp_get.argtypes = [vp, vp, vp, C.POINTER(C.c_double)]
    L.ft_knn_features.argtypes = [vp, vp, vp, i32, i32, vp]
    L.ft_mlp_forward.argtypes = [vp, vp, vp, i32, vp]
    L.ft_mlp_grad.argtypes = [vp, vp, vp, vp, i32, vp, vp]
    L.ft_mlp_train_step.argtypes = [vp, vp, vp, vp, i32, C.c_float]
    L.ft_mlp_last_loss.argtypes = [vp, vp, C.POINTER(C.c_float)]
    L.ft_mlp_seed.argtypes = [vp, vp, vp, vp, i32, i32, vp, vp]
    _bound = True


def features(tree: TreeSoA, k: int, leaf_ids=None, out: DeviceArray | None = None, stream=None):
    """Feature rows (float32[nq, 16], device) for self queries of `tree`."""
    _bind()
    nq = tree.num_leaves if leaf_ids is None else (
        leaf_ids.n if isinstance(leaf_ids, DeviceArray) else len(leaf_ids))
    ids = leaf_ids
    if isinstance(ids, np.ndarray):
        ids = DeviceArray.from_host(ids, np.int32)
    if out is None:
        out = DeviceArray(nq * NF, np.float32)
    _check(lib().ft_knn_features(_sh(stream), C.byref(tree.c), _ptr(ids), int(nq), int(k),
                                 out.ptr))
    return out


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases; the padded input rows of W1 are zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)

    def he(fan_in, shape):
        b = np.sqrt(6.0 / fan_in)
        return rng.uniform(-b, b, shape).astype(np.float32)

    w1 = np.zeros((16, 64), np.float32)
    w1[:12] = he(12, (12, 64))
    p[0:1024] = w1.ravel()
    p[1088:1088 + 4096] = he(64, (64, 64)).ravel()
    p[5248:5248 + 4096] = he(64, (64, 64)).ravel()
    p[9408:9408 + 64] = he(64, 64) * 0.1
    return p


class MLP:
    """Device-resident model: fp32 master weights, Adam state, normalisation."""

    def __init__(self, params=None, norm=None, margin=0.0):
        _bind()
        h = C.c_void_p()
        _check(lib().ft_mlp_create(C.byref(h)))
        self.h = h.value
        self.set(params if params is not None else init_params(),
                 norm if norm is not None else default_norm(), margin)

    def set(self, params=None, norm=None, margin=None):
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        n = None if norm is None else np.ascontiguousarray(norm, np.float32)
        if margin is None:
            margin = self.get()[2]
        _check(lib().ft_mlp_set(self.h, _ptr(p), _ptr(n), float(margin)))

    def get(self):
        p = np.zeros(NPARAM, np.float32)
        n = np.zeros(NNORM, np.float32)
        m = C.c_double()
        _check(lib().ft_mlp_get(self.h, p.ctypes.data, n.ctypes.data, C.byref(m)))
        return p, n, m.value

    def forward(self, feats: DeviceArray, n: int, stream=None) -> np.ndarray:
        out = DeviceArray(n, np.float32)
        _check(lib().ft_mlp_forward(_sh(stream), self.h, feats.ptr, int(n), out.ptr))
        return out.to_host()

    def grad(self, feats: DeviceArray, target: DeviceArray, n: int, stream=None):
        g = DeviceArray(NPARAM + 16, np.float32)
        loss = DeviceArray(1, np.float32)
        _check(lib().ft_mlp_grad(_sh(stream), self.h, feats.ptr, target.ptr, int(n), g.ptr,
                                 loss.ptr))
        return g.to_host()[:NPARAM], float(loss.to_host()[0])

    def train_step(self, feats_ptr: int, target_ptr: int, n: int, lr: float, stream=None):
        _check(lib().ft_mlp_train_step(_sh(stream), self.h, feats_ptr, target_ptr, int(n),
                                       float(lr)))

    def last_loss(self, stream=None) -> float:
        v = C.c_float()
        _check(lib().ft_mlp_last_loss(_sh(stream), self.h, C.byref(v)))
        return v.value

    def seed(self, tree: TreeSoA, k: int, seed_out: DeviceArray, scratch: DeviceArray,
             leaf_ids=None, stream=None):
        nq = tree.num_leaves if leaf_ids is None else leaf_ids.n
        _check(lib().ft_mlp_seed(_sh(stream), self.h, C.byref(tree.c), _ptr(leaf_ids), int(nq),
                                 int(k), scratch.ptr, seed_out.ptr))

    def save(self, path, meta=None):
        p, n, m = self.get()
        np.savez(path, params=p, norm=n, margin=np.array([m]),
                 meta=np.frombuffer(json.dumps(meta or {}).encode(), np.uint8))

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        return cls(z["params"], z["norm"], float(z["margin"][0]))

    def __del__(self):
        try:
            if self.h:
                lib().ft_mlp_destroy(self.h)
                self.h = None
        except Exception:
            pass


def default_norm():
    n = np.zeros(NNORM, np.float32)
    n[NF:2 * NF] = 1.0
    n[2 * NF + 1] = 1.0
    return n


def norm_from(feats: np.ndarray, y: np.ndarray):
    f = feats.reshape(-1, NF).astype(np.float64)
    n = np.zeros(NNORM, np.float32)
    mu = f.mean(0)
    sd = f.std(0)
    n[:NF] = mu
    n[NF:2 * NF] = np.where(sd > 1e-8, 1.0 / np.maximum(sd, 1e-8), 0.0)
    n[12:16] = 0.0
    n[NF + 12:2 * NF] = 0.0
    n[2 * NF] = float(np.mean(y))
    n[2 * NF + 1] = float(max(np.std(y), 1e-6))
    return n


# ---------------------------------------------------------------------------
# Training data: exact kNN of seeded clustered snapshots
# ---------------------------------------------------------------------------
def snapshot_tree(n, seed, bits=21, coord="u64", n_halos=2000):
    """Seeded NFW-clustered snapshot in a periodic box L = 100, built on the
    device: 21-bit fixed-point int64 coordinates ("u64") or the double
    coordinates themselves ("f64"), 63-bit Peano-Hilbert keys."""
    if coord == "f64":
        x, y, z = synth.nfw_clustered(n, seed, box=100.0, n_halos=n_halos)
        cfg = Config(sfc=SFC_PEANO_HILBERT, bits_per_dim=bits, periodic=True, box_length=100.0)
        tree = TreeSoA(n, COORD_F64, cfg)
    else:
        x, y, z = synth.clustered_fixed(n, seed, bits=bits, box=100.0, n_halos=n_halos)
        cfg = Config(sfc=SFC_PEANO_HILBERT, bits_per_dim=bits, periodic=True)
        tree = TreeSoA(n, COORD_U64, cfg)
    p = particles(*(DeviceArray.from_host(a) for a in (x, y, z)))
    build_bvh_dev(p, tree)
    return tree


def build_bvh_dev(p, tree):
    from . import build_bvh
    build_bvh(p, tree)


def target_eps(tree: TreeSoA) -> float:
    """The offset of the regression target log(d2_k + eps): 1 for integer
    coordinates; for FP coordinates 1e-12 * (largest root extent)^2, the
    features' own offset (fasttree_amd/csrc/ft_learned.hip, features_kernel)."""
    if tree.coord_type not in (COORD_F32, COORD_F64):
        return 1.0
    dt = np.float64 if tree.coord_type == COORD_F64 else np.float32
    ext = []
    for lo, hi in (("min_x", "max_x"), ("min_y", "max_y"), ("min_z", "max_z")):
        a, b = np.empty(1, dt), np.empty(1, dt)
        _check(lib().ft_memcpy_d2h(a.ctypes.data, getattr(tree.c, lo), a.nbytes, None))
        _check(lib().ft_memcpy_d2h(b.ctypes.data, getattr(tree.c, hi), b.nbytes, None))
        ext.append(float(b[0]) - float(a[0]))
    e = max(ext)
    return 1e-12 * e * e if e > 0 else 1e-300


def examples(tree: TreeSoA, k: int, sample: int, rng) -> tuple:
    """(features[sample,16], log(d2_k + eps)[sample]) for a random sample of leaves."""
    n = tree.num_leaves
    ids = np.sort(rng.choice(n, size=min(sample, n), replace=False)).astype(np.int32)
    dids = DeviceArray.from_host(ids)
    # the subset query returns exactly the self query's rows for these leaves
    _, dst = self_knn_query_subset(tree, dids, k, algorithm=ALGO_PACKET)
    dk = dst[:, k - 1].astype(np.float64)
    f = features(tree, k, leaf_ids=dids).to_host().reshape(-1, NF)
    return f, np.log(dk + target_eps(tree)).astype(np.float32)


def train(train_seeds, k=32, n=10_000_000, sample=1_000_000, steps=3000, batch=65536, lr=2e-3,
          fallback=0.002, log=print, coord="u64", ks=None, n_halos=2000):
    """Trains on `sample` queries from each seeded snapshot (split evenly over
    the k values in `ks`, default just `k`); returns (MLP, report)."""
    ks = list(ks) if ks else [k]
    rng = np.random.default_rng(1234)
    feats, ys = [], []
    t0 = time.time()
    for s in train_seeds:
        tree = snapshot_tree(n, s, coord=coord, n_halos=n_halos)
        for kk in ks:
            f, y = examples(tree, kk, sample // len(ks), rng)
            feats.append(f)
            ys.append(y)
            log(f"snapshot seed {s}, k={kk}: {len(y)} examples ({time.time() - t0:.1f}s)")
        del tree
    F = np.concatenate(feats)
    Y = np.concatenate(ys)
    m = MLP(init_params(7), norm_from(F, Y), 0.0)
    dF = DeviceArray.from_host(F.ravel(), np.float32)
    dY = DeviceArray.from_host(Y, np.float32)
    total = len(Y)
    nb = max(1, total // batch)
    for step in range(steps):
        b = step % nb
        if b == 0 and step > 0:
            perm = rng.permutation(total)
            F, Y = F[perm], Y[perm]
            dF = DeviceArray.from_host(F.ravel(), np.float32)
            dY = DeviceArray.from_host(Y, np.float32)
        cur_lr = lr * (0.1 ** (step / steps))
        m.train_step(dF.ptr + b * batch * NF * 4, dY.ptr + b * batch * 4, batch, cur_lr)
        if step % 500 == 0 or step == steps - 1:
            log(f"step {step}: loss {m.last_loss():.4f} (standardised MSE)")
    pred = m.forward(dF, total)
    resid = Y - pred
    margin = float(np.quantile(resid, 1.0 - fallback))
    m.set(margin=margin)
    rep = {"train_examples": total, "resid_std": float(resid.std()), "margin": margin,
           "train_fallback": float(np.mean(resid > margin)), "k": ks[0] if len(ks) == 1 else ks,
           "n": n, "coord": coord, "n_halos": n_halos,
           "train_seeds": list(train_seeds), "steps": steps, "batch": batch}
    return m, rep


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", default=os.path.join(MODELS, "nfw10m_int64_ph21_k32.npz"))
    t.add_argument("--seeds", default="1,2,3,4,5,6,7,8")
    t.add_argument("--n", type=int, default=10_000_000)
    t.add_argument("--k", type=int, default=32)
    t.add_argument("--ks", default="", help="comma-separated k values to train on (overrides --k)")
    t.add_argument("--coord", choices=("u64", "f64"), default="u64",
                   help="snapshot coordinates: 21-bit fixed-point int64 or double")
    t.add_argument("--halos", type=int, default=2000, help="NFW halos in each snapshot")
    t.add_argument("--sample", type=int, default=1_000_000)
    t.add_argument("--steps", type=int, default=3000)
    t.add_argument("--fallback", type=float, default=0.002,
                   help="training-set re-run fraction that sets the margin; 0.002 (margin ~0.7-0.8 in "
                        "log d2) times fastest on a training snapshot")
    a = ap.parse_args(argv)
    if a.cmd == "train":
        seeds = [int(s) for s in a.seeds.split(",")]
        ks = [int(v) for v in a.ks.split(",")] if a.ks else [a.k]
        m, rep = train(seeds, k=a.k, n=a.n, sample=a.sample, steps=a.steps, fallback=a.fallback,
                       log=lambda s: print(s, flush=True), coord=a.coord, ks=ks,
                       n_halos=a.halos)
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        m.save(a.out, rep)
        print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
