"""Learned start point for the frequency map (KSParMap.cc:397-440).

The reference's ParMap solves Omega_Kerr(p, e, iota) M_map = Omega_AK(v_map,
e, iota, s_map) M for (v_map, M_map, s_map) with a hybrid solver started at
(1/sqrt(p), M, s).  Here a small network (4-128-128-128-3, GELU, BF16 weights,
FP32 accumulation on the matrix cores; the library's own kernels for
inference, back-propagation and Adam) predicts the map from (a, p, e, iota),
and a Newton polish takes the prediction to the solver's tolerance, with the
hybrid solver as the fallback.  The map is linear in M, so the network works
at M = 1 and predicts M_map / M.

Workload of the learned form: a ~ U[0, 0.95], e ~ U[1e-3, 0.6],
iota ~ U(0.01, pi - 0.01), p ~ U[p_lso(a, e, iota) + 0.05, 15] (p_lso the
reference's interpolated last stable orbit); training draws seed 4, evaluation
draws seed 5.
"""
import ctypes as C
import os

import numpy as np

from . import _lib

N_PARAMS = 4 * 128 + 128 + 2 * (128 * 128 + 128) + 3 * 128 + 3
N_NORM = 14
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "freqmap_mlp.npz")


def p_lso(a, e, iota):
    """The reference's interpolated last-stable-orbit p (GKTrajFast P_lso)."""
    return _lib.lib().aak_p_lso(float(a), float(e), float(iota))


def sample_points(n, seed, threads=0):
    """n points (a, p, e, iota) from the learned form's workload."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.0, 0.95, n)
    e = rng.uniform(1e-3, 0.6, n)
    iota = rng.uniform(0.01, np.pi - 0.01, n)
    aei = np.ascontiguousarray(np.column_stack([a, e, iota]))
    lo = np.zeros(n)
    _lib.check(_lib.lib().aak_p_lso_batch(_lib.dptr(aei), n, _lib.dptr(lo), int(threads)), "p_lso_batch")
    lo += 0.05
    p = lo + rng.uniform(0.0, 1.0, n) * (15.0 - lo)
    return np.ascontiguousarray(np.column_stack([a, p, e, iota]))


def solve_maps(pts, mode=0, model=None, tol=1e-12, threads=0):
    """Frequency maps at M = 1: mode 0 the reference's hybrid solver from
    (1/sqrt(p), 1, a); mode 1 the model's start plus Newton polish (hybrid
    fallback).  Returns (maps, iters, resid, stats) with stats = geodesic ms,
    model ms, solve ms, fallbacks (host times summed over threads)."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    n = len(pts)
    maps = np.zeros((n, 3))
    iters = np.zeros(n, dtype=np.int32)
    resid = np.zeros(n)
    stats = np.zeros(4)
    rc = _lib.lib().aak_freqmap_batch(_lib.dptr(pts), n, int(mode), model._m if model else None, float(tol),
                                      int(threads), _lib.dptr(maps), _lib.iptr(iters), _lib.dptr(resid),
                                      _lib.dptr(stats))
    _lib.check(rc, "freqmap_batch")
    return maps, iters, resid, {"geodesic_ms": stats[0], "model_ms": stats[1], "solve_ms": stats[2],
                                "fallbacks": int(stats[3]), "failed": rc}


def features(pts):
    """(a, 1/sqrt(p), e, cos iota)."""
    return np.column_stack([pts[:, 0], 1.0 / np.sqrt(pts[:, 1]), pts[:, 2], np.cos(pts[:, 3])])


def corrections(pts, maps):
    """(v_map - 1/sqrt(p), M_map/M - 1, s_map - a)."""
    return np.column_stack([maps[:, 0] - 1.0 / np.sqrt(pts[:, 1]), maps[:, 1] - 1.0, maps[:, 2] - pts[:, 0]])


class FreqMapModel:
    """The network on the current GPU (aak_mlp_* of include/gpuAAK.h)."""

    def __init__(self, seed=0):
        self._m = C.c_void_p()
        _lib.check(_lib.lib().aak_mlp_create(int(seed), C.byref(self._m)), "mlp_create")

    def __del__(self):
        m = getattr(self, "_m", None)
        if m is not None and m.value:
            try:
                _lib.lib().aak_mlp_destroy(m)
            except Exception:
                pass
            self._m = None

    # ------------------------------------------------------------ parameters
    def get(self):
        p = np.zeros(N_PARAMS, dtype=np.float32)
        nm = np.zeros(N_NORM, dtype=np.float32)
        _lib.check(_lib.lib().aak_mlp_get_params(self._m, _lib.fptr(p), _lib.fptr(nm)), "mlp_get_params")
        return p, nm

    def set(self, params, norm=None):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != N_PARAMS:
            raise ValueError("expected %d parameters" % N_PARAMS)
        nm = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        _lib.check(_lib.lib().aak_mlp_set_params(self._m, _lib.fptr(p), _lib.fptr(nm) if nm is not None else None),
                   "mlp_set_params")

    def set_norm(self, norm):
        nm = np.ascontiguousarray(norm, dtype=np.float32)
        _lib.check(_lib.lib().aak_mlp_set_norm(self._m, _lib.fptr(nm)), "mlp_set_norm")

    def save(self, path=WEIGHTS, **meta):
        p, nm = self.get()
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=p, norm=nm, **meta)

    @classmethod
    def load(cls, path=WEIGHTS):
        with np.load(path, allow_pickle=False) as z:
            m = cls()
            m.set(z["params"], z["norm"])
        return m

    # ------------------------------------------------------------- compute
    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((len(X), 3), dtype=np.float32)
        _lib.check(_lib.lib().aak_mlp_forward(self._m, _lib.fptr(X), len(X), _lib.fptr(Y)), "mlp_forward")
        return Y

    def gradient(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(N_PARAMS, dtype=np.float32)
        loss = C.c_double(0)
        _lib.check(_lib.lib().aak_mlp_gradient(self._m, _lib.fptr(X), _lib.fptr(T), len(X), _lib.fptr(g),
                                               C.byref(loss)), "mlp_gradient")
        return g, loss.value

    def train(self, X, T, epochs=30, batch=4096, lr0=1e-3, lr1=1e-5, seed=4):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        hist = np.zeros(epochs)
        _lib.check(_lib.lib().aak_mlp_train(self._m, _lib.fptr(X), _lib.fptr(T), len(X), int(epochs), int(batch),
                                            float(lr0), float(lr1), int(seed), _lib.dptr(hist)), "mlp_train")
        return hist

    def predict_maps(self, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        out = np.zeros((len(pts), 3))
        _lib.check(_lib.lib().aak_mlp_predict_maps(self._m, _lib.dptr(pts), len(pts), _lib.dptr(out)),
                   "mlp_predict_maps")
        return out


def training_set(n=1_000_000, seed=4, tol=1e-13, threads=0, robust_z=10.0):
    """Inputs, targets and normalisation from the suite's own solver: n points
    of the workload, their maps to sum|f| <= tol at M = 1, failures dropped.

    Close to the last stable orbit the AK frequencies part from the Kerr ones
    and the corrections grow by orders of magnitude (about 5% of the workload's
    points lie beyond `robust_z` median absolute deviations of a correction);
    those points are left out of the training set -- a mean-squared loss
    would be dominated by them -- and their maps come from the solver
    fallback at inference."""
    pts = sample_points(n, seed, threads)
    maps, iters, resid, _ = solve_maps(pts, 0, None, tol, threads)
    ok = np.isfinite(resid) & (resid <= tol) & np.all(np.isfinite(maps), axis=1)
    pts, maps = pts[ok], maps[ok]
    c = corrections(pts, maps)
    med = np.median(c, 0)
    mad = 1.4826 * np.median(np.abs(c - med), 0)
    keep = np.all(np.abs(c - med) <= robust_z * mad, axis=1)
    pts, maps = pts[keep], maps[keep]
    f, c = features(pts), corrections(pts, maps)
    fm, fs = f.mean(0), f.std(0)
    cm, cs = c.mean(0), c.std(0)
    norm = np.concatenate([fm, fs, cm, cs]).astype(np.float32)
    X = ((f - fm) / fs).astype(np.float32)
    T = ((c - cm) / cs).astype(np.float32)
    return X, T, norm, pts, maps


def train(n=1_000_000, epochs=30, batch=4096, seed=4, out=WEIGHTS, threads=0, log=None):
    """Generate the maps, train on the GPU and save the weights."""
    X, T, norm, pts, maps = training_set(n, seed, threads=threads)
    if log:
        log("training set: %d maps" % len(X))
    model = FreqMapModel(seed)
    model.set_norm(norm)
    hist = model.train(X, T, epochs=epochs, batch=batch, seed=seed)
    if log:
        log("loss per epoch: " + " ".join("%.3e" % h for h in hist))
    if out:
        model.save(out, loss=hist, n_train=len(X), epochs=epochs, batch=batch, seed=seed)
    return model, hist


if __name__ == "__main__":
    import argparse
    import sys
    import time
    ap = argparse.ArgumentParser(description="train the frequency-map network on the GPU")
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--out", default=WEIGHTS)
    a = ap.parse_args()
    t0 = time.time()
    train(a.n, a.epochs, a.batch, a.seed, a.out, log=lambda s: print("%.1fs %s" % (time.time() - t0, s), flush=True))
    print("saved %s after %.1f s" % (a.out, time.time() - t0), flush=True)
    sys.exit(0)
