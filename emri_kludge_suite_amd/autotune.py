"""Autotuner of the batched likelihood.

For one grid (N, dt) and precision it times short batches of real templates
over the pipeline's knobs and keeps the fastest setting:
  lanes       concurrent streams, each with its own device buffers (1..8);
  feed_depth  short host plans kept ready for the GPU (the pool's ordering;
              kept at its default);
  exp_shift   mode-sum expansion blocks of 256 << shift samples (the Taylor
              order adapts, so accuracy does not depend on it).
The search is one coordinate-descent sweep from the defaults (three timed
batches per candidate after one untimed one, at the batch size the grid
serves); a candidate replaces the incumbent only when it is at least 5%
faster, so timing noise does not move the table.  Results are kept in a JSON
table keyed by grid and precision; `apply` looks a grid up (the nearest
length of the same dt and precision when it is absent).
"""
import json
import os
import time

import numpy as np

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tuning.json")
DEFAULTS = {"lanes": 2, "feed_depth": 10, "exp_shift": None}
# the feed depth is not searched: batch-length timings cannot tell its effect
# from the order in which a batch's long plans start (10 is measured best on
# the 1-yr grid: 6 and 16 lose 5% and 23%)
CANDIDATES = {"lanes": (1, 2, 3, 4, 6, 8), "exp_shift": (0, 1, 2, 3)}


def key(N, dt, precision):
    return "%d|%g|%s" % (int(N), float(dt), precision)


def _set(g, s):
    g.set_lanes(s["lanes"])
    g.set_tuning(feed_depth=s["feed_depth"], exp_shift=-2 if s["exp_shift"] is None else s["exp_shift"])


def _time(g, P, nmax, repeats):
    g.LikelihoodBatch(P, nmax=nmax)
    t0 = time.perf_counter()
    for _ in range(repeats):
        g.LikelihoodBatch(P, nmax=nmax)
    return (time.perf_counter() - t0) / (repeats * len(P))


def tune(g, templates, nmax=None, repeats=3, log=None, margin=0.95):
    """Best setting for handle g (already holding data) on `templates`;
    returns (setting, {candidate: seconds per likelihood})."""
    P = np.ascontiguousarray(templates, dtype=np.float64)
    best = dict(DEFAULTS)
    _set(g, best)
    best_t = _time(g, P, nmax, repeats)
    seen = {json.dumps(best, sort_keys=True): best_t}
    for knob, values in CANDIDATES.items():
        for v in values:
            s = dict(best)
            s[knob] = v
            k = json.dumps(s, sort_keys=True)
            if k not in seen:
                _set(g, s)
                seen[k] = _time(g, P, nmax, repeats)
                if log:
                    log("  %s -> %.3f ms per likelihood" % (k, 1e3 * seen[k]))
            if seen[k] < margin * best_t:
                best, best_t = s, seen[k]
    _set(g, best)
    return best, seen


def load(path=TABLE):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save(table, path=TABLE):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def lookup(table, N, dt, precision):
    """The entry of (N, dt, precision), else the one of the nearest length with
    the same dt and precision, else None."""
    k = key(N, dt, precision)
    if k in table:
        return table[k]
    best, dist = None, None
    for kk, v in table.items():
        n2, dt2, p2 = kk.split("|")
        if float(dt2) == float(dt) and p2 == precision:
            d = abs(np.log2(int(n2)) - np.log2(N))
            if dist is None or d < dist:
                best, dist = v, d
    return best


def apply(g, N, dt, precision, table=None):
    """Sets handle g to the tabulated setting of its grid; returns it (None
    when the table has nothing for this precision and dt)."""
    s = lookup(load() if table is None else table, N, dt, precision)
    if s:
        _set(g, {k: s.get(k, DEFAULTS[k]) for k in DEFAULTS})
    return s


def _grid_templates(N, dt, n, seed, uneven):
    """n templates whose host stage succeeds on grid (N, dt): the uneven
    batch's draws, or the production prior (n_max = 50) for the 1-yr grid."""
    from . import _lib
    from .multigrid import config5_workload
    L = _lib.lib()
    out, nm = [], []
    traj, kn = np.zeros(500), np.zeros(5 * 4_000_000)
    info, dp = np.zeros(16, dtype=np.int32), np.zeros(4)
    k = 0
    while len(out) < n:
        P, _, _, nmax, _ = config5_workload(256, seed + k)
        k += 1
        for row, m in zip(P, nmax):
            m = int(m) if uneven else 50
            rc = L.aak_test_host_plan(_lib.dptr(np.ascontiguousarray(row)), 1.0, int(N), float(dt), 1, 0, m,
                                      _lib.dptr(traj), _lib.dptr(kn), 4_000_000, _lib.iptr(info), _lib.dptr(dp))
            if rc > 0:
                out.append(row)
                nm.append(m)
            if len(out) == n:
                break
    return np.array(out), np.array(nm, dtype=np.int32)


if __name__ == "__main__":
    import argparse
    from .multigrid import MultiGridBatch, config5_workload, injected_data, injection_source
    ap = argparse.ArgumentParser(description="tune the batch pipeline on the GPU and write the table")
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--templates", type=int, default=256, help="batch size of the 1-yr grid (2^21 at 15 s)")
    a = ap.parse_args()
    table = load(a.out)
    _, N, dt, _, _ = config5_workload(1024, 6)
    # each grid of the uneven batch is tuned at the size of its group there
    count = {}
    for x, y in zip(N, dt):
        count[(int(x), float(y))] = count.get((int(x), float(y)), 0) + 1
    jobs = [(2 ** 21, 15.0, "mixed", False, a.templates), (2 ** 21, 15.0, "fp64", False, a.templates)] + \
        [(n, d, "mixed", True, max(32, c)) for (n, d), c in sorted(count.items())]
    inj = injection_source(7)
    done = set()
    for n, d, prec, uneven, size in jobs:
        if key(n, d, prec) in done:
            continue  # the 1-yr grid's own entry serves its uneven group as well
        done.add(key(n, d, prec))
        mg = MultiGridBatch(precision=prec, tuning={})
        injected_data(mg, inj, n, d, seed=7)
        g = mg.handle(n, d)
        P, nm = _grid_templates(n, d, size, 11, uneven)
        if not uneven:
            g.set_nmax(50)
            nm = None
        t0 = time.perf_counter()
        best, seen = tune(g, P, nm, log=print)
        table[key(n, d, prec)] = dict(best, ms_per_likelihood=1e3 * seen[json.dumps(best, sort_keys=True)],
                                      default_ms=1e3 * seen[json.dumps(DEFAULTS, sort_keys=True)])
        print("%s: %s (%.1f s)" % (key(n, d, prec), table[key(n, d, prec)], time.perf_counter() - t0), flush=True)
        save(table, a.out)
        del mg, g
