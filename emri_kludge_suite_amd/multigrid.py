"""Batched likelihoods over templates of uneven duration and sampling step.

The reference scores templates one at a time on one grid (pygpuAAK.py:204-232
loops over GPUAAK.WaveformThroughLikelihood, AAK_manager.cu:154-260,352-452).
Here one call takes a batch whose rows carry their own grid (length N, step
dt) and harmonic count: rows are grouped by grid, each grid has its own
handle -- device buffers, one FFT plan of its length (no padding to a common
length) and its own data -- tuned by the autotuner's table, and every group
runs through the pipelined batch path.  Groups run three at a time (each on
its own handle, host pool and streams), so one group's start and drain
overlap the others' steady state (1024 templates over seven grids, solver
start: 2.3e3-2.7e3/s one group at a time, 2.8e3-3.2e3 two, 3.1e3-3.4e3
three).  Each row equals a single-template call on a handle
of its grid.

Lengths are powers of two, which the fused FFT needs: a duration T at step dt
is served on N = 2^floor(log2(T / dt)) samples (the grid's length is within a
factor two below T).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import autotune
from .gpuAAK import GPUAAK

YEAR = 31558149.8  # Constants.h


def grid_length(T, dt):
    """Power-of-two sample count for a duration T (s) at step dt."""
    return 1 << int(np.floor(np.log2(T / dt)))


class MultiGridBatch:
    def __init__(self, T_fit=1.0, LISA=True, backint=False, precision="mixed", map_model=None, tuning=None,
                 concurrent_groups=3):
        self.T_fit, self.LISA, self.backint = T_fit, LISA, backint
        self.concurrent_groups = max(1, int(concurrent_groups))
        self.precision = precision
        self.map_model = map_model
        self.tuning = autotune.load() if tuning is None else tuning
        self.handles = {}

    def handle(self, N, dt):
        """The grid's handle, created on first use with its tuning applied."""
        k = (int(N), float(dt))
        if k not in self.handles:
            g = GPUAAK(self.T_fit, int(N) // 100, int(N), 100.0, float(dt), self.LISA, self.backint)
            g.set_precision(self.precision)
            if self.map_model is not None:
                g.set_map_init("learned", self.map_model)
            autotune.apply(g, N, dt, self.precision, self.tuning)
            self.handles[k] = g
        return self.handles[k]

    def input_data(self, N, dt, hI_f, hII_f, w1, w2):
        """Weighted data and weights of one grid (as GPUAAK.input_data)."""
        self.handle(N, dt).input_data(hI_f, hII_f, w1, w2)

    def likelihood(self, params, N, dt, nmax=None):
        """Rows [4<h|d>, 4<h|h>] of the templates params[i] on grid
        (N[i], dt[i]) with nmax[i] harmonics (0: the reference's rule);
        returns (like, status) in the order of the rows."""
        P = np.ascontiguousarray(params, dtype=np.float64)
        n = len(P)
        N = np.broadcast_to(np.asarray(N, dtype=np.int64), (n,))
        dt = np.broadcast_to(np.asarray(dt, dtype=np.float64), (n,))
        nm = np.zeros(n, dtype=np.int32) if nmax is None else np.asarray(nmax, dtype=np.int32)
        like = np.full((n, 2), np.nan)
        status = np.zeros(n, dtype=np.int32)
        # longest grids first: their host stages are the slowest
        keys = sorted({(int(a), float(b)) for a, b in zip(N, dt)}, key=lambda x: (-x[0], x[1]))
        for k in keys:
            self.handle(*k)  # created here, one at a time

        def run(k):
            rows = np.nonzero((N == k[0]) & (dt == k[1]))[0]
            out, st = self.handles[k].LikelihoodBatch(P[rows], return_status=True, nmax=nm[rows])
            return rows, out, st

        if self.concurrent_groups == 1 or len(keys) == 1:
            results = [run(k) for k in keys]
        else:
            with ThreadPoolExecutor(max_workers=self.concurrent_groups) as ex:
                results = list(ex.map(run, keys))
        for rows, out, st in results:
            like[rows] = out
            status[rows] = st
        return like, status

    def groups(self):
        return sorted(self.handles)


def config5_workload(n=1024, seed=6):
    """The uneven batch: production-prior sources with durations
    T ~ U[0.25, 2] yr, dt in {10, 15} s, e0 ~ U[1e-3, 0.6] and
    n_max = clamp(round(100 e0), 10, 60); returns (params, N, dt, nmax, T)."""
    rng = np.random.default_rng(seed)
    T = rng.uniform(0.25, 2.0, n) * YEAR
    dt = rng.choice([10.0, 15.0], n)
    N = np.array([grid_length(a, b) for a, b in zip(T, dt)])
    P = np.zeros((n, 14))
    P[:, 0] = 0.5
    P[:, 1] = rng.uniform(0, 0.95, n)
    P[:, 2] = rng.uniform(7, 15, n)
    P[:, 3] = rng.uniform(1e-3, 0.6, n)
    P[:, 4] = 10 ** rng.uniform(5, 7, n)
    P[:, 5] = rng.uniform(5, 50, n)
    P[:, 6:9] = rng.uniform(0, 2 * np.pi, (n, 3))
    P[:, 9] = np.arccos(rng.uniform(-1, 1, n))
    P[:, 10] = rng.uniform(0, 2 * np.pi, n)
    P[:, 11] = np.arccos(rng.uniform(-1, 1, n))
    P[:, 12] = rng.uniform(0, 2 * np.pi, n)
    P[:, 13] = 1.0
    nmax = np.clip(np.rint(100 * P[:, 3]), 10, 60).astype(np.int32)
    return P, N, dt, nmax, T


def injection_source(seed=7, N=1 << 22, dt=10.0):
    """The injected source: the first draw of the uneven batch's prior (seed)
    whose host stage succeeds on the longest grid (2^22 samples at 10 s)."""
    from . import _lib
    L = _lib.lib()
    traj, kn = np.zeros(500), np.zeros(5 * 4_000_000)
    info, dp = np.zeros(16, dtype=np.int32), np.zeros(4)
    for k in range(64):
        P = config5_workload(16, seed + 1000 * k)[0]
        for row in P:
            if L.aak_test_host_plan(_lib.dptr(np.ascontiguousarray(row)), 1.0, N, dt, 1, 0, 0, _lib.dptr(traj),
                                    _lib.dptr(kn), 4_000_000, _lib.iptr(info), _lib.dptr(dp)) > 0:
                return row
    raise RuntimeError("no injection source found")


def psd_weights(N, dt, dur):
    """w_k = sqrt(dt / N) / sqrt(S_n(f_k)) on the rFFT grid, f_0 = f_1 / 10
    (pygpuAAK.py:72,86-96)."""
    from .LISAnoise import LISA_Noise
    f = np.fft.rfftfreq(N, dt)
    f[0] = f[1] / 10.0
    return np.sqrt(dt / N) / np.sqrt(LISA_Noise(f, dur=dur))


def injected_data(mg, injection, N, dt, seed=7):
    """Data of grid (N, dt): the injection's waveform plus Gaussian noise of
    the analytic LISA PSD (confusion of the nearest fitted duration), in
    unnormalised rFFT units (Re, Im ~ N(0, sqrt(N S_n / (4 dt))) -- the
    reference's generate_noise_single_channel, LISAnoise.py:56-63, is dt times
    too large in these units), weighted; loaded into the grid's handle."""
    from .LISAnoise import LISA_Noise
    T = N * dt / YEAR
    dur = min((0.5, 1, 2, 4), key=lambda d: abs(np.log(d / T)))
    w = psd_weights(N, dt, dur)
    f = np.fft.rfftfreq(N, dt)
    f[0] = f[1] / 10.0
    sig = np.sqrt(N * LISA_Noise(f, dur=dur) / (4 * dt))
    rng = np.random.default_rng([seed, N, int(dt)])
    g = mg.handle(N, dt)
    g.gpu_gen_AAK(*injection)
    _, hI, hII = g.GetWaveform()
    d = []
    for h in (hI[:N], hII[:N]):
        noise = sig * (rng.normal(size=f.size) + 1j * rng.normal(size=f.size))
        d.append((np.fft.rfft(h) + noise) * w)
    mg.input_data(N, dt, d[0], d[1], w, w)
    return d, w
