"""Benchmark: AAK likelihoods per second on one MI355X.

Workload (production size): 1-yr-class LISA waveforms,
N = 2^21 samples at dt = 15 s (0.997 yr), harmonics n = 1..50, parameters drawn
from the production priors (log10 M ~ U[5,7], mu ~ U[5,50], a ~ U[0,0.95],
e0 ~ U[1e-3,0.5], p0 ~ U[7,15], iota0 = 0.5, angles uniform on the sphere),
noise-weighted inner products against a seeded injected signal under the
analytic LISA instrument + confusion PSD (1 yr) on the rFFT grid.

One step = one batch of templates (default 256, the "hundreds of waveforms per
launch" of the batched form) through the whole path: host
inspiral + knots (thread pool), device spline/phasing/exact folds/mode sum,
the fused FFT + PSD-weighted reduction, pipelined across templates.  The data
(weighted injection) is resident in HBM before timing starts.

The headline value is the mixed-precision form (FP32 harmonic loop on FP64
phases, expansions and sums), accepted only if its noise-weighted mismatch
against the FP64 waveforms of the same templates is <= 1e-6.  The plain FP64
batched path and the one-waveform-per-call FP64 path (the reference's calling
pattern) are timed on the same templates and reported beside it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "AAK likelihoods/sec, 1-yr dt=15s, mixed vs plain FP64 at mismatch<=1e-6"
N_SAMPLES = 2 ** 21
DT = 15.0
N_MAX = 50
MSUN_SEC = 6.67259e-11 * 1.98892e30 / 299792458.0 ** 3
FP64_PEAK_TFLOPS = 78.6  # MI355X published FP64 peak (vector and matrix)
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 peak


def draw_params(rng):
    """One draw from the production priors (14 parameters, reference order)."""
    return [0.5, rng.uniform(0, 0.95), rng.uniform(7, 15), rng.uniform(1e-3, 0.5),
            10 ** rng.uniform(5, 7), rng.uniform(5, 50), rng.uniform(0, 2 * np.pi),
            rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)),
            rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi), 1.0]


def valid_draws(seed, count):
    """First `count` draws whose host stage succeeds (p0 outside the LSO etc.),
    with the active length and harmonic count of each (for the FLOP count)."""
    from emri_kludge_suite_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(seed)
    out, meta = [], []
    cap = 4_000_000
    traj = np.zeros(500)
    kn = np.zeros(5 * cap)
    info = np.zeros(16, dtype=np.int32)
    dp = np.zeros(4)
    while len(out) < count:
        pars = draw_params(rng)
        pv = np.array(pars)
        K = L.aak_test_host_plan(_lib.dptr(pv), 1.0, N_SAMPLES, DT, 1, 0, N_MAX, _lib.dptr(traj),
                                 _lib.dptr(kn), cap, _lib.iptr(info), _lib.dptr(dp))
        if K > 0:
            out.append(pars)
            meta.append((min(N_SAMPLES, int(info[11]) + int(info[12])), int(info[13])))
    return out, meta


def modesum_flops(meta):
    """Algorithmic work of one mode-sum launch: N_active (190 n_modes + 460) FLOP,
    transcendental calls counted as 20 FLOP each."""
    return sum(n_act * (190.0 * nm + 460.0) for n_act, nm in meta)


def weights():
    from emri_kludge_suite_amd.LISAnoise import LISA_Noise, generate_noise_frequencies
    freqs = generate_noise_frequencies(N_SAMPLES * DT, 1 / DT)
    freqs[0] = freqs[1] / 10.0
    return 1.0 / np.sqrt(LISA_Noise(freqs, dur=1)) * np.sqrt(DT / N_SAMPLES)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Threads this process may run on (affinity, bounded by a cgroup quota)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            n = min(n, max(1, -(-int(q) // int(per))))
    except (OSError, ValueError):
        pass
    return max(1, n)


def cpu_baseline(templates, w, data, budget_s):
    """The CPU oracle (plain C restatement of the reference CPU AAK, -O3, no
    fast math) timed on the same workload: waveform + numpy rFFT + weighted
    sums per likelihood.  One core for half the budget, then one template
    per thread on every thread the process may use for the other half (the
    reference is single-threaded and independent across templates; the
    oracle's calls release the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol

    def one(pars):
        r = ol.o_aak_wave(pars, length=N_SAMPLES, dt=DT, LISA=True, backint=False, n_max=N_MAX)
        hI = np.fft.rfft(r["hI"]) * w
        hII = np.fft.rfft(r["hII"]) * w
        dh = 4 * np.sum(np.real(np.conj(hI) * data[0]) + np.real(np.conj(hII) * data[1]))
        hh = 4 * np.sum(np.abs(hI) ** 2 + np.abs(hII) ** 2)
        return dh, hh

    t0 = time.perf_counter()
    n1 = 0
    for pars in templates:
        one(pars)
        n1 += 1
        if time.perf_counter() - t0 > budget_s / 2:
            break
    single = n1 / (time.perf_counter() - t0)
    T = cpu_threads()
    t1 = time.perf_counter()
    nall = 0
    with ThreadPoolExecutor(T) as ex:
        while time.perf_counter() - t1 < budget_s / 2:
            batch = [templates[(nall + k) % len(templates)] for k in range(T)]
            list(ex.map(one, batch))
            nall += T
    allc = nall / (time.perf_counter() - t1)
    return {"value": allc, "unit": "likelihoods/s", "cores": T, "kind": "port",
            "single_core": {"value": single, "unit": "likelihoods/s", "cores": 1},
            "cpu_model": cpu_model(),
            "sample": "%d templates on one core, then %d on %d threads (one template per thread), "
                      "full N=2^21, n=1..50, %.0f s in all" % (n1, nall, T, budget_s)}


def copy_probe():
    """Plain copy of 2^27 doubles (1 GiB read + 1 GiB write), five times: the
    reference point that scales the FETCH_SIZE/WRITE_SIZE counters to bytes."""
    import torch
    x = torch.ones(2 ** 27, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_each_way": 2 ** 30, "launches": 5}))


def flop_probe():
    """In-place multiplies of 2^27 float32, then float64 values, five times each
    (one FLOP per element and launch): the reference point that scales the
    SQ_INSTS_VALU_FLOPS_FP32/FP64 counters to FLOP."""
    import torch
    for dt in (torch.float32, torch.float64):
        x = torch.ones(2 ** 27, dtype=dt, device="cuda")
        for _ in range(5):
            x.mul_(1.0000001)
    torch.cuda.synchronize()
    print(json.dumps({"flop_probe_elements": 2 ** 27, "launches_per_dtype": 5}))


def read_profile_value(name, key):
    """A figure from a committed profile summary (a line "key value"); None
    when the summary or the line is absent."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith(key + " "):
                return float(line.split()[-1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed batches")
    ap.add_argument("--warmup", type=int, default=3, help="untimed batches (also brings the GPU clocks up)")
    ap.add_argument("--batch", type=int, default=256, help="templates per step")
    ap.add_argument("--lanes", type=int, default=0,
                    help="concurrent streams of the batched path (0: the autotuner's table for this grid)")
    ap.add_argument("--single", type=int, default=8, help="one-waveform-per-call likelihoods timed beside")
    ap.add_argument("--mismatch", type=int, default=0,
                    help="templates whose mixed/FP64 mismatch is measured (0: every timed template)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=30.0)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned map-start leg")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-batch leg")
    ap.add_argument("--uneven-batch", type=int, default=1024)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (HBM counter calibration)")
    ap.add_argument("--flop-probe", action="store_true",
                    help="only run in-place float32/float64 multiplies of 2^27 values (FLOP counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.flop_probe:
        flop_probe()
        return
    if args.gpus != 1:
        print("bench.py: measures one GPU; --gpus %d is not supported "
              "(waveform batches shard across GPUs without communication)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from emri_kludge_suite_amd import GPUAAK

    from emri_kludge_suite_amd import autotune
    tuning = autotune.load()
    settings = {}

    injection = valid_draws(2, 1)[0][0]
    ntpl = max(args.batch, args.single)
    templates, meta = valid_draws(3, ntpl)
    print("bench: %d templates drawn" % len(templates), file=sys.stderr, flush=True)
    w = weights()
    g = GPUAAK(1.0, N_SAMPLES // 100, N_SAMPLES, 100.0, DT, True, False)
    g.set_nmax(N_MAX)
    g.gpu_gen_AAK(*injection)
    _, hI, hII = g.GetWaveform()
    dI = np.fft.rfft(hI[:N_SAMPLES]) * w
    dII = np.fft.rfft(hII[:N_SAMPLES]) * w
    g.input_data(dI, dII, w, w)
    P = np.array(templates[:args.batch])
    flops_batch = modesum_flops(meta[:args.batch])

    def configure(mode):
        """The pipeline settings of this grid and precision: the autotuner's
        table (data/tuning.json), --lanes overriding its lane count."""
        g.set_precision(mode)
        s = autotune.lookup(tuning, N_SAMPLES, DT, mode) or {}
        s = {k: s.get(k, autotune.DEFAULTS[k]) for k in autotune.DEFAULTS}
        if args.lanes:
            s["lanes"] = args.lanes
        g.set_lanes(s["lanes"])
        g.set_tuning(feed_depth=s["feed_depth"], exp_shift=-2 if s["exp_shift"] is None else s["exp_shift"])
        settings[mode] = s
        return s

    def profiled_pass(lanes):
        """One batch with HIP events around each stage on the stream of each
        template; returns the stage sums (ms) over the batch."""
        g.set_lanes(lanes)
        g.set_profiling(True)
        g.LikelihoodBatch(P)
        tm = g.batch_timings()
        g.set_profiling(False)
        return np.array([tm[x] for x in ("rates", "fold", "modesum", "fft", "dot")])

    host_plan = {}

    def run_batched(mode):
        lanes = configure(mode)["lanes"]
        for _ in range(args.warmup):
            g.LikelihoodBatch(P)
        t0 = time.perf_counter()
        host = np.zeros(2)
        busy, issue, wait = [], [], []
        for _ in range(args.steps):
            like = g.LikelihoodBatch(P)
            hs, hm = g.batch_host_ms()
            host += (hs, hm)
            ps = g.pool_stats()
            busy.append(ps["busy"])
            issue.append(ps["issue_ms"] / ps["wall_ms"])
            wait.append(ps["wait_ms"] / ps["wall_ms"])
        elapsed = time.perf_counter() - t0
        assert np.all(np.isfinite(like)), "bench: a template failed"
        host_plan[mode] = {"ms_per_template": host[0] / (args.steps * len(P)), "longest_ms": host[1] / args.steps,
                           "pool_threads": int(g.pool_stats()["threads"]), "pool_busy_frac": float(np.mean(busy)),
                           "issue_thread_enqueue_frac": float(np.mean(issue)),
                           "issue_thread_wait_frac": float(np.mean(wait))}
        # stage times: on one lane (no overlap with another template's kernels)
        # and as they run concurrently on the bench's lanes
        return elapsed, profiled_pass(1), profiled_pass(lanes), like

    # batched paths: host stages on the thread pool, device work pipelined
    el_mx, st_mx, st_mx_lanes, like_mx = run_batched("mixed")
    el_64, st_64, st_64_lanes, like_64 = run_batched("fp64")
    launches = args.steps * len(P)
    per_pass = len(P)
    value = launches / el_mx
    value_64 = launches / el_64

    # one waveform per call in FP64 (the reference's calling pattern)
    g.set_precision("fp64")
    g.set_profiling(True)
    single_stage = np.zeros(6)
    t1 = time.perf_counter()
    for k in range(args.single):
        g.WaveformThroughLikelihood(*templates[k % len(templates)])
        tm = g.timings()
        single_stage += np.array([tm[x] for x in ("host", "rates", "fold", "modesum", "fft", "dot")])
    single_elapsed = time.perf_counter() - t1
    g.set_profiling(False)
    value_single = args.single / single_elapsed
    # waveforms alone (host stage + dense stage + mode sum into HBM), FP64, one per call
    t1 = time.perf_counter()
    for k in range(args.single):
        g.gpu_gen_AAK(*templates[k % len(templates)])
    wave_elapsed = time.perf_counter() - t1

    # accuracy of the mixed form, outside the timed region, on every timed
    # template: noise-weighted mismatch against the FP64 waveform of the same
    # template.  The FP64 template's weighted spectrum (left by Likelihood)
    # becomes the data, so the mixed template's likelihood gives
    # 4<h_mixed|h_fp64> and 4<h_mixed|h_mixed> on the device.
    mism = []
    nb = N_SAMPLES // 2 + 1
    for k in range(min(args.mismatch, len(P)) if args.mismatch > 0 else len(P)):
        g.set_precision("fp64")
        g.gpu_gen_AAK(*P[k])
        hh_64 = g.Likelihood()[1]
        _, sI, sII = g.GetWaveform()
        g.input_data(sI[:2 * nb].view(np.complex128), sII[:2 * nb].view(np.complex128), w, w)
        g.set_precision("mixed")
        g.gpu_gen_AAK(*P[k])
        cross, hh_mx = g.Likelihood()
        mism.append(1.0 - cross / np.sqrt(hh_mx * hh_64))
        g.input_data(dI, dII, w, w)
    mism = np.array(mism)
    like_rel = float(np.max(np.abs(like_mx[:, 1] - like_64[:, 1]) / np.abs(like_64[:, 1])))
    n_bad = int(np.count_nonzero(~(mism <= 1e-6)))
    assert n_bad == 0, "bench: %d templates with mixed-precision mismatch above 1e-6 (worst %.3e)" % (
        n_bad, float(np.max(mism)))

    def roof(stage):
        """Mean mode-sum stage time per launch (one lane) and the convention's TFLOP/s."""
        return stage[2] / per_pass, flops_batch / (stage[2] * 1e-3) / 1e12

    ms_mx, ach_mx = roof(st_mx)
    ms_64, ach_64 = roof(st_64)
    HBM_PROF, SQ_PROF, FLOP_PROF = ("aak_likelihood_N2p21_n50_hbm_counters.md", "aak_likelihood_N2p21_n50_sq_counters.md",
                                    "aak_likelihood_N2p21_n50_flops.md")
    traffic = read_profile_value(HBM_PROF, "traffic_bytes_per_launch modesum_stage_mixed")
    valu_frac = read_profile_value(SQ_PROF, "valu_issue_frac modesum_stage_mixed")
    # executed work: FP32 + FP64 VALU FLOP per launch of the mixed stage from
    # the SQ_INSTS_VALU_FLOPS_* counters (calibrated, in the committed
    # summary), over the stage time measured here
    exec_flop = read_profile_value(FLOP_PROF, "valu_flop_per_launch modesum_stage_mixed")
    exec_tf = exec_flop / (ms_mx * 1e-3) / 1e12 if exec_flop else None
    # HBM rooflines of the fold stage and k_rates, algorithmic bytes per
    # active sample: the fold reads a rate and writes a phase per step and
    # phase (3 x 16 B); k_rates writes e, three rates and the amplitude (40 B)
    n_active = sum(n for n, _ in meta[:len(P)]) / len(P)

    def hbm_roof(nbytes, ms, key):
        gbs = nbytes / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": gbs, "peak": 8000.0, "unit": "GB/s", "frac": gbs / 8000.0,
                "bytes_per_launch": nbytes, "ms": ms, "traffic": read_profile_value(HBM_PROF, key)}
    # noise-weighted inner product (fused four-step FFT + PSD weighting + both
    # sums), HBM-bound; algorithmic bytes per channel: 8N (time series) +
    # 24(N/2+1) (weighted data and weights), two channels.  The four-step
    # scratch adds 32N bytes in all (written once, read once).
    ip_bytes = 2 * (8 * N_SAMPLES + 24 * (N_SAMPLES // 2 + 1))
    ip_ms = (st_mx[3] + st_mx[4]) / per_pass
    stages = ("rates", "fold", "modesum", "fft", "dot")
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "likelihoods/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el_mx / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "mixed: f64 phases/expansions/sums, f32 harmonic loop",
        "data": "synthetic (seeded prior draws, injection seed 2, templates seed 3)",
        "config": {"workload": "1-yr LISA AAK likelihoods, N=2^21, dt=15 s, n=1..50, %d templates per step "
                               "(batched), mixed precision" % len(P),
                   "N": N_SAMPLES, "dt": DT, "n_max": N_MAX, "batch": len(P), "pipeline": settings,
                   "psd": "LISA_Noise(dur=1)"},
        "host_stage": {"mixed": host_plan.get("mixed"), "fp64": host_plan.get("fp64"),
                       "note": "host-stage time per template on the pool during the timed batches"},
        "accuracy": {"max_mismatch_vs_fp64": float(np.max(mism)), "median_mismatch_vs_fp64": float(np.median(mism)),
                     "templates_checked": int(mism.size), "templates_above_1e-6": n_bad,
                     "max_rel_diff_hh_vs_fp64": like_rel},
        "stages_ms_per_template": dict(zip(stages, [round(x / per_pass, 4) for x in st_mx])),
        "stages_ms_per_template_overlapped": dict(zip(stages, [round(x / per_pass, 4) for x in st_mx_lanes])),
        "plain_fp64_batched": {"value": value_64, "unit": "likelihoods/s", "ms_per_step": 1e3 * el_64 / args.steps,
                               "stages_ms_per_template": dict(zip(stages, [round(x / per_pass, 4) for x in st_64])),
                               "k_modesum_algorithmic_tflops": ach_64, "k_modesum_ms": ms_64},
        "single_call_fp64": {"value": value_single, "unit": "likelihoods/s",
                             "ms_per_likelihood": 1e3 * single_elapsed / args.single,
                             "stages_ms": dict(zip(("host",) + stages,
                                                   [round(x / args.single, 4) for x in single_stage]))},
        "single_waveform_fp64": {"value": args.single / wave_elapsed, "unit": "waveforms/s",
                                 "ms_per_waveform": 1e3 * wave_elapsed / args.single},
        "speedup": {"mixed_batched_over_single_fp64": value / value_single,
                    "fp64_batched_over_single_fp64": value_64 / value_single},
        "roofline": {"bound": "valu", "achieved": exec_tf, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": exec_tf / FP32_PEAK_TFLOPS if exec_tf else None, "traffic": traffic,
                     "kernel": "mode-sum stage, mixed precision (k_modesum_setup + k_modesum_split + k_modesum + k_modesum_direct)",
                     "kernel_ms": ms_mx,
                     "kernel_timing": "HIP events around the stage on its stream; one profiled pass over the batch on one lane",
                     "flop_per_launch": exec_flop,
                     "convention": "executed FP32 + FP64 VALU FLOP per launch (SQ_INSTS_VALU_FLOPS_FP32/FP64, calibrated; "
                                   "profiles/%s) over the measured stage time; peak = published FP32 vector rate "
                                   "(packed FP32 is not emitted in this build, so its reachable peak is half)" % FLOP_PROF,
                     "valu_issue_frac": valu_frac,
                     "algorithmic_tflops": ach_mx,
                     "algorithmic_convention": "N_active*(190*n_modes+460) FLOP per launch: the reference's per-sample "
                                               "work (transcendentals at 20 FLOP), most of which the Taylor form skips; "
                                               "an effective rate, not a fraction of any peak"},
        "fold": dict(hbm_roof(48.0 * n_active, st_mx[1] / per_pass, "traffic_bytes_per_launch fold_stage"),
                     kernels="k_guard + k_fold_resolve (the replay check rides on the mode-sum setup)",
                     bytes_convention="48 B per active sample (three phases: rate read, phase written)"),
        "rates": dict(hbm_roof(40.0 * n_active, st_mx[0] / per_pass, "traffic_bytes_per_launch k_rates"),
                      kernels="k_rates", bytes_convention="40 B per active sample (e, three rates, amplitude written)"),
        "inner_product": {"bound": "hbm", "achieved": ip_bytes / (ip_ms * 1e-3) / 1e9, "peak": 8000.0, "unit": "GB/s",
                          "frac": ip_bytes / (ip_ms * 1e-3) / 8e12, "bytes_per_launch": ip_bytes, "ms": ip_ms,
                          "kernels": "k_fft_cols (column FFTs x twiddles) + k_fft_rows (row FFTs, real-spectrum split, "
                                     "weighting, both sums; its last workgroup adds the partial sums)",
                          "scratch_bytes_per_launch": 32 * N_SAMPLES},
    }
    del g
    model = learned_model()
    if not args.no_learned:
        result["learned_map_init"] = learned_map_leg(model)
    if not args.no_uneven:
        result["uneven_batch"] = uneven_leg(model, args.uneven_batch)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(templates, w, (dI, dII), args.cpu_seconds)
    print(json.dumps(result))


def learned_model():
    """The frequency-map network trained on the GPU (package data), or None."""
    from emri_kludge_suite_amd import freqmap_mlp as F
    if not os.path.exists(F.WEIGHTS):
        return None
    return F.FreqMapModel.load()


def learned_map_leg(model, n=4096):
    """Frequency maps of 4096 trajectory points (the learned form's workload,
    evaluation seed 5) at M = 1 to sum|f| <= 1e-12: the reference's hybrid
    solver against the network's start + Newton polish (solver fallback)."""
    from emri_kludge_suite_amd import freqmap_mlp as F
    if model is None:
        return {"skipped": "no trained weights"}
    pts = F.sample_points(n, 5)
    out = {}
    for mode, name in ((0, "solver"), (1, "learned")):
        F.solve_maps(pts, mode, model, 1e-12)  # warm
        t0 = time.perf_counter()
        maps, it, res, st = F.solve_maps(pts, mode, model, 1e-12)
        wall = time.perf_counter() - t0
        ok = np.isfinite(res)
        out[name] = {"maps_per_s": n / wall, "wall_ms": 1e3 * wall,
                     "host_ms": {"geodesic": st["geodesic_ms"], "model": st["model_ms"], "solve": st["solve_ms"]},
                     "mean_iterations": float(np.mean(np.abs(it[ok]))), "max_residual": float(np.max(res[ok])),
                     "fallbacks": st["fallbacks"], "failed": st["failed"]}
        if mode == 0:
            ref = maps
    d = np.abs(maps - ref) / np.abs(ref)
    out["max_rel_diff_vs_solver"] = float(np.nanmax(d))
    out["solve_ms_saved_per_map"] = (out["solver"]["host_ms"]["solve"] - out["learned"]["host_ms"]["solve"]) / n
    return out


def uneven_leg(model, n):
    """The uneven batch: n templates with durations U[0.25, 2] yr, dt in
    {10, 15} s, e0 U[1e-3, 0.6], n_max = clamp(round(100 e0), 10, 60), each
    scored against an injection plus Gaussian noise of its grid, mixed
    precision and (with trained weights) the learned map start; grids are
    grouped, each with its FFT length and tuned pipeline."""
    from emri_kludge_suite_amd import multigrid as MG
    P, N, dt, nmax, T = MG.config5_workload(n, 6)
    inj = MG.injection_source(7)
    mg = MG.MultiGridBatch(precision="mixed", map_model=model)
    for k in sorted(set(zip(N.tolist(), dt.tolist()))):
        MG.injected_data(mg, inj, k[0], k[1], seed=7)
    mg.likelihood(P, N, dt, nmax)  # untimed: growth of buffers and pools
    t0 = time.perf_counter()
    like, st = mg.likelihood(P, N, dt, nmax)
    wall = time.perf_counter() - t0
    groups = {"%d@%gs" % k: int(np.count_nonzero((N == k[0]) & (dt == k[1]))) for k in mg.groups()}
    return {"value": int(np.count_nonzero(st == 0)) / wall, "unit": "likelihoods/s", "templates": n,
            "failed_templates": int(np.count_nonzero(st != 0)), "wall_s": wall, "groups": groups,
            "learned_map_init": model is not None, "precision": "mixed"}


if __name__ == "__main__":
    main()
