"""The frequency-map network on the GPU (emri_kludge_suite_amd/csrc/aak_mlp.hip):
its BF16 matrix-core forward pass and its back-propagated gradients against the
float64 reference (tests/mlp_ref.py, itself pinned by finite differences),
and a short training run that must learn the map."""
import numpy as np
import pytest

import mlp_ref as R
import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _model(params, norm=None):
    from emri_kludge_suite_amd.freqmap_mlp import FreqMapModel
    m = FreqMapModel(0)
    m.set(params.astype(np.float32), norm)
    return m


def _random_params(rng, scale=0.15):
    p = rng.normal(0, scale, R.NP)
    # He-like scale for the wide layers, so the activations stay O(1)
    for name, fan in (("W2", R.H), ("W3", R.H), ("W4", R.H)):
        o, n = R.OFF[name]
        p[o:o + n] = rng.normal(0, np.sqrt(2.0 / fan), n)
    return p.astype(np.float32).astype(np.float64)


def test_forward_matches_the_reference():
    rng = np.random.default_rng(71)
    p = _random_params(rng)
    X = rng.normal(size=(1000, 4)).astype(np.float32)  # not a multiple of the 64-row tile
    print("case mlp forward n=1000", flush=True)
    m = _model(p)
    Y = m.forward(X)
    ref, _ = R.forward(p, X.astype(np.float64), bf=True)
    err = np.abs(Y - ref) / (np.abs(ref) + 1e-2)
    print("forward: max rel err %.3e, median %.3e" % (err.max(), np.median(err)), flush=True)
    assert np.median(err) < 1e-5
    assert err.max() < 1e-2  # a BF16 rounding tie decided differently in FP32 and FP64 moves ~4e-3


def test_gradient_matches_the_reference():
    rng = np.random.default_rng(72)
    p = _random_params(rng)
    X = rng.normal(size=(777, 4)).astype(np.float32)
    T = rng.normal(size=(777, 3)).astype(np.float32)
    print("case mlp gradient n=777", flush=True)
    m = _model(p)
    g, loss = m.gradient(X, T)
    ref_loss, ref = R.loss_and_grad(p, X.astype(np.float64), T.astype(np.float64), bf=True)
    assert loss == pytest.approx(ref_loss, rel=1e-5)
    for name, (o, n) in R.OFF.items():
        a, b = g[o:o + n].astype(np.float64), ref[o:o + n]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print("  %s rel L2 %.3e" % (name, rel), flush=True)
        assert rel < 2e-3, name
    # and it is the gradient of the loss: against the unrounded float64 one
    _, exact = R.loss_and_grad(p, X.astype(np.float64), T.astype(np.float64), bf=False)
    assert np.linalg.norm(g - exact) / np.linalg.norm(exact) < 3e-2


def test_training_learns_the_map():
    from emri_kludge_suite_amd import freqmap_mlp as F
    print("case mlp training n=60000 epochs=12", flush=True)
    X, T, norm, pts, maps = F.training_set(60000, seed=4)
    m = F.FreqMapModel(4)
    m.set_norm(norm)
    _, loss0 = m.gradient(X[:8192], T[:8192])
    hist = m.train(X, T, epochs=12, batch=1024, lr0=1e-3, lr1=1e-5, seed=4)
    _, loss1 = m.gradient(X[:8192], T[:8192])
    print("loss before %.3e, per epoch %s, after %.3e" % (loss0, " ".join("%.2e" % h for h in hist), loss1),
          flush=True)
    assert np.all(np.isfinite(hist))
    assert loss1 < 0.1 * loss0
    # the learned starts polish to the solver's tolerance on new points
    ev = F.sample_points(4096, 5)
    maps_s, it_s, res_s, st_s = F.solve_maps(ev, 0, None, 1e-12)
    maps_m, it_m, res_m, st_m = F.solve_maps(ev, 1, m, 1e-12)
    ok = res_s <= 1e-12  # points the reference's solver itself resolves
    assert np.all(res_m[ok] <= 1e-12)
    print("solver: mean iterations %.2f; learned: mean Newton steps %.2f, fallbacks %d of %d" % (
        np.mean(it_s[ok]), np.mean(np.abs(it_m[ok])), st_m["fallbacks"], ok.sum()), flush=True)
    assert st_m["fallbacks"] < 0.15 * ok.sum()


def _sol_fun(x, Om, e, iota):
    import ctypes as C
    f = np.zeros(3)
    xx, oo = np.ascontiguousarray(x, dtype=np.float64), np.ascontiguousarray(Om, dtype=np.float64)
    pd = C.POINTER(C.c_double)
    ol.oracle().o_sol_fun(xx.ctypes.data_as(pd), oo.ctypes.data_as(pd), 1.0, float(e), float(iota),
                          f.ctypes.data_as(pd))
    return f


def _trained():
    import os
    from emri_kludge_suite_amd import freqmap_mlp as F
    if not os.path.exists(F.WEIGHTS):
        pytest.skip("no trained weights in the package")
    return F, F.FreqMapModel.load()


def test_trained_model_on_the_evaluation_points():
    # 4096 trajectory points of the evaluation draws (seed 5): every learned
    # start is polished to sum|f| <= 1e-12 (M = 1) or handed to the solver,
    # and the roots agree with the solver's
    F, m = _trained()
    ev = F.sample_points(4096, 5)
    maps_s, it_s, res_s, st_s = F.solve_maps(ev, 0, None, 1e-12)
    maps_m, it_m, res_m, st_m = F.solve_maps(ev, 1, m, 1e-12)
    ok = res_s <= 1e-12
    assert np.all(res_m[ok] <= 1e-12)
    rel = np.max(np.abs(maps_m[ok] - maps_s[ok]) / np.abs(maps_s[ok]))
    print("trained: fallbacks %d of %d, mean Newton steps %.2f vs solver %.2f iterations, solve ms %.1f vs %.1f, "
          "max rel diff of the roots %.2e" % (st_m["fallbacks"], ok.sum(), np.mean(np.abs(it_m[ok])),
                                              np.mean(it_s[ok]), st_m["solve_ms"], st_s["solve_ms"], rel), flush=True)
    assert st_m["fallbacks"] < 0.1 * ok.sum()
    # both are roots to sum|f| <= 1e-12; where the map is ill-conditioned
    # (near the separatrix, or where a direction leaves the three frequencies
    # unchanged) two such roots differ by up to cond(J) * 1e-12.
    # Same root: the step between them maps through the oracle's Jacobian
    # of f (central differences) to no more than the two residuals
    each = np.max(np.abs(maps_m - maps_s) / np.abs(maps_s), axis=1)
    far = np.nonzero(ok & (each > 1e-9))[0]
    print("roots differing by more than 1e-9: %d of %d" % (len(far), ok.sum()), flush=True)
    worst = 0.0
    for i in far:
        a, pp, e, iota = ev[i]
        Om = ol.o_iekg(pp, e, np.cos(iota), a, freq=True)[13:16]
        J = np.zeros((3, 3))
        for k in range(3):
            hk = 1e-6 * abs(maps_s[i, k])
            xp, xm = maps_s[i].copy(), maps_s[i].copy()
            xp[k] += hk
            xm[k] -= hk
            J[:, k] = (_sol_fun(xp, Om, e, iota) - _sol_fun(xm, Om, e, iota)) / (2 * hk)
        worst = max(worst, np.sum(np.abs(J @ (maps_m[i] - maps_s[i]))))
    print("largest |J dx|_1 over those: %.3e" % worst, flush=True)
    assert worst <= 1e-11


def test_learned_start_in_the_waveform():
    # the waveform path with the learned start.  The reference's solver stops
    # at sum|f| < 1e-6 (f ~ Omega M, so ~1e-10 relative at M = 1e6), the
    # polish at min(1e-6, 1e-12 M): the two maps differ within the solver's
    # own tolerance, and over ~1e6-1e7 radians of phase that is up to ~1e-4
    # in noise-weighted mismatch for the lightest sources; reported, held to 1e-4
    F, m = _trained()
    from emri_kludge_suite_amd import GPUAAK
    from emri_kludge_suite_amd.multigrid import psd_weights
    N, dt = 2 ** 20, 15.0
    w = psd_weights(N, dt, 1)
    rng = np.random.default_rng(12)
    g = GPUAAK(1.0, N // 100, N, 100.0, dt, True, False)
    g.set_nmax(30)
    worst, done = 0.0, 0
    while done < 6:
        pars = [0.5, rng.uniform(0, 0.95), rng.uniform(7, 15), rng.uniform(1e-3, 0.5), 10 ** rng.uniform(5, 7),
                rng.uniform(5, 50), 1.0, 2.0, 3.0, 1.1, 4.0, 0.7, 5.0, 1.0]
        try:
            g.set_map_init("solver")
            g.gpu_gen_AAK(*pars)
        except RuntimeError:
            continue
        _, a1, a2 = g.GetWaveform()
        g.set_map_init("learned", m)
        g.gpu_gen_AAK(*pars)
        _, b1, b2 = g.GetWaveform()
        fa = [np.fft.rfft(x[:N]) * w for x in (a1, a2)]
        fb = [np.fft.rfft(x[:N]) * w for x in (b1, b2)]
        ab = sum(np.vdot(x, y).real for x, y in zip(fa, fb))
        mm = 1 - ab / np.sqrt(sum(np.vdot(x, x).real for x in fa) * sum(np.vdot(y, y).real for y in fb))
        print("learned start: template %d mismatch vs solver start %.3e" % (done, mm), flush=True)
        worst = max(worst, mm)
        done += 1
    assert worst <= 1e-4
