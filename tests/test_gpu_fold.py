"""GPU: the two forms of the exact phase fold.  The single-pass form (one
kernel; each block finds the exact value before it by a look-back over the
blocks before it) is the default; the three-pass form (classification,
one-workgroup chain, write) is kept as a test hook.  Both must reproduce the
sequential FP64 rounding of AAK.cc:244-264, so their dense phases are equal
bit for bit on every draw; the oracle comparisons of the other parity tests
run the default form.  A block that fails its own check hands the phase to
the serial replay, after every other block of the phase has finished."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

PHASES = ("gim", "Phi", "alp")


def _draws(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        out.append([0.5, rng.uniform(0, 0.95), rng.uniform(7, 15), rng.uniform(1e-3, 0.5),
                    10 ** rng.uniform(5, 7), rng.uniform(5, 50), rng.uniform(0, 2 * np.pi),
                    rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)),
                    rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi), 1.0])
    return out


def _dense(g, pars):
    g.gpu_gen_AAK(*pars)
    info = g.info()
    d = g.dense_arrays()
    return info, {k: d[k].copy() for k in PHASES}


@pytest.mark.parametrize("backint", [False, True])
def test_single_pass_and_three_pass_folds_agree(backint):
    from emri_kludge_suite_amd import GPUAAK
    N, dt = 2 ** 21, 15.0
    g = GPUAAK(1.0, N // 10, N, 100.0, dt, True, backint)
    g.set_nmax(4)
    for i, pars in enumerate(_draws(20 + int(backint), 8)):
        print("case fold-forms backint=%d draw %d pars=%s" % (backint, i, np.array2string(np.array(pars), precision=6)),
              flush=True)
        g.fold_mode(1)
        info1, ph1 = _dense(g, pars)
        g.fold_mode(0)
        info0, ph0 = _dense(g, pars)
        assert info0["fold_fallbacks"] == 0 and info1["fold_fallbacks"] == 0
        for k in PHASES:
            assert np.array_equal(ph0[k], ph1[k]), (i, k)


def test_failed_block_hands_the_phase_to_the_replay():
    # the block in the middle of every phase fails its walk; its successors
    # see the failure through the look-back and the last block of the phase
    # replays it serially
    from emri_kludge_suite_amd import GPUAAK
    pars = [np.pi / 6, 0.5, 8.0, 0.1, 1e6, 100.0, 1.0, 2.0, 3.0, 1.1, 4.0, 0.7, 5.0, 1.0]
    N, dt = 2 ** 18, 10.0
    print("case fold-mid-failure N=%d backint=1" % N, flush=True)
    ref = ol.o_aak_wave(pars, length=N, dt=dt, backint=True, want_dense=True)
    g = GPUAAK(1.0, N // 10, N, 100.0, dt, True, True)
    g.fold_mode(0)
    g.force_fold_replay(2)
    g.gpu_gen_AAK(*pars)
    info = g.info()
    dense = g.dense_arrays()
    print(info, flush=True)
    assert info["fold_fallbacks"] == (6 if info["i0"] > 0 else 3)
    for name in PHASES:
        assert np.array_equal(dense[name], ref["dense_" + name]), name
    g.force_fold_replay(0)
    g.gpu_gen_AAK(*pars)
    assert g.info()["fold_fallbacks"] == 0
    dense = g.dense_arrays()
    for name in PHASES:
        assert np.array_equal(dense[name], ref["dense_" + name]), name
