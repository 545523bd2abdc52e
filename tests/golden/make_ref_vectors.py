"""Writes tests/golden/ref_subset.npz: input/output vectors of the reference's
own GSL-free trajectory code (GKTrajFast, IEKG, GKR, CKG and the
Numerical-Recipes numerics), compiled from the reference sources by
`make -C oracle ref` into oracle/_ref/libaakref.so.

The vectors pin the CPU restatement in oracle/ bit for bit where the compiled
reference is not present (it is built only where the reference sources are).
Run from the repository root:  python tests/golden/make_ref_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as ol  # noqa: E402

MSUN_SEC = 6.67259e-11 * 1.98892e30 / 299792458.0 ** 3


def dt_fit(N, dt, M, mu, T_fit=1.0):
    """Trajectory step of the sparse stage (AAKpy.cc:20-26)."""
    tr = N * dt / MSUN_SEC / M / M * mu
    return min(tr, T_fit) / 99


def main():
    if not ol.ref_available():
        raise SystemExit("oracle/_ref/libaakref.so is missing: run `make -C oracle ref` first")
    R = ol.ref()
    rng = np.random.default_rng(20240611)
    out = {}

    # geodesic constants and frequencies over the production range
    pts = []
    while len(pts) < 48:
        a = rng.uniform(0.01, 0.95)
        cosi = rng.uniform(-0.99, 0.99)
        e = rng.uniform(1e-3, 0.7)
        p = rng.uniform(6.0, 20.0)
        # the reference stops the process on invalid Carlson arguments, so only
        # points that the restatement finds stable are handed to it
        try:
            g = ol.o_iekg(p, e, cosi, a)
        except RuntimeError:
            continue
        if np.all(np.isfinite(g)) and g[3] > 0:
            pts.append((p, e, cosi, a))
    pts += [(8.0, 0.2, np.cos(0.5), 0.5), (6.5, 0.6, 0.9, 0.9), (12.0, 1e-3, -0.7, 0.3)]
    pts = np.array(pts)
    out["iekg_in"] = pts
    out["iekg_out"] = np.array([ol.r_iekg(*q) for q in pts])
    pd = np.zeros((len(pts), 3))
    for i, q in enumerate(pts):
        R.ref_param_derivs(*q, ol._ptr(pd[i]))
    out["derivs_out"] = pd

    # LSO table and circular orbits
    lso_in = np.array([(c, a) for c in (1.0, 0.6, -0.2, -1.0) for a in (0.0, 0.5, 0.95)])
    lso = np.zeros((len(lso_in), 4))
    for i, (c, a) in enumerate(lso_in):
        R.ref_plso_table(c, a, ol._ptr(lso[i]))
    out["plso_in"], out["plso_out"] = lso_in, lso
    ckg_in = np.array([(rng.uniform(-0.95, 0.95), rng.uniform(6.0, 20.0), rng.uniform(0.0, 0.95))
                       for _ in range(16)])
    ckg = np.zeros((len(ckg_in), 4))
    for i, q in enumerate(ckg_in):
        R.ref_ckg(*q, ol._ptr(ckg[i]))
    out["ckg_in"], out["ckg_out"] = ckg_in, ckg

    # Carlson integrals and Legendre forms
    xyz = rng.uniform(0.01, 5.0, size=(32, 4))
    out["carlson_in"] = xyz
    out["carlson_out"] = np.array([[R.ref_rf(x, y, z), R.ref_rd(x, y, z), R.ref_rj(x, y, z, p), R.ref_rc(x, y)]
                                   for x, y, z, p in xyz])
    leg = np.column_stack([rng.uniform(0, np.pi / 2, 32), rng.uniform(-0.9, 0.9, 32), rng.uniform(0, 0.99, 32)])
    out["legendre_in"] = leg
    out["legendre_out"] = np.array([[R.ref_ellf(ph, k), R.ref_elle(ph, k), R.ref_ellpi(ph, n, k)]
                                    for ph, n, k in leg])

    # quartic roots (real coefficients, mixed real/complex roots)
    coefs = []
    for _ in range(24):
        r = rng.uniform(-3, 3, 2)
        c = complex(rng.uniform(-3, 3), rng.uniform(0.1, 2))
        coefs.append(np.real(np.poly([r[0], r[1], c, c.conjugate()]))[::-1].copy())
    for _ in range(8):
        coefs.append(np.poly(rng.uniform(0.1, 20, 4))[::-1].copy())
    coefs = np.array(coefs)
    rr = np.zeros((len(coefs), 8))
    for i, c in enumerate(coefs):
        R.ref_quartic_roots(ol._ptr(c), ol._ptr(rr[i, :4]), ol._ptr(rr[i, 4:]))
    out["quartic_in"], out["quartic_out"] = coefs, rr

    # sparse inspirals: the 1-yr (N=2^18, dt=10 s) source, production-prior draws
    # and a source that plunges inside the window
    cases = [(8.0, 0.2, np.cos(0.5), 0.5, dt_fit(2 ** 18, 10.0, 1e6, 10.0)),
             (8.0, 0.1, np.cos(np.pi / 6), 0.5, dt_fit(3000000, 10.0, 1e6, 10.0)),
             (7.2, 0.3, np.cos(0.5), 0.9, dt_fit(2 ** 21, 15.0, 3e5, 30.0)),
             # plunges inside the window (flag 0), and starts inside the LSO (flag -1)
             (6.76, 0.5, 0.3, 0.2, 1 / 99), (6.85, 0.5, 0.3, 0.2, 0.3 / 99),
             (5.0, 0.3, 0.3, 0.2, 1 / 99)]
    while len(cases) < 11:
        a, e, p = rng.uniform(0, 0.95), rng.uniform(1e-3, 0.5), rng.uniform(7, 15)
        M, mu = 10 ** rng.uniform(5, 7), rng.uniform(5, 50)
        cases.append((p, e, np.cos(0.5), a, dt_fit(2 ** 21, 15.0, M, mu)))
    trajs, flags, lens = [], [], []
    for c in cases:
        f, arr = ol.r_eccentric(*c)
        trajs.append(np.pad(arr.reshape(-1, 12), ((0, 100 - len(arr)), (0, 0))))
        flags.append(f)
        lens.append(len(arr))
    out["traj_in"] = np.array(cases)
    out["traj_out"] = np.array(trajs)
    out["traj_flag"] = np.array(flags, dtype=np.int32)
    out["traj_steps"] = np.array(lens, dtype=np.int32)

    np.savez_compressed(os.path.join(HERE, "ref_subset.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_subset.npz"), {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
