"""GPU: the Phidot guard kernel alone (aak_test_phidot_guard) against the
reference's sequential rule (AAK.cc:254): walking i upward through
(i0, guard_last], a rate <= 0 is replaced by the value at i - 1, already
corrected, so it carries the last rate that is not <= 0 (NaN included), or
the rate at i0.  The kernel instead splits the range from the first
non-positive sample into 64 segments joined by a max look-back and scans
each segment in chunks; the cases below cover short tails, bad runs that
cross segment and chunk boundaries, bad samples from i0 + 1 on, NaNs and a
range with nothing to replace."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sequential(r, i0, last):
    r = r.copy()
    for i in range(i0 + 1, last + 1):
        if r[i] <= 0.0:
            r[i] = r[i - 1]
    return r


def _run(r, i0, last):
    from emri_kludge_suite_amd import _lib
    out = np.ascontiguousarray(r.copy())
    n = _lib.lib().aak_test_phidot_guard(_lib.dptr(out), out.size, i0, last)
    assert n >= 0, _lib.last_error()
    return out, n


def _cases():
    rng = np.random.default_rng(11)
    n = 1 << 20
    base = rng.uniform(0.1, 2.0, n)
    c = []
    a = base.copy(); a[rng.choice(n, 300, replace=False)] *= -1.0
    c.append(("sparse", a, 5, n - 10))
    a = base.copy(); a[n - 700:] = -rng.uniform(0, 1, 700)
    c.append(("tail", a, 0, n - 1))
    a = base.copy(); a[1:] = np.where(rng.uniform(size=n - 1) < 0.97, -1.0, a[1:])
    c.append(("mostly-bad", a, 0, n - 1))
    a = base.copy(); a[100:n - 3] = -0.5; a[n // 3] = np.nan
    c.append(("long-run-nan", a, 99, n - 2))
    a = base.copy(); a[40000:90000] = 0.0; a[700000:700001] = -2.0
    c.append(("zeros", a, 17, n - 1))
    a = base.copy()
    c.append(("nothing", a, 0, n - 1))
    a = base.copy(); a[3:9] = -1.0
    c.append(("short", a, 2, 8))
    return c


@pytest.mark.parametrize("case", range(7))
def test_guard_matches_the_sequential_rule(case):
    name, r, i0, last = _cases()[case]
    print("case guard %s i0=%d last=%d" % (name, i0, last), flush=True)
    got, count = _run(r, i0, last)
    want = _sequential(r, i0, last)
    assert count == int(np.count_nonzero(r[i0 + 1:last + 1] <= 0.0))
    assert np.array_equal(got, want, equal_nan=True)
