"""Host-side logic that needs no GPU: the bench's one-GPU guard, the request
stream and latency statistics of the ensemble/serving drivers, the precision
report's formatting, and the serving C-ABI's argument checks."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_stream_sizes_do_not_depend_on_rate():
    import ensemble as E
    a = E.stream_cases(20, seed=3, rate=1.0)
    b = E.stream_cases(20, seed=3, rate=4.0)
    assert [(c.n, c.steps, c.radius, c.pratio, c.centre) for c in a] == \
           [(c.n, c.steps, c.radius, c.pratio, c.centre) for c in b]
    for x, y in zip(a, b):
        assert y.arrival == pytest.approx(x.arrival / 4.0)
    assert all(a[i].arrival < a[i + 1].arrival for i in range(len(a) - 1))
    assert all(32 <= c.n <= 192 and c.n % 16 == 0 and 50 <= c.steps <= 500 for c in a)


def test_ensemble_cases_and_work():
    import ensemble as E
    cs = E.ensemble_cases(64, 100, seed=0)
    assert len(cs) == 64 and {c.n for c in cs} <= {32, 48, 64, 96, 128}
    assert all(0.1 <= c.radius <= 0.3 and 10 <= c.pratio <= 1e4 for c in cs)
    c = cs[0]
    assert c.work == c.n ** 3 * 3 * c.steps
    assert E.ensemble_cases(64, 100, seed=0)[5].radius == cs[5].radius     # seeded


def test_latency_stats():
    import ensemble as E
    v = list(np.arange(1, 101, dtype=float))
    st = E.latency_stats(v)
    assert st["p50_s"] == pytest.approx(50.5)
    assert st["p99_s"] == pytest.approx(np.percentile(v, 99))
    assert st["max_s"] == 100 and st["mean_s"] == pytest.approx(50.5)


def test_precision_report_markdown():
    import precision
    F = precision.FIELDS
    v = {"rel_l1": {f: 1e-7 for f in F}, "rel_linf": {f: 2e-5 for f in F},
         "integral_diff": {f: 3e-9 for f in F}, "dt_own_rel_err_max": 1e-6, "seconds": 1.5,
         "p_rel_l1": 4e-6, "p_rel_linf": 5e-5}
    res = {"workload": "bubble-cloud collapse 32^3, 2 bubbles, 3 LSRK3 steps, CFL 0.3",
           "fp64": {"ms_per_step": 2.0, "t_end": 1e-3}, "variants": {"parity": v, "fast": v},
           "storage_floor": {"E": (3e-8, 6e-8), "p": (2e-6, 9e-5)}}
    md = precision.to_markdown(res, "python3 precision.py --n 32")
    assert md.startswith("# Precision")
    assert "| parity | 1.00e-07 |" in md and "| fast |" in md
    assert "| parity | 1.00e-07 | 2.00e-05 | 4.00e-06 | 5.00e-05 |" in md
    assert "| binary32 storage floor | 3.00e-08 | 6.00e-08 | 2.00e-06 | 9.00e-05 |" in md
    assert "python3 precision.py --n 32" in md


def test_server_and_precision_entry_points_check_arguments(mpcf):
    import ctypes as C
    lib = mpcf.lib()
    assert lib.mpcf_server_create(None, 1, 1 << 20, 0, 0, 4, 0.3) == mpcf.EINVAL
    assert lib.mpcf_server_submit(None, 8, 8, 8, 0.1, None, 1, None, None, None) == mpcf.EINVAL
    n = C.c_int()
    assert lib.mpcf_server_poll(None, None, None, 0, C.cast(C.byref(n), C.c_void_p)) == mpcf.EINVAL
    assert lib.mpcf_server_destroy(None) == mpcf.OK
    assert lib.mpcf_compare(None, None, 0, None) == mpcf.EINVAL
    assert lib.mpcf_integrals(None, 0, None) == mpcf.EINVAL
    assert lib.mpcf_upload_f64(None, 0, None) == mpcf.EINVAL
    assert lib.mpcf_server_clock() > 0.0
