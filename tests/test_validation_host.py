"""Host logic of the validation summary's EMD runner (no device needed):
distances come back in submission order whatever order the worker threads
finish in, at most 2 x workers cost matrices are held at once, and one worker
runs in line."""
import time
from concurrent.futures import Future

import pytest

pytest.importorskip('torch')


def _stub(monkeypatch):
    from wot_amd.ot import emd as E
    monkeypatch.setattr(E, 'emd_inputs', lambda a, b, e: (a, b, e))

    def fake_emd2(p, q, M, numItermax):
        time.sleep(0.002 * ((7 * p) % 5))  # finish out of order
        return float(p) ** 2
    monkeypatch.setattr(E, 'emd2', fake_emd2)


def test_runner_keeps_submission_order_and_bounds_pending(monkeypatch):
    from wot_amd.ot import validation as V
    _stub(monkeypatch)
    r = V._EmdRunner(4)
    out = []
    for i in range(40):
        out.append(r.submit(i, 0, None))
        assert len(r.pending) <= 2 * 4
    r.close()
    got = [d.result() if isinstance(d, Future) else d for d in out]
    assert got == [float(i) for i in range(40)]


def test_runner_with_one_worker_runs_in_line(monkeypatch):
    from wot_amd.ot import validation as V
    _stub(monkeypatch)
    r = V._EmdRunner(1)
    d = r.submit(3, 0, None)
    r.close()
    assert d == 3.0 and not isinstance(d, Future)


def test_host_threads_bounded():
    from wot_amd.ot import validation as V
    assert 1 <= V._host_threads() <= 16
