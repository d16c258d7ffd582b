"""CPU: the rendezvous logic of the multi-trait pass batcher (no GPU work:
the contraction is replaced by a host function) and the SLQ budget rule."""
import threading
import time
import types

import numpy as np
import pytest
import torch

from coherit_amd import slq
from coherit_amd.reml_model import FitConfig
from coherit_amd.traits import PassBatcher, _Request, _slq_budget


def _batcher(n_fits, log, fail_on=None):
    fake = types.SimpleNamespace(panel=types.SimpleNamespace(n_components=2))
    b = PassBatcher(fake, stats=None, n_fits=n_fits)

    def contract(reqs, prec):
        log.append(sorted(r.trait for r in reqs))
        if fail_on is not None and len(log) == fail_on:
            raise RuntimeError("contraction failed")
        for r in reqs:
            r.result = torch.stack([r.V * 2.0, r.V * 3.0])

    b._contract = contract
    return b


def _fit(b, t, calls, out, errors):
    try:
        for c in range(calls):
            time.sleep(0.001 * ((t * 7 + c * 3) % 5))  # host work of differing length
            V = torch.full((4, 1), float(10 * t + c))
            out.append((t, c, b.submit(_Request(t, V, "i8x2"))))
    except RuntimeError as e:
        errors.append((t, str(e)))
    finally:
        b.leave()


def test_every_batch_holds_one_request_of_each_live_fit():
    log, out, errors = [], [], []
    b = _batcher(4, log)
    calls = {0: 1, 1: 2, 2: 3, 3: 5}
    th = [threading.Thread(target=_fit, args=(b, t, calls[t], out, errors)) for t in range(4)]
    for x in th:
        x.start()
    for x in th:
        x.join(timeout=30)
    assert not errors
    # fits retire after 1, 2, 3, 5 calls: batches shrink as they leave
    assert log == [[0, 1, 2, 3], [1, 2, 3], [2, 3], [3], [3]]
    assert len(out) == sum(calls.values())
    for t, c, res in out:
        assert torch.equal(res[0], torch.full((4, 1), 2.0 * (10 * t + c)))
        assert torch.equal(res[1], torch.full((4, 1), 3.0 * (10 * t + c)))


def test_a_failed_contraction_reaches_every_waiting_fit():
    log, out, errors = [], [], []
    b = _batcher(3, log, fail_on=2)
    th = [threading.Thread(target=_fit, args=(b, t, 3, out, errors)) for t in range(3)]
    for x in th:
        x.start()
    for x in th:
        x.join(timeout=30)
    assert sorted(t for t, _ in errors) == [0, 1, 2]
    assert all(msg == "contraction failed" for _, msg in errors)
    assert len(out) == 3  # the first batch completed


@pytest.mark.parametrize("n,S,m", [(300_000, 16, 40), (1000, 8, 12), (50, 64, 40)])
def test_slq_budget_runs_the_whole_probe_block_in_one_batch(n, S, m):
    cfg = FitConfig(slq_samples=S, slq_m=m)
    ws = slq.default_workspace_bytes(_slq_budget(cfg, n))
    assert slq.workspace_layout(n, S, m, ws)["batch_width"] == S
