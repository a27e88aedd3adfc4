"""Host-side accounting of the Lanczos driver (no GPU): a simulated device whose thick restarts keep
fewer vectors than the host assumes, so that many queued step launches do nothing."""
from types import SimpleNamespace

import pytest

from specann_amd import eigsh as E
from specann_amd._lib import SA_MAXITER, SA_RESTART, SA_RUNNING, EigensolverBreakdown


class FakeDevice:
    """Steps advance j until the basis is full (status RESTART; further steps are no-ops); a thick
    restart keeps `kept` vectors (fewer than the host's keep). Stops at maxiter applications."""

    def __init__(self, ncv, kept, maxiter, stall_after=None):
        self.ncv, self.kept, self.maxiter = ncv, kept, maxiter
        self.j, self.steps, self.status, self.cycles = 0, 0, SA_RUNNING, 0
        self.step_launches = 0
        self.stall_after = stall_after

    def step(self):
        self.step_launches += 1
        if self.status != SA_RUNNING:
            return
        if self.stall_after is not None and self.steps >= self.stall_after:
            return                                    # a device that stopped making progress
        self.steps += 1
        self.j += 1
        if self.steps >= self.maxiter:
            self.status = SA_MAXITER
        elif self.j >= self.ncv:
            self.status = SA_RESTART

    def restart(self):
        if self.status == SA_RESTART:
            self.j, self.status = self.kept, SA_RUNNING
            self.cycles += 1

    def head(self):
        return SimpleNamespace(status=self.status, j=self.j, steps=self.steps, cycles=self.cycles, theta=1.0,
                               res=1.0, first_beta=0.5, r=[0.0] * 8)


def _solver(dev, ncv=33, keep=12, chunk=8):
    lz = object.__new__(E.LanczosSolver)
    lz.ncv, lz.keep, lz.chunk, lz.use_graphs = ncv, keep, chunk, False
    lz.work = SimpleNamespace(head=dev.head)
    lz._step = lambda s: dev.step()
    lz.thick_restart = lambda s: dev.restart()
    return lz


@pytest.mark.parametrize("kept,chunk", [(1, 8), (3, 8), (12, 8), (1, 32), (5, 33)])
def test_run_reaches_maxiter_when_restarts_keep_fewer_vectors(kept, chunk):
    """However many queued launches the device skips, the solve ends at the device's maxiter."""
    maxiter = 1000
    dev = FakeDevice(33, kept, maxiter)
    res = _solver(dev, chunk=chunk).run(None, maxiter)
    assert res.status == SA_MAXITER and res.steps == maxiter
    assert dev.step_launches >= maxiter


def test_run_raises_when_the_device_stops_progressing():
    dev = FakeDevice(33, 12, 1000, stall_after=100)
    with pytest.raises(EigensolverBreakdown, match="no progress"):
        _solver(dev).run(None, 1000)
    assert dev.steps == 100 and dev.step_launches < 100 + 4 * 8 + 1
