"""The continuous-batching serve loop's host logic (CPU): a stand-in ensemble with the same
one-step-late snapshot semantics as mahyco_ensemble_poll_async / _poll_result (slots empty
when the snapshot was queued read -100, slots admitted but not yet stepped read 2).  Every
request must be retired with its own final state, also when it takes a slot freed one step
earlier, and on_done must see the request's own case."""
import numpy as np
import pytest

from mahyco_amd import serve
from mahyco_amd import types as T


class _Case:
    def __init__(self, ident, steps, fp=1.0):
        self.ident, self.steps, self.fp = ident, steps, fp


class _FakeEnsemble:
    def __init__(self, max_slots, budget):
        self.max_slots, self.budget = max_slots, budget
        self.slot = [None] * max_slots  # [case, steps done, pending]
        self.snap = None
        self.removed_unfinished = []

    def add(self, case, **kw):
        used = sum(s[0].fp for s in self.slot if s)
        if used + case.fp > self.budget:
            return None
        for i, s in enumerate(self.slot):
            if s is None:
                self.slot[i] = [case, 0, True]
                return i
        return None

    def step(self, n, graph=True):
        for s in self.slot:
            if s is None:
                continue
            if s[2]:
                s[2] = False  # activated at the start of the queued steps
            if s[1] < s[0].steps:
                s[1] += 1

    def poll_async(self):
        st = np.full(self.max_slots, -100, dtype=np.int32)
        it = np.zeros(self.max_slots, dtype=np.int32)
        for i, s in enumerate(self.slot):
            if s is None:
                continue
            if s[2]:
                st[i] = 2
            else:
                st[i] = T.MH_STOPPED if s[1] >= s[0].steps else T.MH_OK
                it[i] = s[1]
        self.snap = (st, it)

    def poll_result(self):
        assert self.snap is not None
        st, it = self.snap
        return int(np.sum((st == 0) | (st == 1))), st.copy(), np.zeros(self.max_slots), it.copy()

    def get(self, slot, field, env=-1, comp=0):
        return np.array([self.slot[slot][0].ident])

    def remove(self, slot):
        s = self.slot[slot]
        if s[1] < s[0].steps:
            self.removed_unfinished.append(s[0].ident)
        self.slot[slot] = None

    def close(self):
        pass


@pytest.mark.parametrize("max_slots,seed", [(1, 0), (2, 1), (3, 2), (5, 3)])
def test_serve_retires_each_request_with_its_own_state(monkeypatch, max_slots, seed):
    from mahyco_amd import lib
    made = []

    def make(n, b):
        made.append(_FakeEnsemble(n, b))
        return made[-1]

    monkeypatch.setattr(lib, "Ensemble", make)
    monkeypatch.setattr(lib, "footprint", lambda c: c.fp)
    rng = np.random.default_rng(seed)
    reqs = [(0.0, _Case(i, int(rng.integers(1, 6)))) for i in range(17)]
    seen = {}
    recs = serve.serve(reqs, budget=100.0, max_slots=max_slots, step_cycles=1,
                       on_done=lambda i, E, slot: seen.__setitem__(i, int(E.get(slot, 0)[0])))
    E = made[0]
    assert not E.removed_unfinished
    for i, (r, (_, c)) in enumerate(zip(recs, reqs)):
        assert r.status == T.MH_STOPPED, (i, r.status)
        assert r.iterations == c.steps, (i, r.iterations, c.steps)
        assert seen[i] == i
    assert all(s is None for s in E.slot)


def test_serve_refuses_a_request_larger_than_the_budget(monkeypatch):
    from mahyco_amd import lib
    monkeypatch.setattr(lib, "Ensemble", _FakeEnsemble)
    monkeypatch.setattr(lib, "footprint", lambda c: c.fp)
    reqs = [(0.0, _Case(0, 2)), (0.0, _Case(1, 2, fp=1e9)), (0.0, _Case(2, 3))]
    recs = serve.serve(reqs, budget=10.0, max_slots=4, step_cycles=1)
    assert [r.status for r in recs] == [T.MH_STOPPED, T.MH_ERR_NOMEM, T.MH_STOPPED]
    assert recs[1].iterations == -1
