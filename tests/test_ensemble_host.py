"""The host-side ensemble loop of solve() (SciMLBase's EnsembleProblem
semantics: batch_size, reduction with early stop, output_func's rerun flag with
prob_func's repeat counter, safetycopy), on the CPU: the one GPU call per batch
is replaced by a stand-in that records what it was asked to solve and returns
each trajectory's initial state as its final state."""
import numpy as np
import pytest

import simplediffeq_hip as sd
from simplediffeq_hip import problem


@pytest.fixture
def calls(monkeypatch):
    log = []

    def fake_solve_many(eprob, alg, ens, idx, repeats, dt0s, order, kw):
        u0s, ps, tspans = problem._materialise(eprob, idx, repeats)
        log.append((list(idx), list(repeats)))
        sols = [problem.ODESolution(np.array([0.0, 1.0]), [u0s[:, j], u0s[:, j]], "Default",
                                    1, 0, alg) for j in range(len(idx))]
        return sols, ("buffers", tuple(idx))

    monkeypatch.setattr(problem, "_solve_many", fake_solve_many)
    return log


def _eprob(**kw):
    base = sd.ODEProblem(sd.rhs.expdecay3, [0.0, 0.0, 0.0], (0.0, 1.0), [-1.0])
    return sd.EnsembleProblem(
        base, prob_func=lambda prob, i, repeat: sd.remake(prob, u0=[float(i), float(repeat), 0.0]),
        **kw)


def test_batches_and_default_reduction(calls):
    esol = sd.solve(_eprob(), sd.GPUSimpleTsit5(), sd.EnsembleHIP(), trajectories=10,
                    batch_size=3, dt=0.1)
    assert [c[0] for c in calls] == [[0, 1, 2], [3, 4, 5], [6, 7, 8], [9]]
    assert [s.u[-1][0] for s in esol.u] == [float(i) for i in range(1, 11)]
    assert esol.converged is False and len(esol.raw) == 4 and esol.elapsedTime >= 0.0
    # one batch by default
    calls.clear()
    esol = sd.solve(_eprob(), sd.GPUSimpleTsit5(), sd.EnsembleHIP(), trajectories=10, dt=0.1)
    assert [c[0] for c in calls] == [list(range(10))] and esol.raw == ("buffers", tuple(range(10)))


def test_reduction_sees_batches_and_stops_when_converged(calls):
    seen = []

    def reduction(u, data, I):
        seen.append(list(I))
        u = u + sum(x for x in data)
        return u, len(seen) == 2

    ep = _eprob(output_func=lambda sol, i: (sol.u[-1][0], False), reduction=reduction, u_init=0.0)
    esol = sd.solve(ep, sd.GPUSimpleATsit5(), sd.EnsembleHIP(), trajectories=10, batch_size=4)
    assert seen == [[1, 2, 3, 4], [5, 6, 7, 8]]
    assert len(calls) == 2 and esol.converged is True
    assert esol.u == float(sum(range(1, 9)))


def test_rerun_resolves_with_next_repeat(calls):
    # odd trajectories ask for a rerun until their third repeat
    def output_func(sol, i):
        i_, rep = int(sol.u[-1][0]), int(sol.u[-1][1])
        assert i_ == i
        return (i, rep), (i % 2 == 1 and rep < 3)

    esol = sd.solve(_eprob(output_func=output_func), sd.GPUSimpleTsit5(), sd.EnsembleHIP(),
                    trajectories=6, batch_size=4, dt=0.1)
    assert esol.u == [(1, 3), (2, 1), (3, 3), (4, 1), (5, 3), (6, 1)]
    # reruns of a batch are solved together, one call per round
    assert calls == [([0, 1, 2, 3], [1, 1, 1, 1]), ([0, 2], [2, 2]), ([0, 2], [3, 3]),
                     ([4, 5], [1, 1]), ([4], [2]), ([4], [3])]


def test_output_func_must_return_a_pair(calls):
    ep = _eprob(output_func=lambda sol, i: sol)
    with pytest.raises(TypeError, match="output_func must return"):
        sd.solve(ep, sd.GPUSimpleTsit5(), sd.EnsembleHIP(), trajectories=2, dt=0.1)


def test_safetycopy(calls):
    # a prob_func that mutates its argument in place sees a fresh copy of the
    # prototype with safetycopy (the default), and the accumulated mutations without
    def mutating(prob, i, repeat):
        prob.u0[0] += 1.0
        return sd.remake(prob, u0=prob.u0.copy())

    base = sd.ODEProblem(sd.rhs.expdecay3, [0.0, 0.0, 0.0], (0.0, 1.0), [-1.0])
    esol = sd.solve(sd.EnsembleProblem(base, prob_func=mutating), sd.GPUSimpleTsit5(),
                    sd.EnsembleHIP(), trajectories=3, dt=0.1)
    assert [s.u[-1][0] for s in esol.u] == [1.0, 1.0, 1.0] and base.u0[0] == 0.0
    esol = sd.solve(sd.EnsembleProblem(base, prob_func=mutating, safetycopy=False),
                    sd.GPUSimpleTsit5(), sd.EnsembleHIP(), trajectories=3, dt=0.1)
    assert [s.u[-1][0] for s in esol.u] == [1.0, 2.0, 3.0] and base.u0[0] == 3.0


def test_array_form_batches_slice_columns(calls):
    u0s = np.arange(12.0).reshape(3, 4)
    base = sd.ODEProblem(sd.rhs.expdecay3, u0s[:, 0], (0.0, 1.0), [-1.0])
    ep = sd.EnsembleProblem(base, u0s=u0s, ps=np.full((1, 4), -1.0))
    esol = sd.solve(ep, sd.GPUSimpleTsit5(), sd.EnsembleHIP(), trajectories=4, batch_size=3,
                    dt=0.1)
    assert np.array_equal(np.stack([s.u[-1] for s in esol.u], axis=1), u0s)
    with pytest.raises(ValueError, match="batch_size"):
        sd.solve(ep, sd.GPUSimpleTsit5(), sd.EnsembleHIP(), trajectories=4, batch_size=0, dt=0.1)
