"""The request server (vti_serve_*): 2D gradient and 3D modeling requests of
different grid sizes, submitted together, run concurrently on worker slots
with resident models.  Every result must equal, bit for bit, the same
driver call made alone; admission must keep resident plus in-flight memory
within the budget; errors of one request must not disturb the others."""
import numpy as np
import pytest

from common2d import args2d, problem2d
from common3d_drivers import problem

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _problems2d(vti):
    return [problem2d(vti, nx=64, nz=48, npad=12, nt=161, ns=2, seed=3),
            problem2d(vti, nx=96, nz=40, npad=10, nt=201, ns=1, seed=4),
            problem2d(vti, nx=80, nz=64, npad=12, nt=121, ns=3, seed=5)]


def _obs2d(vti, ctx, P):
    return vti.synthetic2d(ctx, *P["ctrue"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))


def _submit_mix(vti, srv, ctx):
    """Submit 2D gradients on three grids (each with its own resident model)
    interleaved with 3D modeling on two grids; return the tickets with what
    each would give alone."""
    jobs = []
    P2 = _problems2d(vti)
    h2 = [srv.put_model(*P["c"]) for P in P2]
    P3 = [problem(), problem(nz=36, ny=40, nx=48, npad=8, nt=110, ns=1, seed=12)]
    h3 = [srv.put_model(*P["cij"]) for P in P3]
    for rnd in range(2):
        for P, h in zip(P2, h2):
            # scaled data: a nonzero residual even where the model's
            # perturbation has not reached the receivers yet
            obs = (_obs2d(vti, ctx, P) * np.float32(0.9 if rnd == 0 else 0.5)).astype(np.float32)
            t = srv.grad2d(h, obs, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
            alone = vti.objfunc_gradient2d(ctx, obs, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
            jobs.append(("2d", t, alone))
        for P, h in zip(P3, h3):
            shot = np.arange(len(P["souloc"]))
            t = srv.model3d(h, P["souloc"], shot, P["recloc"], P["w"], P["cij"].shape[1:], P["nt"], P["npad"], P["o"],
                            P["d"], P["dt"], P["rate"])
            alone = vti.model_data3d(ctx, P["souloc"], shot, P["recloc"], P["w"], *P["cij"], P["nt"], P["npad"],
                                     P["o"], P["d"], P["dt"], P["rate"])
            jobs.append(("3d", t, alone))
    return jobs


def _check(srv, jobs):
    for kind, t, alone in jobs:
        res, lat, svc = srv.wait(t)
        assert lat >= svc >= 0.0
        if kind == "2d":
            assert np.float32(res[0]) == np.float32(alone[0])
            assert np.abs(alone[1]).max() > 0
            assert np.array_equal(res[1], alone[1])
        else:
            assert np.abs(alone).max() > 0
            assert np.array_equal(res, alone)


@pytest.mark.parametrize("slots", [1, 4])
def test_mix_equals_alone(vti, ctx, slots):
    _say("serve mix, slots", slots)
    with vti.Server(0, slots=slots) as srv:
        jobs = _submit_mix(vti, srv, ctx)
        _check(srv, jobs)
        st = srv.stats()
        _say("  stats", st)
        assert st["completed"] == len(jobs)
        assert 0.0 < st["p50_s"] <= st["p99_s"]


def test_budget_bounds_memory_in_flight(vti, ctx):
    """A budget that holds the resident models and about two requests: every
    result is still exact and the peak stays within the budget."""
    P2 = _problems2d(vti)
    resident = sum(12 * P["nnx"] * P["nnz"] for P in P2)
    budget = resident + 2.5 * 4 * (P2[0]["nnx"] * P2[0]["nnz"]) * (19 + 2 * 41) + (3 << 20)
    _say("serve budget", budget)
    with vti.Server(0, slots=4, budget_bytes=budget) as srv:
        h = [srv.put_model(*P["c"]) for P in P2]
        jobs = []
        for rnd in range(3):
            for P, hh in zip(P2[:2], h[:2]):
                obs = (_obs2d(vti, ctx, P) * np.float32(0.9)).astype(np.float32)
                t = srv.grad2d(hh, obs, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
                alone = vti.objfunc_gradient2d(ctx, obs, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"],
                                               *args2d(P))
                jobs.append(("2d", t, alone))
        _check(srv, jobs)
        st = srv.stats()
        _say("  stats", st)
        assert st["peak_bytes"] <= budget


def test_refuses_request_over_budget(vti, ctx):
    P = problem2d(vti)
    with vti.Server(0, slots=2, budget_bytes=12 * P["nnx"] * P["nnz"] + (1 << 20)) as srv:
        h = srv.put_model(*P["c"])
        with pytest.raises(vti.VtiError, match="budget"):
            srv.grad2d(h, _obs2d(vti, ctx, P), P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))


def test_failed_request_leaves_others_intact(vti, ctx):
    """A request whose receivers lie outside the interior fails in its driver
    (a host-side check); the requests around it complete unchanged."""
    P = problem2d(vti)
    obs = _obs2d(vti, ctx, P)
    bad = P["rloc"].copy()
    bad[0, 1] = 0
    alone = vti.objfunc_gradient2d(ctx, obs, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    with vti.Server(0, slots=3) as srv:
        h = srv.put_model(*P["c"])
        t1 = srv.grad2d(h, obs, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
        tb = srv.grad2d(h, obs, P["w"], P["sloc"], bad, P["taper"], *args2d(P))
        t2 = srv.grad2d(h, obs, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
        with pytest.raises(vti.VtiError, match="receiver"):
            srv.wait(tb)
        for t in (t1, t2):
            (f, g), _, _ = srv.wait(t)
            assert np.array_equal(g, alone[1])
        with pytest.raises(vti.VtiError, match="no request"):
            srv.wait(tb)


def test_merged_requests_equal_alone(vti, ctx):
    """2D gradient requests of one schedule on models of different widths and
    shot counts, queued behind a 3D request on a single slot, run as merged
    batched runs; each result equals the same call made alone, bit for bit."""
    Ps = [problem2d(vti, nx=nx, nz=48, npad=10, nt=161, ns=ns, seed=sd)
          for nx, ns, sd in ((64, 2, 21), (96, 1, 22), (80, 3, 23))]
    obs, alone = [], []
    for P in Ps:
        o = (_obs2d(vti, ctx, P) * np.float32(0.9)).astype(np.float32)
        obs.append(o)
        alone.append(vti.objfunc_gradient2d(ctx, o, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P)))
    P3 = problem(nz=36, ny=40, nx=48, npad=8, nt=110, ns=1, seed=12)
    _say("serve merged: 3 models x 2 rounds behind one 3D request")
    with vti.Server(0, slots=1, batch=8) as srv:
        h2 = [srv.put_model(*P["c"]) for P in Ps]
        h3 = srv.put_model(*P3["cij"])
        shot = np.arange(len(P3["souloc"]))
        t3 = srv.model3d(h3, P3["souloc"], shot, P3["recloc"], P3["w"], P3["cij"].shape[1:], P3["nt"], P3["npad"],
                         P3["o"], P3["d"], P3["dt"], P3["rate"])
        tickets = [srv.grad2d(h, o, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
                   for _ in range(2) for P, h, o in zip(Ps, h2, obs)]
        srv.wait(t3)
        for k, t in enumerate(tickets):
            (f, g), lat, svc = srv.wait(t)
            fa, ga = alone[k % 3]
            assert lat >= svc >= 0.0
            assert np.abs(ga).max() > 0
            assert np.float32(f) == np.float32(fa)
            assert np.array_equal(g, ga)
        st = srv.stats()
        _say("  stats", st)
        assert st["merged_runs"] >= 1 and st["merged_requests"] >= 2


def test_batching_off_runs_requests_alone(vti, ctx):
    P = problem2d(vti)
    obs = _obs2d(vti, ctx, P)
    alone = vti.objfunc_gradient2d(ctx, obs, *P["c"], P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P))
    with vti.Server(0, slots=1, batch=1) as srv:
        h = srv.put_model(*P["c"])
        ts = [srv.grad2d(h, obs, P["w"], P["sloc"], P["rloc"], P["taper"], *args2d(P)) for _ in range(3)]
        for t in ts:
            (f, g), _, _ = srv.wait(t)
            assert np.array_equal(g, alone[1])
        assert srv.stats()["merged_runs"] == 0
