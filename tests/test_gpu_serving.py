"""Serving loop on the device (MI355X only): every result is the same when the
request is served alone and when it is served in the mix with others in
flight (relative difference <= 1e-10), all residuals meet the bar, and
admission respects the slot count and the memory budget."""
import dataclasses

import numpy as np
import pytest

from kore_amd import backend, serving

pytestmark = pytest.mark.gpu


def small_trace():
    tr = serving.serving_trace(nreq=8, seed=7, e_range=(1e-4, 1e-3), n_cap=32, forced_frac=0.6, nev=4)
    return [dataclasses.replace(r, arrival=0.0) for r in tr]


@pytest.mark.parametrize("precision", [backend.PREC_FP64, backend.PREC_C64])
def test_alone_equals_mix(precision):
    trace = small_trace()
    print("case: serving alone vs mix, precision", precision, [(r.kind, r.N, r.m) for r in trace], flush=True)
    mix = serving.Server(slots=4, precision=precision, keep_results=True)
    res_mix = {r.rid: r for r in mix.run(trace)}
    inflight = mix.max_inflight
    mix.close()
    alone = serving.Server(slots=1, precision=precision, keep_results=True)
    res_alone = {}
    for r in trace:
        res_alone[r.rid] = alone.run([r])[0]
    alone.close()
    assert inflight >= 2            # requests really overlapped
    for r in trace:
        a, b = res_alone[r.rid], res_mix[r.rid]
        assert a.ok and b.ok, (r, a.error, b.error, a.residual, b.residual)
        if r.kind == "forced":
            d = np.linalg.norm(a.x - b.x) / np.linalg.norm(a.x)
        else:
            d = max(np.min(np.abs(b.lam - l)) / abs(l) for l in a.lam)
        assert d <= 1e-10, (r.rid, r.kind, d)


def test_admission_respects_budget():
    trace = small_trace()[:4]
    need = max(serving.footprint(r.params(), backend.PREC_FP64, r.nev) for r in trace)
    srv = serving.Server(slots=4, precision=backend.PREC_FP64, mem_budget=int(need * 1.5))
    out = srv.run(trace)
    assert all(r.ok for r in out)
    assert srv.max_inflight == 1        # the budget admits one request at a time
    assert srv.mem_peak <= int(need * 1.5)
    srv.close()


def test_oversized_request_is_refused_not_crashed():
    trace = small_trace()[:2]
    srv = serving.Server(slots=2, precision=backend.PREC_FP64, mem_budget=1 << 20)
    out = srv.run(trace)
    assert all((not r.ok) and "budget" in r.error for r in out)
    srv.close()


def test_failed_host_work_is_reported_and_the_rest_served():
    """A request whose host work raises (a failure forced in its operator
    assembly) comes back failed with the error; the requests around it are
    served and meet their checks."""
    trace = small_trace()[:3]
    bad = trace[1].rid

    class FailingServer(serving.Server):
        def _prepare(self, req):
            if req.rid == bad:
                raise ValueError("assembly failed for request %d" % req.rid)
            return super()._prepare(req)

    print("case: failed host work, request", bad, flush=True)
    srv = FailingServer(slots=2, precision=backend.PREC_FP64)
    out = srv.run(trace)
    srv.close()
    assert not out[1].ok and "assembly failed" in out[1].error
    assert out[0].ok and out[2].ok, (out[0].error, out[2].error)


def test_affine_plan_matches_formed_operator():
    """A system staged from a cached structure's A0, A1 and E
    (kore_sys_prepare_affine) is the system of the formed A = A0 + E A1: the
    same FP64 and complex64 + GMRES solves, bit for bit."""
    r = next(q for q in small_trace() if q.kind == "forced")
    cache = serving.OperatorCache()
    cache.operators(dataclasses.replace(r, Ek=r.Ek * 1.5))
    cache.operators(dataclasses.replace(r, Ek=r.Ek * 0.7))
    par, A, B = cache.operators(r)
    assert isinstance(A, serving.AffineA)
    print("case: affine plan vs formed operator, N", r.N, "m", r.m, flush=True)
    lay = backend.layout_from_params(par)
    ctx = backend.Context(0)
    S1 = backend.System(ctx, lay, None, None,
                        plan=backend.SystemPlan.affine(lay, A.indptr, A.indices, A.a0, A.a1, A.e, B, A.shape))
    S2 = backend.System(ctx, lay, A.materialize(), B)
    b = serving.rhs_for(r, S1.n)
    for prec in (backend.PREC_FP64, backend.PREC_C64):
        S1.factor(r.sigma, prec)
        S2.factor(r.sigma, prec)
        x1, _ = S1.solve_gmres(b, tol=1e-15, maxit=60)
        x2, _ = S2.solve_gmres(b, tol=1e-15, maxit=60)
        assert np.array_equal(x1, x2), prec
    S1.close()
    S2.close()
    ctx.close()


def test_device_residual_matches_host():
    """kore_residual + kore_sys_norm give the host formula's residuals (forced
    and eigen) for perturbed solutions (residuals well above rounding)."""
    trace = small_trace()
    rng = np.random.default_rng(3)
    ctx = backend.Context(0)
    for kind in ("forced", "eigen"):
        r = next(q for q in trace if q.kind == kind)
        print("case: device residual", kind, r.N, r.m, flush=True)
        par, A0, B0 = serving.assemble_request(r)
        S = backend.System(ctx, backend.layout_from_params(par), A0, B0)
        n = S.n
        if kind == "forced":
            b = serving.rhs_for(r, n)
            x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
            host = serving.residual(A0, B0, r, x=x, b=b)
            dev = serving.residual_device(S, r, x=x, b=b)
        else:
            X = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
            lam = rng.standard_normal(3) + 1j * rng.standard_normal(3)
            host = serving.residual(A0, B0, r, lam=lam, X=X)
            dev = serving.residual_device(S, r, lam=lam, X=X)
        assert abs(dev - host) <= 1e-10 * host, (kind, dev, host)
        S.close()
    ctx.close()


@pytest.mark.parametrize("precision", [backend.PREC_FP64, backend.PREC_C64])
def test_gmres_backward_error_is_the_check(precision):
    """The serving loop takes the backward error that kore_solve_gmres reports
    as the check of a forced result: it agrees with the host formula
    ||A x - b|| / (||A||_F ||x|| + ||b||) on the returned solution."""
    r = next(q for q in small_trace() if q.kind == "forced")
    print("case: gmres backward error as check, precision", precision, r.N, r.m, flush=True)
    par, A0, B0 = serving.assemble_request(r)
    ctx = backend.Context(0)
    S = backend.System(ctx, backend.layout_from_params(par), A0, B0)
    S.factor(r.sigma, precision)
    b = serving.rhs_for(r, S.n)
    x, info = S.solve_gmres(b, tol=1e-15, maxit=60)
    host = serving.residual(A0, B0, r, x=x, b=b)
    print("  iterations", info.iterations, "device", info.rel_correction, "host", host, flush=True)
    # both are at the rounding level of a residual evaluation (~eps): they
    # agree to eps, not relatively
    assert info.converged and host <= 1e-15 and info.rel_correction <= 1e-15
    assert abs(info.rel_correction - host) <= 2.3e-16
    S.close()
    ctx.close()
