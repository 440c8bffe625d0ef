"""Serving loop: admission by slots and HBM budget, continuous batching, and
results that are the same alone and in the mix.

The CPU tests drive the loop with a stand-in context that records calls (no
GPU); the GPU test runs real replicas through the C ABI."""
import numpy as np
import pytest

from gpumdpd_amd import serve


class FakeSys:
    """Records calls; keeps a device-style step counter and step-completion
    clock (ctx.clock ticks 1 ms per step of any launch)."""

    def __init__(self, log, req, ctx):
        self.log, self.req, self.step, self.ctx = log, req, 0, ctx
        self.midstep = False
        self.dev_step, self.times = 0, {}

    def complete_step(self):
        self.ctx.clock += 1.0
        self.times[self.dev_step] = self.ctx.clock
        self.dev_step += 1

    def zero_step(self):
        self.midstep = True
        self.log.append(("zero", self.req.k))
        self.complete_step()

    def finish(self):
        self.log.append(("finish", self.req.k, self.step))
        self.complete_step()

    def destroy(self):
        self.log.append(("destroy", self.req.k))

    def download(self):
        return {"pos": np.zeros((1, 3)), "vel": np.zeros((1, 3)), "force": np.zeros((1, 3))}

    def thermo_rows(self):
        return np.zeros((0, 13))


class FakeCtx:
    def __init__(self):
        self.log = []
        self.max_batch = 0
        self.clock = 0.0

    def batch_advance(self, systems, k):
        assert all(s.midstep for s in systems)
        self.max_batch = max(self.max_batch, len(systems))
        for _ in range(k):
            for s in systems:
                s.complete_step()
        for s in systems:
            s.step += k
        self.log.append(("batch", tuple(s.req.k for s in systems), k))

    def sync(self):
        pass

    def admission(self, on):
        self.log.append(("admission", bool(on)))

    def step_times(self, systems):
        cap = 256
        t = np.zeros((len(systems), cap))
        for k, s in enumerate(systems):
            for st, v in s.times.items():
                t[k, st % cap] = v
        return t, np.array([s.dev_step for s in systems])


@pytest.fixture
def fake(monkeypatch):
    ctx = FakeCtx()
    monkeypatch.setattr(serve, "_create", lambda c, r, *a: FakeSys(c.log, r, c))
    return ctx


def reqs_of(sizes, steps, arrivals=None, nbytes=None):
    out = []
    for k, n in enumerate(sizes):
        r = serve.Request(k, "dpd", n, k + 1, arrivals[k] if arrivals else 0.0, steps[k])
        r.bytes = nbytes[k] if nbytes else n * 300
        out.append(r)
    return out


def test_every_request_runs_its_steps_exactly(fake):
    reqs = reqs_of([100] * 7, [5, 12, 1, 33, 16, 17, 2])
    m = serve.serve(fake, reqs, max_slots=3, chunk=4, realtime=False)
    fin = {e[1]: e[2] for e in fake.log if e[0] == "finish"}
    # zero_step is step 0 of steps; advance covers steps - 1 more
    assert fin == {r.k: r.steps - 1 for r in reqs}
    assert m["replica_steps"] == sum(r.steps for r in reqs)
    assert fake.max_batch == 3


def test_slots_and_budget_bound_admission(fake):
    reqs = reqs_of([100] * 6, [9] * 6, nbytes=[40, 40, 40, 40, 40, 40])
    serve.serve(fake, reqs, max_slots=5, hbm_budget=100, chunk=4, realtime=False)
    assert fake.max_batch == 2  # the budget (100 B) holds two 40-B replicas
    live, peak = set(), 0
    for e in fake.log:
        if e[0] == "zero":
            live.add(e[1])
        elif e[0] == "destroy":
            live.discard(e[1])
        peak = max(peak, len(live))
    assert peak == 2
    assert {e[1] for e in fake.log if e[0] == "destroy"} == set(range(6))


def test_latency_samples_cover_every_step(fake):
    reqs = reqs_of([100, 200], [10, 21])
    m = serve.serve(fake, reqs, max_slots=2, chunk=8, realtime=False)
    assert m["step_latency_p50_ms"] >= 0 and m["step_latency_p99_ms"] >= m["step_latency_p50_ms"]
    assert m["requests"] == 2
    # one sample per step of every request: the steps after the zero step, finish included
    assert m["step_latency_samples"] == sum(r.steps for r in reqs)


def test_request_stream_is_deterministic_and_in_range():
    a = serve.make_requests(5, n=512)
    b = serve.make_requests(5, n=512)
    assert [(r.style, r.n, r.seed, r.arrival) for r in a] == [(r.style, r.n, r.seed, r.arrival) for r in b]
    n = np.array([r.n for r in a])
    assert n.min() >= 2000 and n.max() <= 200000
    frac_dpd = np.mean([r.style == "dpd" for r in a])
    assert 0.6 < frac_dpd < 0.8
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert 0.004 < gaps.mean() < 0.006  # Poisson at 200/s


@pytest.mark.gpu
def test_served_mix_equals_alone():
    import gpumdpd_amd as g
    ctx = g.Context(0)
    try:
        reqs = serve.make_requests(5, n=10, steps=41, n_max=12000, n_scale=0.2)
        serve.prepare_inputs(reqs)
        keep = [0, 3, 7, 9]
        print("case: served mix of 10 replicas, 4 slots, 7-step chunks", flush=True)
        m = serve.serve(ctx, reqs, max_slots=4, chunk=7, log_every=10, keep=keep, realtime=False)
        assert m["replica_steps"] == sum(r.steps for r in reqs)
        for k in keep:
            r = reqs[k]
            s = g.system_from_config(ctx, r.cfg, precision=g.PREC_F32, log_every=10)
            s.zero_step()
            s.advance(r.steps - 1)
            s.finish()
            d = s.download()
            pos, vel, rows, _ = m["kept"][k]
            assert np.array_equal(d["pos"], pos) and np.array_equal(d["vel"], vel), f"request {k}"
            assert np.array_equal(s.thermo_rows(), rows)
            s.destroy()
    finally:
        ctx.close()


@pytest.mark.gpu
def test_c4_full_stream_served_equals_alone_and_oracle():
    """The whole C4 stream at full size (512 Poisson requests, 2,000-200,000 beads,
    70% DPD / 30% mDPD, 2,000 steps each, FP32 force).  Kept replicas -- the
    smallest and largest of each style and two in between -- give the same bits
    as alone; each one's final force (at its last step, from the served state)
    and its step-0 thermo row agree with the FP64 oracle."""
    import gpumdpd_amd as g
    from oracle import pyoracle
    ctx = g.Context(0)
    try:
        reqs = serve.make_requests(5, n=512, steps=2000)
        serve.prepare_inputs(reqs)
        keep = set()
        for style in ("dpd", "mdpd"):
            sub = sorted((r for r in reqs if r.style == style), key=lambda r: r.n)
            keep |= {sub[0].k, sub[-1].k, sub[len(sub) // 3].k}
        print(f"case: C4 full stream, {len(reqs)} requests, "
              f"{sum(r.n for r in reqs)} beads, kept {sorted(keep)}", flush=True)
        m = serve.serve(ctx, reqs, max_slots=128, chunk=32, log_every=100, keep=keep, realtime=False)
        print(f"served: {m['replica_steps_per_s']:.0f} replica-steps/s, p50 {m['step_latency_p50_ms']:.3f} ms, "
              f"p99 {m['step_latency_p99_ms']:.3f} ms over {m['step_latency_samples']} steps", flush=True)
        assert m["replica_steps"] == 512 * 2000
        assert m["step_latency_samples"] == 512 * 2000
        for k in sorted(keep):
            r = reqs[k]
            print(f"case: kept request {k} ({r.style}, {r.n} beads)", flush=True)
            s = g.system_from_config(ctx, r.cfg, precision=g.PREC_F32, log_every=100)
            s.zero_step()
            s.advance(r.steps - 1)
            s.finish()
            d = s.download()
            rows = s.thermo_rows()
            s.destroy()
            pos, vel, srows, force = m["kept"][k]
            assert np.array_equal(d["pos"], pos) and np.array_equal(d["vel"], vel), f"request {k}"
            assert np.array_equal(d["force"], force) and np.array_equal(rows, srows)
            # the final force against the oracle on the served state: the force
            # of step 2000 used the half-step velocity v - dt/2 F / m
            m_b = np.asarray(r.cfg["masses"])[np.asarray(r.cfg["type"])[np.argsort(r.cfg["id"])] - 1]
            v_half = vel - (0.5 * r.cfg["dt"]) * force / m_b[:, None]
            cfg_id = dict(r.cfg, id=np.arange(1, r.n + 1, dtype=np.int32),
                          type=np.asarray(r.cfg["type"])[np.argsort(r.cfg["id"])])
            f_ref = pyoracle.forces(cfg_id, pos=pos, vel=v_half, step=r.steps)[0]
            err = np.linalg.norm(force - f_ref) / np.linalg.norm(f_ref)
            assert err < 1e-5, (k, err)
            ref_rows = pyoracle.run(r.cfg, 0, log_every=100)[0]
            assert srows[0][11] == 0.0
            assert np.allclose(srows[0][:11], ref_rows[0], rtol=1e-5, atol=1e-6), k
    finally:
        ctx.close()
