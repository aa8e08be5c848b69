"""Serving runtime: every request served in a mix (continuous batching,
HBM-budget admission, requests overlapping on the device) returns exactly
the oracle's bytes for that request, i.e. what it returns alone."""
import ctypes

import numpy as np
import pytest

from conftest import oracle_codec

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def mods():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frsz_amd
    from frsz_amd import serve

    return frsz_amd, serve


def arrays(seed, count):
    from frsz_amd.workloads import smooth_signal_torch

    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        f64 = bool(rng.random() < 0.5)
        dt = torch.float64 if f64 else torch.float32
        bits = int(rng.choice([8, 12, 16, 21, 32] + ([40, 64] if f64 else [])))
        B = int(rng.choice([16, 32, 64, 5, 1]))
        n = int(np.exp(rng.uniform(np.log(10), np.log(400_000))))
        x = smooth_signal_torch(n, seed=500 + i, dtype=dt) * float(2.0 ** rng.integers(-10, 10))
        out.append((x, bits, B))
    return out


@pytest.mark.parametrize("schedule,xcd_order", [(0, 0), (1, 0), (0, 1)], ids=["tiles", "persistent", "xcd"])
def test_mix_equals_oracle(mods, port, schedule, xcd_order):
    fz, serve = mods
    specs = arrays(1, 40)
    rng = np.random.default_rng(2)
    with serve.Server(hbm_budget=1 << 30, slots=3, schedule=schedule, xcd_order=xcd_order) as srv:
        comp, dec, gat = [], [], []
        for x, bits, B in specs:
            t = srv.submit_compress(x, bits, B)
            comp.append(t)
            # readers submitted right behind their stream: they wait for it
            dec.append(srv.submit_decompress(t, x.numel(), x.dtype, bits, B))
            idx = torch.from_numpy(rng.integers(0, x.numel(), int(rng.integers(1, 4097)))).cuda()
            gat.append((srv.submit_gather(t, x.numel(), x.dtype, bits, B, idx), idx))
        for (x, bits, B), tc, td, (tg, idx) in zip(specs, comp, dec, gat):
            alone_s, alone_y = oracle_codec(port, x, bits, B)
            alone_g = alone_y[idx]
            rc = srv.wait(tc)
            assert rc["status"] == 0
            assert torch.equal(srv.output(rc, torch.uint8, alone_s.numel()), alone_s)
            rd = srv.wait(td)
            assert torch.equal(srv.output(rd, x.dtype, x.numel()).view(torch.uint8), alone_y.view(torch.uint8))
            rg = srv.wait(tg)
            assert torch.equal(srv.output(rg, x.dtype, idx.numel()).view(torch.uint8), alone_g.view(torch.uint8))
            assert rc["t_arrival"] <= rc["t_admit"] <= rc["t_launch"] <= rc["t_done"]
            for t in (tg, td, tc):
                srv.release(t)
        st = srv.stats()
        assert st["completed"] == 120 and st["pool_used"] == 0
        assert st["launched_requests"] == 120


def test_budget_admission_and_release(mods, port):
    """Outputs that together exceed the budget are admitted as memory frees."""
    fz, serve = mods
    x = torch.randn(200_000, device="cuda", dtype=torch.float32)
    size = fz.compressed_size(torch.float32, 32, 32, x.numel())
    slot = -(-size // 256) * 256  # pool allocations are 256-byte granular
    with serve.Server(hbm_budget=3 * slot + 256, slots=2) as srv:
        big = torch.randn(4 * size, device="cuda", dtype=torch.float32)
        with pytest.raises(fz.Frsz2Error) as e:
            srv.submit_compress(big, 32, 32)
        assert e.value.code == 3
        held = [srv.submit_compress(x, 32, 32) for _ in range(3)]
        late = srv.submit_compress(x * 2, 32, 32)  # no room until one is released
        for t in held:
            srv.wait(t)
        assert srv.stats()["pool_used"] >= 3 * size - 3 * 256
        srv.release(held[0])
        r = srv.wait(late)
        assert torch.equal(srv.output(r, torch.uint8, size), oracle_codec(port, x * 2, 32, 32)[0])
        for t in held[1:] + [late]:
            srv.release(t)


def test_replay_trace_equals_oracle(mods, port):
    fz, serve = mods
    specs = arrays(3, 24)
    rng = np.random.default_rng(4)
    trace = np.zeros(0, dtype=serve.REQUEST_DTYPE)
    rows, keep_alive, meta = [], [], []
    t = 0.0
    for i, (x, bits, B) in enumerate(specs):
        t += float(rng.exponential(2e-4))
        ci = len(rows)
        rows.append((0, int(x.dtype == torch.float64), bits, 0, B, x.numel(), x.data_ptr(), 0, 0, 0, t))
        meta.append(("c", i, None))
        keep_alive.append(x)
        t += float(rng.exponential(2e-4))
        rows.append((1, int(x.dtype == torch.float64), bits, 0, B, x.numel(), 0, ci, 0, 0, t))
        meta.append(("d", i, None))
        idx = torch.from_numpy(rng.integers(0, x.numel(), 100)).cuda()
        keep_alive.append(idx)
        t += float(rng.exponential(2e-4))
        rows.append((2, int(x.dtype == torch.float64), bits, 0, B, x.numel(), 0, ci, idx.data_ptr(), 100, t))
        meta.append(("g", i, idx))
    trace = np.array(rows, dtype=serve.REQUEST_DTYPE)
    with serve.Server(hbm_budget=1 << 30, slots=4) as srv:
        res, tickets = srv.replay(trace)
        assert np.all(res["status"] == 0)
        assert np.all(res["t_done"] >= res["t_arrival"])
        want_all = [oracle_codec(port, x, bits, B) for x, bits, B in specs]
        for r, tk, (kind, i, idx) in zip(res, tickets, meta):
            x, bits, B = specs[i]
            s, y = want_all[i]
            rd = {k: r[k] for k in r.dtype.names}
            if kind == "c":
                got, want = srv.output(rd, torch.uint8, s.numel()), s
            elif kind == "d":
                got = srv.output(rd, x.dtype, x.numel()).view(torch.uint8)
                want = y.view(torch.uint8)
            else:
                got = srv.output(rd, x.dtype, idx.numel()).view(torch.uint8)
                want = y[idx].view(torch.uint8)
            assert torch.equal(got, want), (kind, i)
        for tk in tickets:
            srv.release(int(tk))
        assert srv.stats()["pool_used"] == 0


def test_large_requests_batch_apart(mods, port):
    """Requests moving at least large_bytes are never batched with smaller
    ones (a launch completes as a whole, so a small request batched with a
    large one would wait for its bytes), and their launches hold at most
    slots - 1 slots at once.  Every result equals the oracle's."""
    fz, serve = mods
    from frsz_amd.workloads import smooth_signal_torch

    large_bytes = 1 << 20
    rng = np.random.default_rng(9)
    specs = []
    for i in range(30):
        big = i % 5 == 0
        n = int(rng.integers(300_000, 500_000)) if big else int(rng.integers(100, 4000))
        x = smooth_signal_torch(n, seed=700 + i, dtype=torch.float32)
        specs.append((x, 16, 32, big))
    rows = [(0, 0, bits, 0, B, x.numel(), x.data_ptr(), 0, 0, 0, 0.0) for x, bits, B, _ in specs]
    trace = np.array(rows, dtype=serve.REQUEST_DTYPE)  # all released at once: admission batches what it can
    with serve.Server(hbm_budget=1 << 30, slots=3, large_bytes=large_bytes) as srv:
        res, tickets = srv.replay(trace)
        assert np.all(res["status"] == 0)
        launch = res["t_launch"]
        for i, (x, bits, B, big) in enumerate(specs):
            assert big == (x.numel() * 4 + fz.compressed_size(torch.float32, bits, B, x.numel()) >= large_bytes)
            together = [j for j in range(len(specs)) if launch[j] == launch[i]]
            assert all(specs[j][3] == big for j in together), f"request {i} shared a launch across sizes"
            want = oracle_codec(port, x, bits, B)[0]
            r = {k: res[i][k] for k in res.dtype.names}
            assert torch.equal(srv.output(r, torch.uint8, want.numel()), want), i
            srv.release(int(tickets[i]))
        # at most slots - 1 = 2 large launches overlap at any time
        big_launches = sorted({(float(res["t_launch"][i]), float(res["t_done"][i])) for i, s in enumerate(specs)
                               if s[3]})
        for t0, _ in big_launches:
            assert sum(1 for a, b in big_launches if a <= t0 < b) <= 2


def test_reader_waits_for_queued_large_source(mods, port):
    """A decompress or gather whose stream's compress is still queued (a
    large request waiting for one of the slots - 1 large slots) waits for it
    instead of failing."""
    fz, serve = mods
    from frsz_amd.workloads import smooth_signal_torch

    xs = [smooth_signal_torch(4_000_000 + 1000 * i, seed=800 + i, dtype=torch.float32) for i in range(3)]
    idx = torch.arange(0, 4_000_000, 997, device="cuda", dtype=torch.int64)
    rows = []
    for i, x in enumerate(xs):
        rows.append((0, 0, 16, 0, 32, x.numel(), x.data_ptr(), 0, 0, 0, 0.0))
    for i, x in enumerate(xs):
        rows.append((2, 0, 16, 0, 32, x.numel(), 0, i, idx.data_ptr(), idx.numel(), 0.0))
        rows.append((1, 0, 16, 0, 32, x.numel(), 0, i, 0, 0, 0.0))
    trace = np.array(rows, dtype=serve.REQUEST_DTYPE)
    with serve.Server(hbm_budget=1 << 30, slots=2, large_bytes=1 << 20) as srv:
        res, tickets = srv.replay(trace)
        assert np.all(res["status"] == 0), res["status"]
        for i, x in enumerate(xs):
            s, y = oracle_codec(port, x, 16, 32)
            r = {k: res[i][k] for k in res.dtype.names}
            assert torch.equal(srv.output(r, torch.uint8, s.numel()), s)
            rg = {k: res[3 + 2 * i][k] for k in res.dtype.names}
            assert torch.equal(srv.output(rg, torch.float32, idx.numel()).view(torch.uint8), y[idx].view(torch.uint8))
            rd = {k: res[4 + 2 * i][k] for k in res.dtype.names}
            assert torch.equal(srv.output(rd, torch.float32, x.numel()).view(torch.uint8), y.view(torch.uint8))
        for t in tickets:
            if t:
                srv.release(int(t))


def test_held_pool_fails_waiting_requests_together(mods, port):
    """A pool held by outputs the client does not release: once nothing has
    run for stall_ms, every queued request that has waited that long fails
    with code 3 in one pass (not one per stall_ms), each after waiting its
    own stall_ms; after a release the server admits again."""
    fz, serve = mods
    x = torch.randn(100_000, device="cuda", dtype=torch.float32)
    size = fz.compressed_size(torch.float32, 16, 32, x.numel())
    slot = -(-size // 256) * 256
    stall = 0.2
    with serve.Server(hbm_budget=2 * slot + 256, slots=2, stall_ms=int(stall * 1000)) as srv:
        held = [srv.submit_compress(x, 16, 32) for _ in range(2)]
        for t in held:
            assert srv.wait(t)["status"] == 0
        late = [srv.submit_compress(x * (k + 2), 16, 32) for k in range(4)]
        res = []
        for t in late:
            # the C call fills the result of a failed request too (wait()
            # raises on it)
            r = serve.Result()
            rc = serve._L.frsz2_server_wait(srv._h, t, ctypes.byref(r))
            assert rc == 3
            res.append(serve.result_dict(r))
        assert [r["status"] for r in res] == [3, 3, 3, 3]
        for r in res:
            assert r["t_done"] - r["t_arrival"] >= stall
        # one pass fails them all: the old form spaced the failures stall_ms apart
        done = [r["t_done"] for r in res]
        assert max(done) - min(done) < stall / 2
        srv.release(held[0])
        t = srv.submit_compress(x * 7, 16, 32)
        r = srv.wait(t)
        assert r["status"] == 0
        assert torch.equal(srv.output(r, torch.uint8, size), oracle_codec(port, x * 7, 16, 32)[0])
        for t2 in (held[1], t):
            srv.release(t2)
