"""Batched path: every job of a graph-captured batch must produce exactly the
oracle's bytes (stream, decode, gather) for its array, for any launch
configuration and on every graph replay."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def fz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frsz_amd
    from frsz_amd import batch

    return frsz_amd, batch


def make_arrays(seed, count=60):
    from frsz_amd.workloads import smooth_signal_torch

    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        f64 = rng.random() < 0.4
        dt = torch.float64 if f64 else torch.float32
        W = 64 if f64 else 32
        bits = int(rng.choice([4, 8, 9, 12, 16, 21, 32] + ([40, 52, 64] if f64 else [])))
        if f64 and bits <= 8:
            bits = 16
        B = int(rng.choice([1, 2, 4, 16, 32, 64, 256, 1024, 5, 100]))
        n = int(np.exp(rng.uniform(0, np.log(300_000)))) if i % 10 else int(rng.choice([0, 1, 31]))
        x = smooth_signal_torch(max(n, 1), seed=1000 + i, dtype=dt)[:n].contiguous()
        x.mul_(float(2.0 ** rng.integers(-20, 20)))
        specs.append((x, bits, B))
        assert bits <= W
    return specs


def oracle_outputs(port, specs):
    """The oracle's stream and decode of every array, as device tensors."""
    from conftest import oracle_codec

    streams, values = [], []
    for x, bits, B in specs:
        s, y = oracle_codec(port, x, bits, B)
        streams.append(s)
        values.append(y)
    return streams, values


CONFIGS = [None,
           {"schedule": 1, "use_graph": 1},
           {"schedule": 1, "workgroups": 1, "use_graph": 0},
           {"xcd_order": 1, "use_graph": 1},
           {"schedule": 1, "xcd_order": 1, "workgroups": 2, "use_graph": 1},
           {"plain_launches": 1, "use_graph": 1},
           {"use_graph": 0}]


@pytest.mark.parametrize("cfg", CONFIGS, ids=["default", "persistent", "persistent-1-per-cu-nograph", "xcd",
                                             "persistent-xcd", "plain-graph", "nograph"])
def test_batch_matches_oracle(fz, port, cfg):
    fz_mod, batch = fz
    specs = make_arrays(1)
    ref_s, ref_v = oracle_outputs(port, specs)
    outs = [torch.full_like(s, 0x5A) for s in ref_s]
    plan = batch.BatchPlan([batch.compress_job(x, o, b, B) for (x, b, B), o in zip(specs, outs)], cfg)
    plan.launch()
    vals = [torch.full_like(x, float("nan")) for x, _, _ in specs]
    dplan = batch.BatchPlan([batch.decompress_job(s, v, b, B) for (x, b, B), s, v in zip(specs, ref_s, vals)], cfg)
    dplan.launch()
    torch.cuda.synchronize()
    for i, ((x, b, B), o, r) in enumerate(zip(specs, outs, ref_s)):
        assert torch.equal(o, r), f"job {i}: dtype={x.dtype} bits={b} B={B} n={x.numel()}"
    for i, (v, r) in enumerate(zip(vals, ref_v)):
        assert torch.equal(v.view(torch.uint8), r.view(torch.uint8)), f"decompress job {i}"
    plan.close()
    dplan.close()


@pytest.mark.parametrize("count,cfg", [(1, None), (3, {"use_graph": 0}), (8, {"use_graph": 1}),
                                       (9, {"use_graph": 0}), (5, {"schedule": 1}),
                                       (4, {"schedule": 1, "xcd_order": 1, "use_graph": 1}),
                                       (3, {"xcd_order": 1, "use_graph": 0}), (8, {"xcd_order": 1, "use_graph": 1}),
                                       (9, {"xcd_order": 1, "use_graph": 1})])
def test_few_jobs(fz, port, count, cfg):
    """Up to eight descriptors of a class travel in the kernel arguments of a
    one-workgroup-per-tile launch (no descriptor copy); nine, or the
    persistent schedule, read them from device memory."""
    fz_mod, batch = fz
    specs = make_arrays(30 + count, count=count)
    ref_s, ref_v = oracle_outputs(port, specs)
    outs = [torch.full_like(s, 0x77) for s in ref_s]
    plan = batch.BatchPlan([batch.compress_job(x, o, b, B) for (x, b, B), o in zip(specs, outs)], cfg)
    vals = [torch.full_like(x, float("nan")) for x, _, _ in specs]
    dplan = batch.BatchPlan([batch.decompress_job(s, v, b, B) for (x, b, B), s, v in zip(specs, ref_s, vals)], cfg)
    for rep in range(2):
        plan.launch()
        dplan.launch()
        torch.cuda.synchronize()
        for i, (o, r) in enumerate(zip(outs, ref_s)):
            assert torch.equal(o, r), f"rep {rep} job {i}"
        for i, (v, r) in enumerate(zip(vals, ref_v)):
            assert torch.equal(v.view(torch.uint8), r.view(torch.uint8)), f"rep {rep} decompress job {i}"
        for o in outs:
            o.fill_(0x33)
        for v in vals:
            v.fill_(float("nan"))
    plan.close()
    dplan.close()


@pytest.mark.parametrize("schedule", [0, 1])
@pytest.mark.parametrize("count", [1, 5, 8, 12, 13, 64])
def test_xcd_order_tile_counts(fz, port, schedule, count):
    """XCD order: a class of `count` one-tile jobs (float32, bits 16, B 32),
    tile counts below, at and above a multiple of the 8 XCDs, on both
    schedules; every tile is taken exactly once."""
    fz_mod, batch = fz
    from frsz_amd.workloads import smooth_signal_torch

    specs = [(smooth_signal_torch(1000 - 7 * i, seed=500 + i, dtype=torch.float32), 16, 32) for i in range(count)]
    ref_s, ref_v = oracle_outputs(port, specs)
    cfg = {"schedule": schedule, "xcd_order": 1, "use_graph": 1}
    outs = [torch.full_like(s, 0x5A) for s in ref_s]
    plan = batch.BatchPlan([batch.compress_job(x, o, b, B) for (x, b, B), o in zip(specs, outs)], cfg)
    vals = [torch.full_like(x, float("nan")) for x, _, _ in specs]
    dplan = batch.BatchPlan([batch.decompress_job(s, v, b, B) for (x, b, B), s, v in zip(specs, ref_s, vals)], cfg)
    plan.launch()
    dplan.launch()
    torch.cuda.synchronize()
    for i, (o, r) in enumerate(zip(outs, ref_s)):
        assert torch.equal(o, r), f"{count} jobs schedule {schedule} job {i}"
    for i, (v, r) in enumerate(zip(vals, ref_v)):
        assert torch.equal(v.view(torch.uint8), r.view(torch.uint8)), f"{count} jobs decompress job {i}"
    plan.close()
    dplan.close()


def test_graph_replays_with_eager_work_between(fz, port):
    fz_mod, batch = fz
    specs = make_arrays(2, count=30)
    ref_s, _ = oracle_outputs(port, specs)
    outs = [torch.empty_like(s) for s in ref_s]
    plan = batch.BatchPlan([batch.compress_job(x, o, b, B) for (x, b, B), o in zip(specs, outs)])
    assert plan.config["use_graph"] == 1
    for rep in range(4):
        for o in outs:
            o.fill_(rep + 1)
        torch.cuda.synchronize()
        plan.launch()
        torch.cuda.synchronize()
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        for i, (o, r) in enumerate(zip(outs, ref_s)):
            assert torch.equal(o, r), f"replay {rep} job {i}"
        del junk


SHAPES = [(v, w) for v in (2, 4, 8, 16) for w in (64, 256)]


@pytest.mark.parametrize("vpt,wgs", SHAPES, ids=[f"{v}x{w}" for v, w in SHAPES])
def test_tile_shapes_match_oracle(fz, port, vpt, wgs):
    """Every class present in one tile shape (values per thread x workgroup
    size; float32 classes take 4 values per thread where 2 is asked): the
    bytes stay the oracle's, and jobs whose block exceeds the tile run as
    plain launches.  The plan then switches to the built-in shapes, which
    rebuilds its descriptors, and runs again."""
    fz_mod, batch = fz
    specs = make_arrays(5, count=40)
    ref_s, ref_v = oracle_outputs(port, specs)

    def shapes(op):
        d = {}
        for x, b, B in specs:
            two = x.dtype == torch.float64 and b > 8  # U >= 16
            d[batch.job_class(op, x.dtype, b)] = [vpt if vpt != 2 or two else 4, wgs]
        return d

    outs = [torch.full_like(s, 0x5A) for s in ref_s]
    plan = batch.BatchPlan([batch.compress_job(x, o, b, B) for (x, b, B), o in zip(specs, outs)],
                           {"tile_shapes": shapes(0)})
    assert plan.config["tile_shapes"] == shapes(0)
    vals = [torch.full_like(x, float("nan")) for x, _, _ in specs]
    dplan = batch.BatchPlan([batch.decompress_job(s, v, b, B) for (x, b, B), s, v in zip(specs, ref_s, vals)],
                            {"tile_shapes": shapes(1), "use_graph": 0})
    for rnd in range(2):
        print(f"tile shape {vpt}x{wgs} round {rnd}", flush=True)
        for o in outs:
            o.fill_(0x5A)
        for v in vals:
            v.fill_(float("nan"))
        plan.launch()
        dplan.launch()
        torch.cuda.synchronize()
        for i, ((x, b, B), o, r) in enumerate(zip(specs, outs, ref_s)):
            assert torch.equal(o, r), f"round {rnd} job {i}: dtype={x.dtype} bits={b} B={B} n={x.numel()}"
        for i, (v, r) in enumerate(zip(vals, ref_v)):
            assert torch.equal(v.view(torch.uint8), r.view(torch.uint8)), f"round {rnd} decompress job {i}"
        plan.config = {"tile_shapes": {}}
        dplan.config = {"tile_shapes": {}}
        assert plan.config["tile_shapes"] == {}


def test_autotune_keeps_bytes(fz, port):
    fz_mod, batch = fz
    specs = make_arrays(3, count=40)
    ref_s, _ = oracle_outputs(port, specs)
    outs = [torch.empty_like(s) for s in ref_s]
    plan = batch.BatchPlan([batch.compress_job(x, o, b, B) for (x, b, B), o in zip(specs, outs)])
    chosen = plan.autotune(reps=1)
    # whichever candidate wins the timing, the plan keeps a valid schedule,
    # valid tile shapes and the bytes
    assert chosen["schedule"] in (0, 1) and chosen["xcd_order"] in (0, 1)
    for cls, (vpt, wgs) in chosen["tile_shapes"].items():
        assert chosen["schedule"] == 0 and 0 <= cls < 14
        assert vpt in (0, 2, 4, 8, 16) and wgs in (0, 64, 256)
    assert chosen == plan.config
    for o in outs:
        o.zero_()
    plan.launch()
    torch.cuda.synchronize()
    for o, r in zip(outs, ref_s):
        assert torch.equal(o, r)


def test_batch_rejects_bad_jobs(fz):
    fz_mod, batch = fz
    x = torch.ones(100, device="cuda")
    small = torch.empty(10, dtype=torch.uint8, device="cuda")
    with pytest.raises(fz_mod.Frsz2Error) as e:
        batch.BatchPlan([batch.compress_job(x, small, 16, 32)])
    assert e.value.code == 3
    out = torch.empty(fz_mod.compressed_size(torch.float64, 33, 32, 100), dtype=torch.uint8, device="cuda")
    with pytest.raises(fz_mod.Frsz2Error) as e:
        batch.BatchPlan([batch.compress_job(x, out, 33, 32)])
    assert e.value.code == 2 and "Unsupported number of bits: 33" in str(e.value)


@pytest.mark.parametrize("cfg", [None, {"schedule": 1, "use_graph": 0}, {"xcd_order": 1, "use_graph": 0},
                                 {"schedule": 1, "xcd_order": 1, "use_graph": 1}],
                         ids=["default", "persistent", "xcd", "persistent-xcd"])
def test_decompress_and_gather_of_arbitrary_streams(fz, port, cfg):
    """Streams no compress writes (random payload and block exponents over
    the range where the reference decode is defined, E + bias < 0 included):
    every batched decompress and gather equals the oracle's, for each tile
    schedule."""
    fz_mod, batch = fz
    from oracle.oracle import compressed_bytes, storage_bits

    rng = np.random.default_rng(17)
    streams, outs, gouts, idxs, specs = [], [], [], [], []
    for i, (dt, bits, B) in enumerate([(torch.float32, 12, 32), (torch.float32, 21, 16), (torch.float32, 16, 64),
                                       (torch.float32, 7, 8), (torch.float64, 21, 32), (torch.float64, 40, 128),
                                       (torch.float64, 32, 32), (torch.float64, 9, 4), (torch.float32, 32, 1024)]):
        n = int(rng.integers(1, 60_000))
        U = storage_bits(bits)
        nb = -(-n // B)
        s = rng.integers(0, 256, compressed_bytes(bits, B, n), dtype=np.uint8)
        et = {8: np.int8, 16: np.int16, 32: np.int32, 64: np.int64}[U]
        lo = max(int(np.iinfo(et).min), -151 if dt == torch.float32 else -1076)
        hi = min(int(np.iinfo(et).max), 40_000)
        s[: nb * U // 8] = rng.integers(lo, hi, nb, endpoint=True).astype(et).view(np.uint8)
        sd = torch.from_numpy(s).cuda()
        streams.append(sd)
        outs.append(torch.full((n,), float("nan"), dtype=dt, device="cuda"))
        idx = torch.from_numpy(rng.integers(0, n, 777).astype(np.int64)).cuda()
        idxs.append(idx)
        gouts.append(torch.full((777,), float("nan"), dtype=dt, device="cuda"))
        specs.append((n, dt, bits, B))
    jobs = [batch.decompress_job(s, o, b, B) for s, o, (n, dt, b, B) in zip(streams, outs, specs)]
    jobs += [batch.gather_job(s, n, g, ix, b, B) for s, g, ix, (n, dt, b, B) in zip(streams, gouts, idxs, specs)]
    plan = batch.BatchPlan(jobs, cfg)
    plan.launch()
    torch.cuda.synchronize()
    for i, (s, o, g, ix, (n, dt, b, B)) in enumerate(zip(streams, outs, gouts, idxs, specs)):
        npdt = np.float32 if dt == torch.float32 else np.float64
        want = port.decompress(s.cpu().numpy(), n, npdt, b, B)
        assert np.array_equal(o.cpu().numpy().view(np.uint8), want.view(np.uint8)), \
            f"decompress job {i}: {dt} b{b} B{B} n{n}"
        wg = port.gather(s.cpu().numpy(), n, npdt, b, B, ix.cpu().numpy().astype(np.uint64))
        assert np.array_equal(g.cpu().numpy().view(np.uint8), wg.view(np.uint8)), f"gather job {i}"
    plan.close()


def test_many_small_jobs(fz, port):
    """8000 small jobs of every kind in one plan (compress, decompress and
    gather; float32 and float64; power-of-two and other block sizes, so
    hundreds of plain launches sit in the graph beside the grouped
    kernels; empty arrays too): every job's bytes equal the oracle's, on
    two graph replays."""
    fz_mod, batch = fz
    rng = np.random.default_rng(8000)
    jobs, outs, wants = [], [], []
    keep = []
    for i in range(8000):
        f64 = rng.random() < 0.4
        npdt = np.float64 if f64 else np.float32
        bits = int(rng.choice([5, 8, 12, 16, 21, 32] + ([40, 64] if f64 else [])))
        if f64 and bits <= 8:
            bits = 12
        B = int(rng.choice([2, 4, 8, 16, 32, 64, 128, 3, 5, 7, 12, 100]))
        n = int(rng.integers(0, 3000)) if i % 50 else 0
        x = (rng.standard_normal(n) * np.exp2(rng.integers(-30, 30, n))).astype(npdt)
        s = port.compress(x, bits, B)
        kind = i % 3
        if kind == 0:
            xd = torch.from_numpy(x).cuda()
            o = torch.full((s.size,), 0x5A, dtype=torch.uint8, device="cuda")
            jobs.append(batch.compress_job(xd, o, bits, B))
            want = s
            keep.append(xd)
        elif kind == 1:
            sd = torch.from_numpy(s).cuda()
            o = torch.full((n,), float("nan"), dtype=torch.float64 if f64 else torch.float32, device="cuda")
            jobs.append(batch.decompress_job(sd, o, bits, B))
            want = port.decompress(s, n, npdt, bits, B).view(np.uint8)
            keep.append(sd)
        else:
            if n == 0:
                continue
            sd = torch.from_numpy(s).cuda()
            idx = rng.integers(0, n, int(rng.integers(1, 300))).astype(np.int64)
            idd = torch.from_numpy(idx).cuda()
            o = torch.full((idx.size,), float("nan"), dtype=torch.float64 if f64 else torch.float32, device="cuda")
            jobs.append(batch.gather_job(sd, n, o, idd, bits, B))
            want = port.decompress(s, n, npdt, bits, B)[idx].view(np.uint8)
            keep += [sd, idd]
        outs.append(o)
        wants.append(want)
    plan = batch.BatchPlan(jobs)
    for replay in range(2):
        plan.launch()
        torch.cuda.synchronize()
        for i, (o, w) in enumerate(zip(outs, wants)):
            got = o.view(torch.uint8).cpu().numpy()
            assert np.array_equal(got, w), f"replay {replay} job {i}"
            o.view(torch.uint8).fill_(0x5A)
        torch.cuda.synchronize()
