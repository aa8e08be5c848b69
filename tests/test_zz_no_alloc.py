"""The allocation-free batched entry point (frsz2_batch_run_noalloc, called
through the C ABI) against the existing one (frsz2_batch_create +
frsz2_batch_launch): the same jobs give the same bytes, bit for bit, on two
calls in a row, the second with fewer and smaller jobs (the shared
descriptor scratch keeps the first call's contents).  The jobs mix every
direction, float32 / float64, packed and unpacked widths, more descriptors
of one class than one store kernel carries, and jobs the tiles cannot take
(B = 5, unaligned values)."""
import ctypes

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


def make_jobs(fz_mod, batch, seed, count, max_n, n_same_class):
    """(jobs for the allocation-free call, the same jobs over fresh outputs
    for the plan, the output pairs)."""
    from frsz_amd.workloads import smooth_signal_torch

    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        if i < n_same_class:
            dt, bits, B = torch.float32, 16, 32  # one class, many descriptors
        else:
            f64 = bool(rng.random() < 0.4)
            dt = torch.float64 if f64 else torch.float32
            bits = int(rng.choice([8, 12, 21, 32] + ([40, 64] if f64 else [])))
            B = int(rng.choice([16, 32, 64, 5]))
        n = int(rng.integers(1, max_n))
        x = smooth_signal_torch(n, seed=seed * 1000 + i, dtype=dt) * float(2.0 ** rng.integers(-8, 8))
        specs.append((x, bits, B))
    # a job whose values are off 16-byte alignment (a plain launch inside the batch)
    base = smooth_signal_torch(5001, seed=seed, dtype=torch.float32)
    specs.append((base[1:], 12, 32))

    sets = ([], [])
    pairs = []
    for x, bits, B in specs:
        size = fz_mod.compressed_size(x.dtype, bits, B, x.numel())
        src = fz_mod.compress(x, bits, B)
        idx = torch.from_numpy(rng.integers(0, x.numel(), int(rng.integers(1, 3000)))).cuda()
        outs = []
        for k in range(2):
            s = torch.full((size,), 0x5A + k, dtype=torch.uint8, device="cuda")
            y = torch.full((x.numel(),), float("nan"), dtype=x.dtype, device="cuda")
            g = torch.full((idx.numel(),), float("nan"), dtype=x.dtype, device="cuda")
            sets[k].append(batch.compress_job(x, s, bits, B))
            sets[k].append(batch.decompress_job(src, y, bits, B))
            sets[k].append(batch.gather_job(src, x.numel(), g, idx, bits, B))
            outs.append((s, y, g))
        pairs.append(outs)
    torch.cuda.synchronize()
    return sets, pairs


def run_and_compare(fz_mod, batch, seed, count, max_n, n_same_class):
    (noalloc_jobs, plan_jobs), pairs = make_jobs(fz_mod, batch, seed, count, max_n, n_same_class)
    batch.run_noalloc(noalloc_jobs)
    plan = batch.BatchPlan(plan_jobs)
    plan.launch()
    torch.cuda.synchronize()
    for i, ((s0, y0, g0), (s1, y1, g1)) in enumerate(pairs):
        assert torch.equal(s0, s1), f"seed {seed} job {i}: stream"
        assert torch.equal(y0.view(torch.uint8), y1.view(torch.uint8)), f"seed {seed} job {i}: decode"
        assert torch.equal(g0.view(torch.uint8), g1.view(torch.uint8)), f"seed {seed} job {i}: gather"
    plan.close()
    return noalloc_jobs, pairs


def test_noalloc_matches_batch_plan(fz, port):
    fz_mod, batch = fz
    print("allocation-free batch: call 1 (190 arrays, 571 jobs)", flush=True)
    run_and_compare(fz_mod, batch, seed=1, count=190, max_n=200_000, n_same_class=100)
    print("allocation-free batch: call 2 (9 arrays, 30 jobs, smaller)", flush=True)
    jobs, pairs = run_and_compare(fz_mod, batch, seed=2, count=9, max_n=20_000, n_same_class=2)
    # the second call's streams are also the oracle's
    from conftest import oracle_codec

    for (job, keep), outs in zip(jobs[0::3], pairs):
        x = keep[0]
        s, _ = oracle_codec(port, x, job.bits, job.block)
        assert torch.equal(outs[0][0], s)


def test_noalloc_rejects_too_many_jobs(fz):
    fz_mod, batch = fz
    from frsz_amd import _capi

    fake = 1 << 20
    job = batch.Job(0, 0, 16, 0, 32, 1000, fake, fake, 0, 1 << 20)
    arr = (batch.Job * (batch.NOALLOC_MAX_JOBS + 1))(*([job] * (batch.NOALLOC_MAX_JOBS + 1)))
    rc = _capi.lib.frsz2_batch_run_noalloc(arr, len(arr), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == _capi.ECAPACITY
    assert "at most 10922 tiled jobs" in _capi.last_error()
