"""FRSZ2 on one MI355X: compress + decompress throughput of the plain kernels.

Default workload: a float64 Gaussian random field of 512^3 values (1 GiB,
power spectrum k^-3, offset 1e3) at bits = 32, B = 32.  One step = one
compress launch + one decompress launch over the resident array; inputs are
in HBM before the timed region starts.  Throughput counts algorithmic bytes:
compress reads n*8 and writes the stream, decompress the reverse.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--side 512]
                    [--no-verify] [--no-cpu]

Prints one JSON line.  ``roofline`` is for the dominant kernel (timed with
HIP events on the stream the kernels run on).  The CPU leg (``cpu_leg``) is
the only code here that touches the oracle: it checks the GPU outputs of the
field, the small case and the edge-case gather against it (``bit_exact``),
and ``cpu_baseline`` times the reference CPU codec (oracle/_ref, built from
the reference sources) on one pinned core, best of 5: the whole field, the
whole small case, every edge-case cell on a 2^24-value sample and the 10^7
random-access decodes, single-threaded like the reference.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Hardware queues for this process, read when HIP initialises.  The serving
# runtime launches on 4 streams of its own, the host-memory pipeline keeps 3,
# and torch has its default stream; with HIP's default of 4 queues, streams
# share queues, and a small request's kernel can then wait in a queue behind
# a large one.  8 gives every stream here a queue of its own (serving, 10 s
# at 20k req/s: p50 0.027 -> 0.023 ms, p99 of requests up to 1 MiB 0.076 -> 0.068 ms).
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ramp-seconds", type=float, default=0.2,
                    help="untimed work before the warmup steps, so the device runs at its sustained clocks")
    ap.add_argument("--side", type=int, default=512)
    ap.add_argument("--bits", type=int, default=32)
    ap.add_argument("--block", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-arrays", type=int, default=4096)
    ap.add_argument("--batched-steps", type=int, default=5)
    ap.add_argument("--no-copy", action="store_true", help="skip the 1 GiB device-copy calibration")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-edge", action="store_true")
    ap.add_argument("--no-generic", action="store_true", help="skip the non-power-of-two block sizes")
    ap.add_argument("--no-blocks", action="store_true",
                    help="skip the power-of-two block sizes beyond the edge sweep's (2-8, 128-1024)")
    ap.add_argument("--no-host", action="store_true")
    ap.add_argument("--no-small", action="store_true", help="skip the float32 2^20-value case")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=20000.0)
    ap.add_argument("--serve-budget-gib", type=int, default=64)
    ap.add_argument("--serve-load-rates", default="100000,200000",
                    help="comma-separated request rates of the short load runs after the serving workload "
                         "('' for none)")
    ap.add_argument("--serve-load-seconds", type=float, default=3.0)
    return ap.parse_args()


def bench_batched(count, steps, seed=4, target_input_bytes=None):
    """4096 arrays, lengths log-uniform over [1e3, 1.6e7], 60% float32 /
    40% float64, bits from {16, 32}, B = 32: one autotuned graph launch
    vs a C++ loop of plain per-array launches, per direction.
    With target_input_bytes every length is divided by one factor so the
    input totals that many bytes (the workload's stated "about 4 GiB")."""
    import numpy as np
    import torch

    import frsz_amd
    from frsz_amd import batch
    from frsz_amd.workloads import batched_plan_arrays, smooth_signal_torch

    lengths, is_f64, bits, scale = batched_plan_arrays(count, seed=seed, target_input_bytes=target_input_bytes)
    B = 32
    xs, streams, outs = [], [], []
    alg = 0
    for i, (n, f64, b) in enumerate(zip(lengths.tolist(), is_f64.tolist(), bits.tolist())):
        dt = torch.float64 if f64 else torch.float32
        x = smooth_signal_torch(n, seed=10_000 + i, dtype=dt)
        size = frsz_amd.compressed_size(dt, b, B, n)
        xs.append(x)
        streams.append(torch.empty(size, dtype=torch.uint8, device="cuda"))
        outs.append(torch.empty_like(x))
        alg += n * x.element_size() + size
    torch.cuda.synchronize()
    cjobs = [batch.compress_job(x, s, b, B) for x, s, b in zip(xs, streams, bits.tolist())]
    djobs = [batch.decompress_job(s, o, b, B) for s, o, b in zip(streams, outs, bits.tolist())]

    def timed(plan, k):
        plan.launch()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(k):
            plan.launch()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k * 1e-3

    res = {"arrays": count, "total_values": int(lengths.sum()), "float64_share": round(float(is_f64.mean()), 3),
           "algorithmic_bytes_per_direction": int(alg), "block": B,
           "length_scale": round(scale, 3), "min_len": int(lengths.min()), "max_len": int(lengths.max())}
    # plain per-array launches from a C++ loop (the baseline), and the same captured in a graph
    plain_c = batch.BatchPlan(cjobs, {"plain_launches": 1, "use_graph": 0})
    plain_d = batch.BatchPlan(djobs, {"plain_launches": 1, "use_graph": 0})
    t_pc = timed(plain_c, steps)
    ref_streams = [s.clone() for s in streams]
    t_pd = timed(plain_d, steps)
    plain_cg = batch.BatchPlan(cjobs, {"plain_launches": 1, "use_graph": 1})
    t_pcg = timed(plain_cg, steps)
    plan_c = batch.BatchPlan(cjobs)
    tuned_c = plan_c.autotune(reps=2)
    t_bc = timed(plan_c, steps)
    same_c = all(torch.equal(a, b) for a, b in zip(streams, ref_streams))
    del ref_streams
    # decompress: plain outputs of each array compared with the batch's
    plan_d = batch.BatchPlan(djobs)
    tuned_d = plan_d.autotune(reps=2)
    t_bd = timed(plan_d, steps)
    same_d = True
    for s, o, x, b in zip(streams, outs, xs, bits.tolist()):
        r = frsz_amd.decompress(s, x.numel(), x.dtype, b, B)
        same_d &= torch.equal(r.view(torch.uint8), o.view(torch.uint8))
    torch.cuda.synchronize()
    res.update({
        "batched_compress_ms": round(t_bc * 1e3, 3), "plain_compress_ms": round(t_pc * 1e3, 3),
        "plain_graph_compress_ms": round(t_pcg * 1e3, 3),
        "batched_decompress_ms": round(t_bd * 1e3, 3), "plain_decompress_ms": round(t_pd * 1e3, 3),
        "batched_compress_gbs": round(alg / t_bc / 1e9, 1), "batched_decompress_gbs": round(alg / t_bd / 1e9, 1),
        "plain_compress_gbs": round(alg / t_pc / 1e9, 1), "plain_decompress_gbs": round(alg / t_pd / 1e9, 1),
        "speedup_compress": round(t_pc / t_bc, 2), "speedup_decompress": round(t_pd / t_bd, 2),
        "bit_exact_vs_plain": bool(same_c and same_d),
        "autotuned": {"compress": tuned_c, "decompress": tuned_d},
        # the most any batched launch could gain over this plain loop: the
        # loop's time over the time of these bytes at the HBM peak
        "speedup_ceiling_at_hbm_peak": {"compress": round(t_pc / (alg / (HBM_PEAK_GBS * 1e9)), 2),
                                        "decompress": round(t_pd / (alg / (HBM_PEAK_GBS * 1e9)), 2)},
        "bound_note": "the plain loop moves most bytes at full bandwidth on its large arrays; the batch removes "
                      "the per-launch dispatch gaps, so the speedup cannot exceed speedup_ceiling_at_hbm_peak",
    })
    for p_ in (plain_c, plain_d, plain_cg, plan_c, plan_d):
        p_.close()
    del xs, streams, outs
    torch.cuda.empty_cache()
    return res


def bench_edge(n=100_000_037, gathers=10_000_000, reps=5):
    """float32 lognormal magnitudes (exponents over about +-30), 10% all-zero
    64-value runs, 1% subnormals, random signs, n not a multiple of any block
    size: compress / decompress GB/s for bits {8,12,16,21,32} x B {16,32,64}
    (best of 3 rounds of `reps` launches), and a uniform random-access decode
    of 10^7 single values (bits 16, B 32).  Every output is compared with the
    first launch's (determinism).  Returns the result and, for the CPU leg,
    the host values, the bits-16 B-32 stream, the indices and the GPU's
    gathered values."""
    import numpy as np
    import torch

    import frsz_amd
    from frsz_amd.workloads import lognormal_edge

    x_host = lognormal_edge(n, seed=3)
    x = torch.from_numpy(x_host).cuda()
    out = {"n": n, "cases": []}
    host = {"x": x_host}

    def timed(fn):
        # best of 3 rounds of `reps` launches, after 2 untimed ones
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        best = None
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            t = e0.elapsed_time(e1) / reps * 1e-3
            best = t if best is None else min(best, t)
        return best

    y = torch.empty_like(x)
    for bits in (8, 12, 16, 21, 32):
        for B in (16, 32, 64):
            size = frsz_amd.compressed_size(torch.float32, bits, B, n)
            s = torch.empty(size, dtype=torch.uint8, device="cuda")
            frsz_amd.compress_into(x, s, bits, B)
            first = s.clone()
            tc = timed(lambda: frsz_amd.compress_into(x, s, bits, B))
            same = torch.equal(s, first)
            td = timed(lambda: frsz_amd.decompress_into(s, y, bits, B))
            alg = n * 4 + size
            out["cases"].append({"bits": bits, "block": B, "stream_bytes": size,
                                 "compress_gbs": round(alg / tc / 1e9, 1), "decompress_gbs": round(alg / td / 1e9, 1),
                                 "deterministic": bool(same)})
            if (bits, B) == (16, 32):
                g_host = np.random.default_rng(7).integers(0, n, gathers).astype(np.int64)
                g = torch.from_numpy(g_host).cuda()
                go = torch.empty(gathers, dtype=torch.float32, device="cuda")
                from frsz_amd import _capi

                def gat():
                    _capi.check(_capi.lib.frsz2_gather(0, bits, B, s.data_ptr(), s.numel(), n, g.data_ptr(),
                                                       gathers, go.data_ptr(), torch.cuda.current_stream().cuda_stream))

                tg = timed(gat)
                ref = y[g]
                U = 16
                out["gather"] = {"indices": gathers, "bits": bits, "block": B, "us": round(tg * 1e6, 1),
                                 "melem_s": round(gathers / tg / 1e6, 1),
                                 "algorithmic_gbs": round(gathers * (8 + 4 + U // 8 + U // 8) / tg / 1e9, 1),
                                 "bound": "transaction (random 64-byte touches), not bandwidth",
                                 "matches_decompress": bool(torch.equal(go.view(torch.int32), ref.view(torch.int32)))}
                host.update(stream=s.cpu().numpy(), idx=g_host.view(np.uint64), gathered=go.cpu().numpy(),
                            bits=bits, block=B)
            del s, first
    del x, y
    torch.cuda.empty_cache()
    return out, host


def bench_blocks(n=100_000_037, reps=5):
    """The power-of-two block sizes the plugin takes beyond the edge sweep's
    16-64 (2, 4, 8, 128, 256, 512, 1024), on the edge-case data (float32,
    n = 100,000,037; float64 on its first 5*10^7 values widened): compress /
    decompress TB/s of algorithmic bytes, every output compared with the
    first launch's."""
    import torch

    import frsz_amd
    from frsz_amd.workloads import lognormal_edge

    x32 = torch.from_numpy(lognormal_edge(n, seed=3)).cuda()
    x64 = x32[: 50_000_000].double()
    out = {"n_float32": n, "n_float64": x64.numel(), "cases": []}
    for x, widths in ((x32, (8, 12, 16, 21, 32)), (x64, (12, 21, 32, 40, 64))):
        y = torch.empty_like(x)
        for B in (2, 4, 8, 128, 256, 512, 1024):
            for bits in widths:
                size = frsz_amd.compressed_size(x.dtype, bits, B, x.numel())
                s = torch.empty(size, dtype=torch.uint8, device="cuda")
                frsz_amd.compress_into(x, s, bits, B)
                first = s.clone()
                tc = _best_time(lambda: frsz_amd.compress_into(x, s, bits, B), reps)
                same = torch.equal(s, first)
                frsz_amd.decompress_into(s, y, bits, B)
                y0 = y.clone()
                td = _best_time(lambda: frsz_amd.decompress_into(s, y, bits, B), reps)
                same = same and torch.equal(y.view(torch.uint8), y0.view(torch.uint8))
                alg = x.numel() * x.element_size() + size
                out["cases"].append({"dtype": str(x.dtype).replace("torch.", ""), "bits": bits, "block": B,
                                     "compress_tbs": round(alg / tc / 1e12, 3),
                                     "decompress_tbs": round(alg / td / 1e12, 3), "deterministic": bool(same)})
                del s, first, y0
        del y
    del x32, x64
    torch.cuda.empty_cache()
    return out


def _best_time(fn, reps):
    """Seconds per call of fn: best of 3 rounds of `reps` calls, after 2
    untimed ones, timed with events on torch's current stream (the stream
    the codec launches on)."""
    import torch

    for _ in range(2):
        fn()
    torch.cuda.synchronize()
    best = None
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) / reps * 1e-3
        best = t if best is None else min(best, t)
    return best


def bench_generic(n=100_000_037, reps=5):
    """Block sizes that are not powers of two (the generic tiles) and
    streams whose payload is off 16-byte alignment, on the edge-case data
    (float32, n = 100,000,037; float64 cases on its first 5*10^7 values
    widened): compress / decompress TB/s of algorithmic bytes (values +
    stream).  Every output is compared with the first launch's."""
    import torch

    import frsz_amd
    from frsz_amd.workloads import lognormal_edge

    x32 = torch.from_numpy(lognormal_edge(n, seed=3)).cuda()
    x64 = x32[: 50_000_000].double()
    out = {"n_float32": n, "n_float64": x64.numel(), "cases": []}

    def run(x, bits, B, offset=0):
        n_ = x.numel()
        size = frsz_amd.compressed_size(x.dtype, bits, B, n_)
        buf = torch.zeros(size + offset, dtype=torch.uint8, device="cuda")
        s = buf[offset:]
        y = torch.empty_like(x)
        frsz_amd.compress_into(x, s, bits, B)
        first = s.clone()
        tc = _best_time(lambda: frsz_amd.compress_into(x, s, bits, B), reps)
        same = torch.equal(s, first)
        frsz_amd.decompress_into(s, y, bits, B)
        y0 = y.clone()
        td = _best_time(lambda: frsz_amd.decompress_into(s, y, bits, B), reps)
        same = same and torch.equal(y.view(torch.uint8), y0.view(torch.uint8))
        alg = n_ * x.element_size() + size
        out["cases"].append({"dtype": str(x.dtype).replace("torch.", ""), "bits": bits, "block": B,
                             "stream_offset": offset, "compress_tbs": round(alg / tc / 1e12, 3),
                             "decompress_tbs": round(alg / td / 1e12, 3), "deterministic": bool(same)})
        del buf, s, y, first, y0

    for B in (3, 5, 7, 9, 12, 100):
        for bits in (8, 12, 16, 21, 32):
            run(x32, bits, B)
    for B in (5, 100):
        for bits in (12, 21, 40):
            run(x64, bits, B)
    # payload 8 bytes off 16-byte alignment: power-of-two B on the generic tiles
    for bits in (12, 21, 32):
        run(x32, bits, 32, offset=8)
    del x32, x64
    torch.cuda.empty_cache()
    return out


def bench_small(reps=200, bits=16, block=32):
    """The reference's own CPU-sized case: float32 x_i = sin(i/500) +
    0.3 sin(i/37) + N(0, 1e-3), 2^20 values (4 MiB), bits 16, B 32.  At this
    size a launch is bound by dispatch, not bandwidth: reported are the
    average over `reps` back-to-back launches from Python, the same launches
    replayed from a captured graph (20 per graph, no Python between them;
    every replay's output is checked against the eager one), the median
    single-launch time (synchronised before and after, event-timed) and the
    host-memory call (pageable in, pageable out, wall clock), the plugin's
    meaning.  Returns
    the result and (values, GPU stream, GPU decode) for the checker."""
    import numpy as np
    import torch

    import frsz_amd
    from frsz_amd import _capi
    from frsz_amd.workloads import smooth_signal

    n = 1 << 20
    x_host = smooth_signal(n, seed=1, dtype=np.float32)
    x = torch.from_numpy(x_host).cuda()
    size = frsz_amd.compressed_size(torch.float32, bits, block, n)
    stream = torch.empty(size, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    alg = n * 4 + size

    def comp():
        frsz_amd.compress_into(x, stream, bits, block)

    def decomp():
        frsz_amd.decompress_into(stream, y, bits, block)

    res = {"workload": f"float32 smooth signal, {n} values, bits {bits}, B {block}", "n": n, "stream_bytes": size}
    for name, fn in (("compress", comp), ("decompress", decomp)):
        for _ in range(20):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        t_b2b = e0.elapsed_time(e1) / reps * 1e-3
        singles = []
        for _ in range(50):
            torch.cuda.synchronize()
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            singles.append(e0.elapsed_time(e1) * 1e-3)
        t_one = float(np.median(singles))
        res[f"{name}_us_back_to_back"] = round(t_b2b * 1e6, 2)
        res[f"{name}_us_single"] = round(t_one * 1e6, 2)
        res[f"{name}_gbs_back_to_back"] = round(alg / t_b2b / 1e9, 1)
        # graph replays: the launches without Python dispatch between them
        out = stream if name == "compress" else y
        fn()
        torch.cuda.synchronize()
        want = out.clone()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            for _ in range(20):
                fn()
        same = True
        best = 1e9
        for _ in range(5):
            out.zero_()
            torch.cuda.synchronize()
            e0.record()
            graph.replay()
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1) * 1e-3 / 20)
            same &= bool(torch.equal(out, want))
        del graph
        res[f"{name}_us_graph"] = round(best * 1e6, 2)
        res[f"{name}_gbs_graph"] = round(alg / best / 1e9, 1)
        res[f"{name}_graph_replays_exact"] = same
    comp()
    decomp()
    torch.cuda.synchronize()
    s_host = np.empty(size, np.uint8)
    y_host = np.empty_like(x_host)
    for name, call in (("compress", lambda: _capi.lib.frsz2_compress_host(0, bits, block, x_host.ctypes.data, n,
                                                                          s_host.ctypes.data, size)),
                       ("decompress", lambda: _capi.lib.frsz2_decompress_host(0, bits, block, s_host.ctypes.data,
                                                                              size, n, y_host.ctypes.data))):
        ts = []
        for _ in range(21):
            t0 = time.perf_counter()
            _capi.check(call())
            ts.append(time.perf_counter() - t0)
        res[f"host_{name}_us"] = round(float(np.median(ts[1:])) * 1e6, 1)
    gpu_s, gpu_y = stream.cpu().numpy(), y.cpu().numpy()
    res["host_path_same_bytes"] = bool(np.array_equal(s_host, gpu_s) and
                                       np.array_equal(y_host.view(np.uint32), gpu_y.view(np.uint32)))
    return res, (x_host, gpu_s, gpu_y, bits, block)


def bench_host(x_host, bits, block, want_stream, want_y, reps=3):
    """Host-resident data (the plugin's host-in / host-out meaning): the same
    field from pageable host memory to a host stream and back through the
    pipelined frsz2_compress_host / frsz2_decompress_host, beside (a) the
    unpipelined sequence a caller would otherwise write (torch pageable H2D,
    device compress, D2H) and (b) the raw PCIe rates of pinned 1 GiB copies.
    Rates are algorithmic bytes (values + stream) per second, PCIe included."""
    import numpy as np
    import torch

    import frsz_amd
    from frsz_amd import _capi

    n = x_host.size
    size = frsz_amd.compressed_size(torch.float64, bits, block, n)
    alg = n * 8 + size
    s_host = np.empty(size, np.uint8)
    y_host = np.empty_like(x_host)

    def best(fn):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return min(ts)

    tc = best(lambda: _capi.check(_capi.lib.frsz2_compress_host(1, bits, block, x_host.ctypes.data, n,
                                                                 s_host.ctypes.data, size)))
    td = best(lambda: _capi.check(_capi.lib.frsz2_decompress_host(1, bits, block, s_host.ctypes.data, size, n,
                                                                   y_host.ctypes.data)))
    xt = torch.from_numpy(x_host)
    st = torch.from_numpy(s_host)

    def naive_c():
        xd = xt.cuda()
        sd = frsz_amd.compress(xd, bits, block)
        sd.cpu()
        torch.cuda.synchronize()

    def naive_d():
        sd = st.cuda()
        yd = frsz_amd.decompress(sd, n, torch.float64, bits, block)
        yd.cpu()
        torch.cuda.synchronize()

    tnc, tnd = best(naive_c), best(naive_d)
    pin = torch.empty(1 << 27, dtype=torch.float64, pin_memory=True)
    dev = torch.empty(1 << 27, dtype=torch.float64, device="cuda")

    def h2d():
        dev.copy_(pin, non_blocking=True)
        torch.cuda.synchronize()

    def d2h():
        pin.copy_(dev, non_blocking=True)
        torch.cuda.synchronize()

    th, tdh = best(h2d), best(d2h)
    same = bool(np.array_equal(s_host, want_stream) and np.array_equal(y_host.view(np.uint64), want_y.view(np.uint64)))
    del pin, dev
    torch.cuda.empty_cache()
    return {"workload": f"float64 field {n} values, bits {bits}, B {block}, pageable host memory",
            "compress_gbs": round(alg / tc / 1e9, 2), "decompress_gbs": round(alg / td / 1e9, 2),
            "unpipelined_compress_gbs": round(alg / tnc / 1e9, 2),
            "unpipelined_decompress_gbs": round(alg / tnd / 1e9, 2),
            "pcie_pinned_h2d_gbs": round((1 << 30) / th / 1e9, 2),
            "pcie_pinned_d2h_gbs": round((1 << 30) / tdh / 1e9, 2),
            "bit_exact_vs_device_path": same}


def bench_serve(seconds, rate, budget_gib, seed=5, verify_every=499, server_options=None, source_gib=4):
    """Poisson arrivals at `rate` requests/s for `seconds`: 50% compress, 30%
    full decompress, 20% random-access reads of 1-4096 values from resident
    streams; sizes log-uniform 4 KiB-256 MiB, float32 / float64 (half each),
    bits from {16, 32}, B = 32; HBM budget `budget_gib` GiB.  Compress
    requests read slices of two `source_gib` GiB source buffers (one per
    dtype), far larger than the on-die caches.  Open-loop replay through the
    serving runtime; latency = completion - scheduled arrival.  Every
    `verify_every`-th request keeps its output, which is then compared with
    the same request run alone through the single-array API."""
    import numpy as np
    import torch

    import frsz_amd
    from frsz_amd import serve
    from frsz_amd.workloads import serving_requests, serving_trace, smooth_signal_torch

    src_values = {False: (source_gib << 30) // 4, True: (source_gib << 30) // 8}
    tr = serving_trace(seconds, rate, src_values, seed=seed)
    kind, f64, bits, n, offs, source, m, arrival = (tr[k] for k in ("kind", "f64", "bits", "n", "offset", "source",
                                                                  "m", "arrival"))
    nreq = kind.size
    B = 32
    src = {False: smooth_signal_torch(src_values[False], seed=77, dtype=torch.float32),
           True: smooth_signal_torch(src_values[True], seed=78, dtype=torch.float64)}
    # kept outputs stay in the pool until the replay ends: at most about 2400
    # of them (about 34 GiB at the mix's mean size), so that they cannot
    # fill the budget and stall admission on longer or faster runs
    verify_every = max(verify_every, -(-nreq // 2400))
    keep = (np.arange(nreq) % verify_every) == 0
    trace, idx_all, idx_off = serving_requests(tr, src, B, keep)
    torch.cuda.synchronize()
    print(f"bench: serving {nreq} requests over {seconds:.0f} s", file=sys.stderr, flush=True)
    opts = dict(slots=4)
    opts.update(server_options or {})
    srv = serve.Server(hbm_budget=budget_gib << 30, **opts)
    res, tickets = srv.replay(trace)
    stats = srv.stats()
    lat = (res["t_done"] - res["t_arrival"]) * 1e3
    ok = bool(np.all(res["status"] == 0))
    vs = np.where(f64, 8, 4)
    makespan = float(res["t_done"].max() - 0.0)
    # bit-exact check of the kept sample against the request run alone
    print(f"bench: checking {int(keep.sum())} kept outputs", file=sys.stderr, flush=True)
    same = True
    checked = 0
    for i in np.nonzero(keep)[0]:
        dt = torch.float64 if f64[i] else torch.float32
        r = {k: res[i][k] for k in res.dtype.names}
        if kind[i] == 0:
            x = src[bool(f64[i])][offs[i]: offs[i] + n[i]]
            want = frsz_amd.compress(x, int(bits[i]), B)
            got = srv.output(r, torch.uint8, want.numel())
        else:
            s = int(source[i])
            x = src[bool(f64[s])][offs[s]: offs[s] + n[s]]
            st = frsz_amd.compress(x, int(bits[i]), B)
            if kind[i] == 1:
                want = frsz_amd.decompress(st, int(n[i]), dt, int(bits[i]), B)
                got = srv.output(r, dt, int(n[i]))
            else:
                idx = idx_all[idx_off[i]: idx_off[i] + m[i]]
                want = frsz_amd.gather(st, int(n[i]), dt, int(bits[i]), B, idx)
                got = srv.output(r, dt, int(m[i]))
        same &= torch.equal(got.view(torch.uint8), want.view(torch.uint8))
        checked += 1
        srv.release(int(tickets[i]))
    srv.close()

    def pct(a, q):
        return round(float(np.percentile(a, q)), 4) if a.size else None

    names = {0: "compress", 1: "decompress", 2: "gather"}
    per = {names[k]: {"count": int((kind == k).sum()), "p50_ms": pct(lat[kind == k], 50),
                      "p99_ms": pct(lat[kind == k], 99)} for k in (0, 1, 2)}
    payload = float(np.sum(np.where(kind == 2, m * vs, n * vs)))
    # each request's own HBM bytes (values + stream; a gather: index, value,
    # exponent and code per element) and the time they take at peak
    # bandwidth: the floor below which no scheduling could bring its latency
    nb = (n + B - 1) // B
    U = bits  # bits 16 / 32 are stored unpacked
    X = nb * U // 8
    stream_bytes = X + 256 - X % 256 + nb * (B * bits // 8)
    own = np.where(kind == 2, m * (8 + vs + 2 * (U // 8)), n * vs + stream_bytes).astype(np.float64)
    floor_ms = own / 8e12 * 1e3
    small, large = own <= (1 << 20), own >= (64 << 20)
    per_size = {"floor_at_peak_p99_ms": pct(floor_ms, 99),
                "p99_minus_floor_ms": pct(lat - floor_ms, 99),
                "p99_requests_to_1mib_ms": pct(lat[small], 99), "requests_to_1mib": int(small.sum()),
                "p99_requests_from_64mib_ms": pct(lat[large], 99), "requests_from_64mib": int(large.sum()),
                # where the requests to 1 MiB spend their p99: queued -> admitted,
                # admitted -> launched, launched -> completion seen
                "stages_p99_requests_to_1mib_ms": {
                    "admit": pct(((res["t_admit"] - res["t_arrival"]) * 1e3)[small], 99),
                    "launch": pct(((res["t_launch"] - res["t_admit"]) * 1e3)[small], 99),
                    "run": pct(((res["t_done"] - res["t_launch"]) * 1e3)[small], 99)}}
    del src, idx_all
    torch.cuda.empty_cache()
    return {"requests": nreq, "rate_offered": rate, "seconds": seconds, "hbm_budget_gib": budget_gib,
            "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "p999_ms": pct(lat, 99.9), "max_ms": pct(lat, 100),
            "per_op": per, "by_size": per_size,
            # median of each stage: queued -> admitted (output carved from the
            # pool), admitted -> launched, launched -> completion seen
            "stages_p50_ms": {"admit": pct((res["t_admit"] - res["t_arrival"]) * 1e3, 50),
                              "launch": pct((res["t_launch"] - res["t_admit"]) * 1e3, 50),
                              "run": pct((res["t_done"] - res["t_launch"]) * 1e3, 50)},
            "throughput_req_s": round(nreq / makespan, 1),
            "payload_gbs": round(payload / makespan / 1e9, 2), "all_ok": ok,
            "bit_exact_sample": {"checked": checked, "same": bool(same)},
            "launches": stats["launches"], "mean_requests_per_launch":
                round(stats["launched_requests"] / max(stats["launches"], 1), 2),
            "pool_peak_gib": round(stats["pool_peak"] / 2**30, 3), "max_queue": stats["max_queue"]}


def bench_ceiling(mib=1024):
    """The streaming ceilings of this device from tests/native/stream_ceiling
    (a child process; plain 16-byte streaming kernels with every read:write
    mix of the codec): the best GB/s per mix, or None when it is not built."""
    import subprocess

    exe = os.path.join(ROOT, "tests", "native", "stream_ceiling")
    if not os.path.exists(exe):
        return None
    res = subprocess.run([exe, "--mib", str(mib)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=300)
    if res.returncode != 0:
        return {"error": res.stdout[-500:]}
    best = {}
    for ln in res.stdout.splitlines():
        if ln.startswith("{"):
            r = json.loads(ln)
            key = {(1, 0): "read_only", (0, 1): "write_only", (1, 1): "copy_1to1", (2, 1): "read2_write1",
                   (1, 2): "read1_write2", (1, 4): "read1_write4",
                   (4, 1): "read4_write1"}[(r["read"], r["write"])]
            best[key] = max(best.get(key, 0.0), r["gbs"])
    return {"side_mib": mib, "gbs": best, "source": "tests/native/stream_ceiling.hip, best over workgroup size, "
            "vectors per lane, non-temporal hint and tile order"}


def bench_random_ceiling(indices=10_000_000, values=100_000_037):
    """The random-access ceiling of this device from tests/native/stream_ceiling
    --random (a child process): the gather's memory pattern (an 8-byte index,
    a random 2-byte code of a 2n-byte payload and its block's 2-byte exponent,
    a 4-byte store per index) without the decode.  Returns the per-variant
    times, or None when the program is not built."""
    import subprocess

    exe = os.path.join(ROOT, "tests", "native", "stream_ceiling")
    if not os.path.exists(exe):
        return None
    res = subprocess.run([exe, "--random", str(indices), "--values", str(values)], stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True, timeout=300)
    if res.returncode != 0:
        return {"error": res.stdout[-500:]}
    rows = [json.loads(ln) for ln in res.stdout.splitlines() if ln.startswith("{")]
    us = {}
    for r in rows:
        key = r["loads"] + (" (non-temporal code)" if r["nt_code"] else "")
        us[key] = r["us"]
    best = min(r["us"] for r in rows if r["loads"] == "code+exponent")
    return {"indices": indices, "values": values, "us": us, "best_code_exponent_us": best,
            "source": "tests/native/stream_ceiling.hip --random"}


def _best_of(fn, reps=5):
    """Wall seconds of fn: one untimed warm-up call, then the best of `reps`."""
    fn()
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        t = time.perf_counter() - t0
        best = t if best is None else min(best, t)
    return best


def cpu_leg(checks, x_host, bits, block, small_values, edge, time_cpu=True, edge_sample=1 << 24, reps=5):
    """The CPU leg, the one part of this file that touches the oracle.

    checks: {name: (values, GPU stream, GPU decode, bits, block)}; each is
    compared byte-for-byte with the oracle (the multithreaded C restatement),
    giving {name: bool}; with `edge` (bench_edge's host data) the GPU's 10^7
    gathered values are compared with the oracle's gather as well.

    With time_cpu, the reference CPU codec (the restatement when the
    reference build is absent) runs on one thread pinned to one core (the
    calling thread's affinity narrowed to its lowest allowed CPU, restored
    afterwards), each workload after one untimed warm-up as the best of
    `reps`, into buffers allocated beforehand:
      * the float32 2^20-value smooth signal, bits 16, B 32;
      * the whole float64 512^3 field at (bits, block): the cpu_baseline
        value is its compress+decompress GB/s;
      * each (bits, B) cell of the edge-case sweep on its first
        `edge_sample` values, and the 10^7-index random-access decode
        (element decode) over the whole 10^8-value stream;
    and then the block-parallel restatement on every allowed core over the
    field.  Returns (verdicts, cpu_baseline or None)."""
    import numpy as np

    from oracle.oracle import Port, Reference, compressed_bytes, reference_available

    port = Port().with_threads(0)
    verdicts = {}
    for name, (vals, got_s, got_y, b, blk) in checks.items():
        want = port.compress(vals, b, blk)
        ui = np.uint32 if vals.dtype == np.float32 else np.uint64
        verdicts[name] = bool(np.array_equal(got_s, want) and
                              np.array_equal(got_y.view(ui), port.decompress(want, vals.size, vals.dtype, b,
                                                                              blk).view(ui)))
    if edge is not None and "stream" in edge:
        ex = edge["x"]
        want = port.gather(edge["stream"], ex.size, ex.dtype, edge["bits"], edge["block"], edge["idx"])
        verdicts["edge_gather"] = bool(np.array_equal(port.compress(ex, edge["bits"], edge["block"]), edge["stream"])
                                       and np.array_equal(want.view(np.uint32), edge["gathered"].view(np.uint32)))
    if not time_cpu:
        return verdicts, None
    kind = "reference" if reference_available() else "port"
    codec = Reference() if kind == "reference" else Port()

    def directions(vals, b, blk):
        # (compress seconds, decompress seconds, algorithmic bytes) of one array
        stream = np.zeros(compressed_bytes(b, blk, vals.size), dtype=np.uint8)
        back = np.zeros_like(vals)
        tc = _best_of(lambda: codec.compress_into(vals, stream, b, blk), reps)
        td = _best_of(lambda: codec.decompress_into(stream, back, b, blk), reps)
        return tc, td, vals.nbytes + stream.size

    allowed = os.sched_getaffinity(0)
    core = min(allowed)
    os.sched_setaffinity(0, {core})
    try:
        res = {"unit": "GB/s", "cores": 1, "kind": kind, "pinned_cpu": core}
        if small_values is not None:
            tc, td, nb = directions(small_values, 16, 32)
            res["small_case"] = {"values": int(small_values.size), "bits": 16, "block": 32,
                                 "compress_us": round(tc * 1e6, 1), "decompress_us": round(td * 1e6, 1),
                                 "compress_gbs": round(nb / tc / 1e9, 3), "decompress_gbs": round(nb / td / 1e9, 3)}
        xf = np.ascontiguousarray(x_host)
        tc, td, nb_field = directions(xf, bits, block)
        res.update(value=round(2 * nb_field / (tc + td) / 1e9, 4), compress_gbs=round(nb_field / tc / 1e9, 4),
                   decompress_gbs=round(nb_field / td / 1e9, 4), field_compress_s=round(tc, 4),
                   field_decompress_s=round(td, 4))
        res["sample"] = (f"the whole float64 {xf.size}-value field ({xf.nbytes / 2**30:.2f} GiB), bits {bits}, "
                         f"B {block}: compress+decompress, warm-up then best of {reps}, one thread pinned to CPU "
                         f"{core}; edge cells on the first {edge_sample} of 10^8 values")
        if edge is not None:
            ex = np.ascontiguousarray(edge["x"][:edge_sample])
            cells = []
            for b in (8, 12, 16, 21, 32):
                for blk in (16, 32, 64):
                    tc, td, nb = directions(ex, b, blk)
                    cells.append({"bits": b, "block": blk, "compress_gbs": round(nb / tc / 1e9, 3),
                                  "decompress_gbs": round(nb / td / 1e9, 3)})
            res["edge_cells"] = {"values": int(ex.size), "cells": cells}
            if "stream" in edge:
                ex = edge["x"]
                tg = _best_of(lambda: codec.gather(edge["stream"], ex.size, ex.dtype, edge["bits"], edge["block"],
                                                   edge["idx"]), reps)
                res["edge_gather"] = {"indices": int(edge["idx"].size), "values": int(ex.size),
                                      "bits": edge["bits"], "block": edge["block"], "ms": round(tg * 1e3, 1),
                                      "melem_s": round(edge["idx"].size / tg / 1e6, 2)}
    finally:
        os.sched_setaffinity(0, allowed)
    # the block-parallel restatement on every allowed core (same bytes)
    stream = np.zeros(compressed_bytes(bits, block, xf.size), dtype=np.uint8)
    back = np.zeros_like(xf)
    tc = _best_of(lambda: port.compress_into(xf, stream, bits, block), 2)
    td = _best_of(lambda: port.decompress_into(stream, back, bits, block), 2)
    res["port_all_cores"] = {"threads": port.max_threads(), "value": round(2 * nb_field / (tc + td) / 1e9, 4),
                             "unit": "GB/s", "best_of": 2}
    res["cpu_model"] = cpu_model()
    res["nproc"] = os.cpu_count()
    res["allowed_cpus"] = len(allowed)
    return verdicts, res


def profiled_traffic(op, side, bits, block):
    """HBM bytes per launch of the plain kernel for `op`, from the committed
    FETCH_SIZE / WRITE_SIZE counter summary of this exact workload (None for
    any other configuration)."""
    if (side, bits, block) != (512, 32, 32):
        return None
    path = os.path.join(ROOT, "profiles", "plain_f64_512cubed_b32_B32_hbm_traffic.md")
    try:
        text = open(path).read()
        data = json.loads(text.split("```json", 1)[1].split("```", 1)[0])
        return int(data["kernels"][op]["traffic_bytes"])
    except (OSError, IndexError, KeyError, ValueError):
        return None


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported (the codec "
                                   "shards by exponent-block ranges, see the design notes)"}))
        sys.exit(2)
    import torch

    import frsz_amd
    from frsz_amd.workloads import gaussian_field_torch

    dev = torch.device("cuda:0")
    bits, block = args.bits, args.block
    side = args.side
    print(f"bench: generating float64 field {side}^3", file=sys.stderr, flush=True)
    x = gaussian_field_torch(side, seed=2, device=dev).reshape(-1).contiguous()
    n = x.numel()
    size = frsz_amd.compressed_size(torch.float64, bits, block, n)
    stream = torch.empty(size, dtype=torch.uint8, device=dev)
    y = torch.empty_like(x)
    torch.cuda.synchronize()

    def step(ev=None):
        # timed steps: one event after each launch (an event record costs
        # about 1% of a step), the previous step's last event opening the
        # next compress
        frsz_amd.compress_into(x, stream, bits, block)
        if ev is not None:
            ev[0].record()
        frsz_amd.decompress_into(stream, y, bits, block)
        if ev is not None:
            ev[1].record()

    # bring the device to its sustained clocks (about 0.2 s of the same work,
    # untimed), then the W warmup steps
    t_ramp = time.perf_counter()
    while time.perf_counter() - t_ramp < args.ramp_seconds:
        for _ in range(20):
            step()
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e_open = torch.cuda.Event(enable_timing=True)
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e_open.record()
    for i in range(args.steps):
        step(evs[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    starts = [e_open] + [e[1] for e in evs[:-1]]
    tc = [a.elapsed_time(e[0]) * 1e-3 for a, e in zip(starts, evs)]
    td = [e[0].elapsed_time(e[1]) * 1e-3 for e in evs]
    alg = n * 8 + size  # algorithmic bytes per launch, each direction
    comp_gbs = alg / (sum(tc) / len(tc)) / 1e9
    decomp_gbs = alg / (sum(td) / len(td)) / 1e9
    ms_per_step = wall / args.steps * 1e3
    value = 2 * alg / (wall / args.steps) / 1e9

    x_host = s_host = y_host = None
    if not args.no_verify or not args.no_cpu or not args.no_host:
        x_host, s_host, y_host = x.cpu().numpy(), stream.cpu().numpy(), y.cpu().numpy()

    dom = "compress" if sum(tc) >= sum(td) else "decompress"
    dom_t = (sum(tc) if dom == "compress" else sum(td)) / args.steps
    achieved = alg / dom_t / 1e9
    line = {
        "metric": "frsz2 comp/decomp GB/s + %HBM peak, batched vs plain (bit-exact); serve p99 ms",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Gaussian random field k^-3 + 1e3, seeded, generated on the GPU)",
        "config": {"workload": f"plain single array: float64 {side}^3 ({n * 8 / 2**30:.2f} GiB), bits={bits}, "
                               f"B={block}, one compress + one decompress launch per step",
                   "n": n, "bits": bits, "block": block, "stream_bytes": size},
        "compress_gbs": round(comp_gbs, 2),
        "decompress_gbs": round(decomp_gbs, 2),
        "compress_pct_hbm": round(100 * comp_gbs / HBM_PEAK_GBS, 2),
        "decompress_pct_hbm": round(100 * decomp_gbs / HBM_PEAK_GBS, 2),
        "compress_us": round(sum(tc) / len(tc) * 1e6, 2),
        "decompress_us": round(sum(td) / len(td) * 1e6, 2),
        "bit_exact": None,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(dom, side, bits, block), "algorithmic_bytes": alg,
                     "traffic_source": "not measured in this run: HBM bytes per launch of this kernel on this "
                                       "workload from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary "
                                       "profiles/plain_f64_512cubed_b32_B32_hbm_traffic.md (x2.0 / x1.0)"},
    }
    if not args.no_copy:
        # a plain 1 GiB device copy (torch), the calibration kernel for the
        # HBM byte counters and a practical bandwidth reference
        a = torch.empty(1 << 27, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        c0.record()
        for _ in range(5):
            b.copy_(a)
        c1.record()
        torch.cuda.synchronize()
        tcopy = c0.elapsed_time(c1) / 5 * 1e-3
        line["copy_calibration"] = {"bytes_read": a.numel() * 8, "bytes_written": a.numel() * 8,
                                    "us": round(tcopy * 1e6, 1), "gbs": round(2 * a.numel() * 8 / tcopy / 1e9, 1)}
        del a, b
    if not args.no_copy:
        print("bench: streaming ceilings", file=sys.stderr, flush=True)
        ceil = bench_ceiling()
        if ceil is not None:
            line["stream_ceiling"] = ceil
            g = ceil.get("gbs", {})
            # float64 at 32 bits: compress reads two bytes per byte written, decompress the reverse
            mix = {"compress": "read2_write1", "decompress": "read1_write2"}[dom] if (bits, block) == (32, 32) else None
            if mix in g:
                line["roofline"]["ceiling_mix"] = mix
                line["roofline"]["ceiling_gbs"] = g[mix]
                line["roofline"]["frac_of_ceiling"] = round(achieved / g[mix], 4)
    if not args.no_host:
        print("bench: host-resident path", file=sys.stderr, flush=True)
        line["host_path"] = bench_host(x_host, bits, block, s_host, y_host)
    del x, y, stream
    torch.cuda.empty_cache()
    small_check = None
    edge_host = None
    if not args.no_small:
        print("bench: small case", file=sys.stderr, flush=True)
        line["small_case"], small_check = bench_small()
    if not args.no_batched:
        print("bench: batched workload", file=sys.stderr, flush=True)
        line["batched"] = bench_batched(args.batched_arrays, args.batched_steps)
        print("bench: batched workload scaled to 4 GiB of input", file=sys.stderr, flush=True)
        line["batched_4gib"] = bench_batched(args.batched_arrays, args.batched_steps, target_input_bytes=1 << 32)
    if not args.no_edge:
        print("bench: edge-case sweep", file=sys.stderr, flush=True)
        line["edge"], edge_host = bench_edge()
        if not args.no_copy:
            print("bench: random-access ceiling", file=sys.stderr, flush=True)
            rc = bench_random_ceiling(line["edge"]["gather"]["indices"], line["edge"]["n"])
            if rc is not None:
                line["edge"]["gather"]["random_access_ceiling"] = rc
                if "best_code_exponent_us" in rc:
                    line["edge"]["gather"]["frac_of_random_ceiling"] = round(
                        rc["best_code_exponent_us"] / line["edge"]["gather"]["us"], 4)
    if not args.no_blocks:
        print("bench: power-of-two block sizes 2-8 and 128-1024", file=sys.stderr, flush=True)
        line["blocks"] = bench_blocks()
    if not args.no_generic:
        print("bench: block sizes that are not powers of two, unaligned streams", file=sys.stderr, flush=True)
        line["generic"] = bench_generic()
    if not args.no_serve:
        line["serve"] = bench_serve(args.serve_seconds, args.serve_rate, args.serve_budget_gib)
        # the same mix at higher offered rates: continuous batching under load,
        # up to the rate the device and the HBM budget sustain
        load = []
        for rate in [float(r) for r in args.serve_load_rates.split(",") if r.strip()]:
            r = bench_serve(args.serve_load_seconds, rate, args.serve_budget_gib)
            load.append({k: r[k] for k in ("rate_offered", "seconds", "requests", "p50_ms", "p99_ms", "p999_ms",
                                           "throughput_req_s", "payload_gbs", "all_ok", "bit_exact_sample",
                                           "mean_requests_per_launch", "pool_peak_gib", "max_queue",
                                           "stages_p50_ms", "by_size")})
        if load:
            line["serve_load"] = load
    if not args.no_verify or not args.no_cpu:
        print("bench: CPU leg (checker and CPU baseline)", file=sys.stderr, flush=True)
        checks = {}
        if not args.no_verify:
            checks["field"] = (x_host, s_host, y_host, bits, block)
            if small_check is not None:
                checks["small"] = small_check
        verdicts, cpu = cpu_leg(checks, x_host, bits, block, None if small_check is None else small_check[0],
                                edge_host, time_cpu=not args.no_cpu)
        if "field" in verdicts:
            line["bit_exact"] = verdicts["field"]
        if "small" in verdicts:
            line["small_case"]["bit_exact"] = verdicts["small"]
        if "edge_gather" in verdicts:
            line["edge"]["gather"]["bit_exact"] = verdicts["edge_gather"]
        if cpu is not None:
            line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
