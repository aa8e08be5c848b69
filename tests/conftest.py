import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running")


class GoldenCase:
    def __init__(self, name, dtype, bits, B, x, stream, y, gidx, gval):
        self.name, self.dtype, self.bits, self.B = name, dtype, bits, B
        self.x, self.stream, self.y, self.gidx, self.gval = x, stream, y, gidx, gval
        self.n = x.size

    def __repr__(self):
        return f"{self.name}-{np.dtype(self.dtype).name}-b{self.bits}-B{self.B}-n{self.n}"


def load_golden(dtype):
    tag = "f32" if dtype == np.float32 else "f64"
    d = np.load(os.path.join(GOLDEN, f"cases_{tag}.npz"), allow_pickle=False)
    out = []
    for name, (bits, B, n, xo, so, sn, go, m) in zip(d["names"], d["table"]):
        out.append(GoldenCase(str(name), dtype, int(bits), int(B), d["x"][xo:xo + n], d["streams"][so:so + sn],
                              d["y"][xo:xo + n], d["gidx"][go:go + m], d["gval"][go:go + m]))
    return out


GOLDEN_CASES = load_golden(np.float32) + load_golden(np.float64)


@pytest.fixture(scope="session")
def port():
    from oracle.oracle import Port

    return Port()


@pytest.fixture(scope="session")
def reference():
    from oracle.oracle import Reference, reference_available

    if not reference_available():
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    return Reference()


def oracle_codec(port, x, bits, B):
    """The oracle's stream and decode of the device tensor ``x`` (computed on
    the host from a copy of ``x``), returned as device tensors for comparison
    with what the HIP path wrote."""
    import torch

    xh = x.detach().reshape(-1).cpu().numpy()
    s = port.compress(xh, bits, B)
    y = port.decompress(s, xh.size, xh.dtype, bits, B)

    def dev(a, dtype):
        # a fresh contiguous tensor, also for an empty result
        t = torch.empty(a.size, dtype=dtype, device=x.device)
        if a.size:
            t.copy_(torch.from_numpy(np.ascontiguousarray(a)))
        return t

    return dev(s, torch.uint8), dev(y, x.dtype)


@pytest.fixture(scope="session")
def port_mt(port):
    """The oracle on every host core (block-parallel, byte-identical to one
    thread), for the large cases."""
    return port.with_threads(0)
