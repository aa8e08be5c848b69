"""Every tuned variant of the plain kernels (values per thread 2/4/8/16,
non-temporal hint, staged decompress stores, tile order, 64- or 256-thread
workgroups) against the oracle:
the defaults can change with tuning, so all variants must be bit-exact."""
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CASES = [(np.float64, 32, 32), (np.float64, 16, 1), (np.float64, 40, 64), (np.float64, 64, 512),
         (np.float64, 12, 16), (np.float64, 32, 1024), (np.float32, 16, 32), (np.float32, 8, 2),
         (np.float32, 12, 16), (np.float32, 21, 1024), (np.float32, 32, 4), (np.float32, 12, 1),
         (np.float32, 5, 8), (np.float64, 21, 2), (np.float64, 33, 128), (np.float32, 16, 4), (np.float64, 12, 4),
         (np.float64, 64, 1), (np.float32, 8, 8)]


@pytest.fixture(scope="module")
def fz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frsz_amd

    return frsz_amd


@pytest.mark.parametrize("dtype,bits,B", CASES, ids=lambda v: str(v))
def test_all_variants(fz, port, dtype, bits, B):
    from frsz_amd import _capi

    rng = np.random.default_rng(bits * 1000 + B)
    n = 3 * 4096 * 2 + 77
    x = (np.exp2(rng.uniform(-40, 40, n)) * rng.standard_normal(n)).astype(dtype)
    x[rng.random(n) < 0.05] = 0
    code = 0 if dtype == np.float32 else 1
    want = port.compress(x, bits, B)
    want_y = port.decompress(want, n, dtype, bits, B)
    xd = torch.from_numpy(x).cuda()
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    saved = [_capi.tune_get(op, code, bits) for op in (0, 1)]
    try:
        for op, vpt, nt, st, lin, sm in itertools.product((0, 1), (2, 4, 8, 16), (0, 1), (0, 1), (0, 1), (0, 1)):
            if (op == 0 and st) or (sm and not nt):
                continue  # staged stores: decompress only; 64-thread workgroups imply the non-temporal hint
            _capi.tune_set(op, code, bits, vpt, nt, st, lin, sm)
            if op == 0:
                got = fz.compress(xd, bits, B).cpu().numpy()
                assert np.array_equal(got, want), ("compress", vpt, nt, lin, sm)
            else:
                y = fz.decompress(torch.from_numpy(want).cuda(), n, tdt, bits, B).cpu().numpy()
                assert np.array_equal(y.view(np.uint8), want_y.view(np.uint8)), ("decompress", vpt, nt, st, lin, sm)
    finally:
        for op, cfg in zip((0, 1), saved):
            _capi.tune_set(op, code, bits, *cfg)
