"""The plain-kernel autotuner (frsz2_tune_auto): it times every variant on
the caller's data, keeps one, and never changes a result byte."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def mods():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import frsz_amd
    from frsz_amd import tune

    return frsz_amd, tune


@pytest.mark.parametrize("dtype,bits,B,n", [(torch.float32, 16, 32, 2_000_003), (torch.float64, 21, 64, 1_000_001),
                                            (torch.float32, 12, 5, 300_007)])
def test_autotune_keeps_results(mods, dtype, bits, B, n):
    fz, tune = mods
    g = torch.Generator(device="cuda")
    g.manual_seed(n)
    x = (torch.randn(n, device="cuda", generator=g, dtype=torch.float64) * 1e4).to(dtype)
    s0 = fz.compress(x, bits, B)
    y0 = fz.decompress(s0, n, dtype, bits, B)
    saved = [tune.get_variant(op, dtype, bits) for op in (0, 1)]
    try:
        chosen = tune.autotune(x, bits, B, reps=2)
        for op, name in ((0, "compress"), (1, "decompress")):
            assert tune.get_variant(op, dtype, bits) == chosen[name]
            assert chosen[name]["values_per_thread"] in (2, 4, 8, 16)
        s1 = fz.compress(x, bits, B)
        assert torch.equal(s1, s0)
        assert torch.equal(fz.decompress(s1, n, dtype, bits, B).view(torch.uint8), y0.view(torch.uint8))
    finally:
        for op, v in zip((0, 1), saved):
            tune.set_variant(op, dtype, bits, **v)


def test_autotune_rejects_bad_arguments(mods):
    fz, tune = mods
    from frsz_amd import _capi

    x = torch.randn(1000, device="cuda")
    with pytest.raises(fz.Frsz2Error) as e:
        tune.autotune(x, 3, 32)
    assert e.value.code == 2
    # value buffer smaller than n values
    assert _capi.lib.frsz2_tune_auto(0, 0, 16, 32, 1000, x.data_ptr(), 100, x.data_ptr(), 1 << 20, 1, None, None,
                                     None) == _capi.ECAPACITY
