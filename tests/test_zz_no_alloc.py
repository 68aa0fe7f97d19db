"""reno_hop_apply_noalloc (scratch in a static device array) against
reno_hop_apply (caller workspace) through the C ABI: identical results on
two calls in a row, the second with a smaller operator, and a clean refusal
of an operator whose scratch exceeds the static array."""
import ctypes

import pytest
import torch

from renormalizer_amd import native

pytestmark = pytest.mark.gpu


def _operands(ml, d, mr, dl, dr, dev, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)

    def c(*shape):
        return torch.complex(torch.randn(*shape, generator=g, dtype=torch.float64),
                             torch.randn(*shape, generator=g, dtype=torch.float64)).to(dev)

    L = c(ml, dl, ml)
    R = c(mr, dr, mr)
    W = torch.randn(dl, d, d, dr, generator=g, dtype=torch.float64).to(dev)
    x = c(ml, d, mr)
    return L, W, R, x


def _noalloc(L, W, R, x):
    ml, d, mr = x.shape
    desc = native.HopDesc(1, native.PREC_FP64, ml, d, mr, W.shape[0], W.shape[3],
                          L.data_ptr(), W.data_ptr(), R.data_ptr())
    out = torch.empty_like(x)
    native.check(native.lib().reno_hop_apply_noalloc(ctypes.byref(desc), ctypes.c_void_p(x.data_ptr()),
                                                     ctypes.c_void_p(out.data_ptr()), native._stream()),
                 "reno_hop_apply_noalloc")
    return out


def test_noalloc_matches_workspace_entry(dev):
    print("[noalloc hop: 256x10x256 then 96x8x64]", flush=True)
    for shape, seed in (((256, 10, 256, 3, 3), 1), ((96, 8, 64, 3, 4), 2)):
        L, W, R, x = _operands(*shape, dev, seed)
        ref = native.hop_apply(L, W, R, x)
        got = _noalloc(L, W, R, x)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (shape, (got - ref).abs().max().item())


def test_noalloc_rejects_oversized_operator(dev):
    print("[noalloc hop: oversized operator refused]", flush=True)
    L = torch.zeros((256, 3, 256), dtype=torch.complex128, device=dev)
    W = torch.zeros((3, 20, 20, 3), dtype=torch.float64, device=dev)
    x = torch.zeros((256, 20, 256), dtype=torch.complex128, device=dev)
    with pytest.raises(ValueError, match="static array"):
        _noalloc(L, W, L, x)
