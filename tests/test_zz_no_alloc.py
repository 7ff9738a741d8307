"""aok_ovl2c_noalloc (no reference counterpart; for callers that cannot allocate device
memory): through the C ABI, against aok_ovl2c on the same inputs.  Both run the same
kernels in the same order per element, so the outputs must be bit for bit equal.  Two
calls in a row, the second with a smaller batch (the shared scratch array keeps the first
call's tables), then the larger one again; a batch beyond the fixed scratch size is refused
before any launch; the calls take no device memory."""
import numpy as np
import pytest
import torch

import aok
from aok import _lib, configs
from helpers import dev_f, dev_rows, log

pytestmark = pytest.mark.gpu


def _inputs(d, B, dtype, dev):
    pA = dev_rows(d["posA"][:B], dtype, dev)
    pB = dev_rows(d["posB"][:B], dtype, dev)
    sA = torch.as_tensor(d["sidxA"][:B], dtype=torch.int64, device=dev)
    sB = torch.as_tensor(d["sidxB"][:B], dtype=torch.int64, device=dev)
    return pA, pB, sA, sB


def _out(basis, B, dtype, dev):
    N = basis.nbf_total
    return torch.full((B, N, N), float("nan"), dtype=dtype, device=dev).permute(2, 1, 0)


def _call(ctx, name, basis, coef, zeta, pA, pB, sA, sB, out, dtype):
    B = pA.shape[0]
    ns, K, NZ = zeta.shape
    code = _lib.AOK_F32 if dtype == torch.float32 else _lib.AOK_F64
    ctx.call(name, code, 0, pA.data_ptr(), pB.data_ptr(), sA.data_ptr(), sB.data_ptr(), _lib.AOK_I64, B, ns, K, NZ,
             _lib.i64_array(basis.ls), coef.data_ptr(), zeta.data_ptr(), float(basis.lengthscale), out.data_ptr())
    return out


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
def test_noalloc_matches_ovl2c(dev, dtype):
    d = configs.c1(1, B=65536)
    cb = d["basis"]
    spec, _, _ = cb.ao_maps()
    R = aok.GaussianTypeRadials(dev_f(cb.zeta, dtype, dev), dev_f(cb.D, dtype, dev), spec, None, (1, 2, 3))
    basis = aok.compile_basis(aok.AtomicOrbitals(R, 2, length_unit="bohr"))
    coef = aok.as_fortran(basis.coef.to(dtype))
    zeta = aok.as_fortran(basis.zeta.to(dtype))
    sizes = (65536, 1000, 65536)  # second call smaller, then the first size again
    ins = [_inputs(d, B, dtype, dev) for B in sizes]
    ref_ctx = aok.Context(dev.index or 0)
    refs = [_call(ref_ctx, "aok_ovl2c", basis, coef, zeta, *x, _out(basis, B, dtype, dev), dtype)
            for B, x in zip(sizes, ins)]
    outs = [_out(basis, B, dtype, dev) for B in sizes]  # every tensor exists before the reading below
    ctx = aok.Context(dev.index or 0)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info(dev)[0]
    for B, x, out, ref in zip(sizes, ins, outs, refs):
        log(f"noalloc 2c forward B={B} dtype={dtype}")
        _call(ctx, "aok_ovl2c_noalloc", basis, coef, zeta, *x, out, dtype)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info(dev)[0]
    # torch allocated nothing in between, so free device memory moves only if the library did
    assert free0 - free1 < (1 << 20), (free0, free1)
    for B, out, ref in zip(sizes, outs, refs):
        assert torch.equal(out, ref), B
    # beyond the fixed scratch: refused before any launch (small buffers; nothing reads them)
    buf = torch.zeros(64, dtype=dtype, device=dev)
    p = buf.data_ptr()
    with pytest.raises(aok.AOKArgumentError, match="scratch"):
        ctx.call("aok_ovl2c_noalloc", _lib.AOK_F32 if dtype == torch.float32 else _lib.AOK_F64, 0, p, p, None, None,
                 _lib.AOK_I64, 1 << 25, cb.ns, cb.K, cb.NZ, _lib.i64_array(cb.ls), p, p, 1.0, p)
    with pytest.raises(aok.AOKArgumentError):  # fast kernels only
        ctx.call("aok_ovl2c_noalloc", _lib.AOK_F32, 1, p, p, None, None, _lib.AOK_I64, 1, cb.ns, cb.K, cb.NZ,
                 _lib.i64_array(cb.ls), p, p, 1.0, p)
    # an empty batch is a no-op
    ctx.call("aok_ovl2c_noalloc", _lib.AOK_F32, 0, None, None, None, None, _lib.AOK_I64, 0, cb.ns, cb.K, cb.NZ,
             _lib.i64_array(cb.ls), p, p, 1.0, None)
    ctx.close()
    ref_ctx.close()
