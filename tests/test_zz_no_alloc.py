"""cosmo_op_apply_device_noalloc (the product without allocation, scratch in a fixed device
array per GPU) against cosmo_op_apply_device on the same product-form operator: the same
bits, call after call, the second call of each pair with a smaller input; against the
oracle's CSB product; its limits refused with an error."""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pair(gpu, op, m, ld, prec, seed, stream=None):
    """Y from both entry points for one random X (n x m in a pitch-ld buffer)."""
    import torch
    dt = torch.float64 if prec == gpu.PREC_FP64 else torch.float32
    g = torch.Generator(device="cuda").manual_seed(seed)
    Xb = torch.randn(op.n, ld, device="cuda", dtype=dt, generator=g)
    X = Xb[:, :m]
    Ya = torch.full((op.n, ld), float("nan"), device="cuda", dtype=dt)[:, :m]
    Yb = torch.full((op.n, ld), float("nan"), device="cuda", dtype=dt)[:, :m]
    op.apply_device(X, Ya, prec, stream)
    op.apply_device_noalloc(X, Yb, prec, stream)
    torch.cuda.synchronize()
    return X, Ya, Yb


def test_noalloc_equals_the_allocating_apply(gpu, golden):
    """Two calls in a row on each operator, the second with a smaller input (fewer vectors
    in a narrower pitch, then a smaller operator), FP64 and FP32, on the caller's stream
    and on the null stream: the same bits as cosmo_op_apply_device, columns past m left as
    they were, FP64 within 1e-12 of the oracle's CSB product."""
    import torch
    big = golden("sd4p4n_s0")   # n = 28,503
    small = golden("fp2p2n_s5")  # n = 4,000
    ops = []
    for g in (big, small):
        ops.append((g, gpu.Operator.build(gpu.System.from_xml(g.xml), gpu.OP_FACTORED)))
    st = torch.cuda.Stream()
    for prec in (gpu.PREC_FP64, gpu.PREC_FP32):
        for (g, op) in ops:
            for (m, ld, stream) in ((16, 16, st), (3, 4, None)):
                _flush("noalloc %s n=%d prec=%d m=%d ld=%d stream=%s" % (g.name, op.n, prec, m, ld, stream is not None))
                X, Ya, Yb = _pair(gpu, op, m, ld, prec, 7 * m + prec, stream)
                assert torch.equal(Ya, Yb)
                full = Yb.as_strided((op.n, ld), (ld, 1))
                if m < ld:
                    assert torch.isnan(full[:, m:]).all()
                if prec == gpu.PREC_FP64:
                    rp, c, v = oracle.System(g.xml).h_upper(drop_small=True)
                    ref = oracle.csb_times(op.n, rp, c, v, X.cpu().numpy())
                    assert np.max(np.abs(Yb.cpu().numpy() - ref)) <= 1e-12 * np.abs(ref).max()


def test_noalloc_large_system_and_limits(gpu, tmp_path):
    """fp 4p4n (n = 1,963,461): FP32 at pitch 4 fits the 64 MiB scratch and equals the
    allocating apply bit for bit; FP64 at pitch 16 needs 502 MB and is refused with an
    error, as are stored operators and the ordered FP64 mode."""
    import torch
    from cosmo_amd import synth
    xml = synth.write_system(str(tmp_path), "fp", "fp4p4n", 4, 4, 0, "+", 2)
    s = gpu.System.from_xml(xml)
    op = gpu.Operator.build(s, gpu.OP_FACTORED)
    _flush("noalloc fp4p4n n=%d FP32 m=4" % op.n)
    X, Ya, Yb = _pair(gpu, op, 4, 4, gpu.PREC_FP32, 3)
    assert torch.equal(Ya, Yb)
    _flush("noalloc fp4p4n FP32 m=2 after m=4")
    X, Ya, Yb = _pair(gpu, op, 2, 4, gpu.PREC_FP32, 4)
    assert torch.equal(Ya, Yb)
    X = torch.zeros(op.n, 16, device="cuda", dtype=torch.float64)
    Y = torch.zeros_like(X)
    with pytest.raises(gpu.CosmoError):
        op.apply_device_noalloc(X, Y, gpu.PREC_FP64)
    with pytest.raises(gpu.CosmoError):
        op.apply_device_noalloc(X[:, :2], Y[:, :2], gpu.PREC_FP64_ORDERED)
    op.close()
    small = gpu.System.from_xml(synth.write_system(str(tmp_path), "sd", "s", 2, 2, 0, "+", 1))
    stored = gpu.Operator.build(small)
    X = torch.zeros(stored.n, 4, device="cuda", dtype=torch.float64)
    with pytest.raises(gpu.CosmoError):
        stored.apply_device_noalloc(X, torch.zeros_like(X), gpu.PREC_FP64)
    torch.cuda.synchronize()
