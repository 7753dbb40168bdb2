"""Product-form (proton-neutron factorised) virtual H.X against the oracle's CSB product
and against the stored GPU operator; Davidson on it against the stored solve."""
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN_CASES

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", GOLDEN_CASES)
@pytest.mark.parametrize("m", [1, 3, 8, 16])
def test_factored_vs_oracle_product(gpu, golden, name, m):
    """FP64 to 1e-12 and FP32 to 2e-5 of max|y| against the oracle CSB product of the
    same synthetic H (every golden system: even and odd A, M=0 and M!=0, sd and fp)."""
    g = golden(name)
    _flush("factored %s m=%d" % (name, m))
    s = gpu.System.from_xml(g.xml)
    op = gpu.Operator.build(s, gpu.OP_FACTORED)
    o = oracle.System(g.xml)
    rp, c, v = o.h_upper(drop_small=True)
    n = op.n
    X = np.random.default_rng(m + 17).standard_normal((n, m))
    ref = oracle.csb_times(n, rp, c, v, X)
    scale = np.abs(ref).max()
    y64 = op.apply(X, gpu.PREC_FP64)
    assert np.max(np.abs(y64 - ref)) < 1e-12 * scale
    y32 = op.apply(X, gpu.PREC_FP32)
    assert np.max(np.abs(y32 - ref)) < 2e-5 * scale


@pytest.mark.parametrize("shell,nv,zv,twojz", [("sd", 4, 4, 0), ("sd", 5, 3, 2), ("fp", 3, 3, 0), ("fp", 2, 4, 4)])
def test_factored_vs_stored(gpu, tmp_path, shell, nv, zv, twojz):
    """Larger synthetic systems (n up to ~1e5): FP64 product form equals the stored FP64
    product to 1e-12 of max|y|; FP32 to 2e-5."""
    from cosmo_amd import synth
    _flush("factored vs stored %s %d+%d 2Jz=%d" % (shell, nv, zv, twojz))
    par = "+" if shell == "sd" else ("-" if (nv + zv) % 2 else "+")
    xml = synth.write_system(str(tmp_path), shell, "f", nv, zv, twojz, par, 5 + nv + 3 * zv)
    s = gpu.System.from_xml(xml)
    a = gpu.Operator.build(s)
    b = gpu.Operator.build(s, gpu.OP_FACTORED)
    n = a.n
    X = np.random.default_rng(3).standard_normal((n, 10))
    ya = a.apply(X, gpu.PREC_FP64)
    scale = np.abs(ya).max()
    yb = b.apply(X, gpu.PREC_FP64)
    assert np.max(np.abs(yb - ya)) < 1e-12 * scale
    yb32 = b.apply(X, gpu.PREC_FP32)
    assert np.max(np.abs(yb32 - ya)) < 2e-5 * scale
    # the ordered FP64 mode of a product-form operator regenerates rows in reference order
    yo = b.apply(X[:, :2], gpu.PREC_FP64_ORDERED)
    assert np.max(np.abs(yo - ya[:, :2])) < 1e-12 * scale
    assert b.nnz() == (0, 0)
    assert b.bytes() < a.bytes()


def test_factored_tuned_width_same_bits(gpu, tmp_path):
    """The first apply of each (precision, pitch) times every slice width of k_fact and keeps
    the fastest; every width sums each output in the same order, so the first (tuning) call
    and the later calls give the same bits, and a second operator of the same system too."""
    from cosmo_amd import synth
    xml = synth.write_system(str(tmp_path), "fp", "t", 3, 3, 0, "+", 9)
    s = gpu.System.from_xml(xml)
    b1 = gpu.Operator.build(s, gpu.OP_FACTORED)
    b2 = gpu.Operator.build(s, gpu.OP_FACTORED)
    X = np.random.default_rng(4).standard_normal((b1.n, 16))
    for prec in (gpu.PREC_FP64, gpu.PREC_FP32):
        for m in (3, 8, 16):
            y0 = b1.apply(X[:, :m], prec)
            y1 = b1.apply(X[:, :m], prec)
            y2 = b2.apply(X[:, :m], prec)
            y3 = b1.apply(X[:, :m], prec)  # captured as a graph
            y4 = b1.apply(X[:, :m], prec)  # graph replay
            assert np.array_equal(y0, y1) and np.array_equal(y0, y2)
            assert np.array_equal(y0, y3) and np.array_equal(y0, y4)


def test_factored_graph_replay_same_bits(gpu, golden):
    """Repeated applies on the same device buffers and stream are captured once into a HIP
    graph (second call) and replayed afterwards.  Each replay reads the current contents of
    X, so new inputs are copied in by ordinary stream work between calls, with host waits in
    between; every result must equal, bit for bit, an eager apply of the same input on
    buffers used only once."""
    import torch
    g = golden("sd4p4n_s0")
    s = gpu.System.from_xml(g.xml)
    b = gpu.Operator.build(s, gpu.OP_FACTORED)
    n = b.n
    st = torch.cuda.Stream()
    gen = torch.Generator(device="cuda").manual_seed(5)
    for prec, dt in ((gpu.PREC_FP64, torch.float64), (gpu.PREC_FP32, torch.float32)):
        inputs = [torch.randn(n, 16, dtype=dt, device="cuda", generator=gen) for _ in range(4)]
        outs = [torch.empty_like(x) for x in inputs]  # all kept alive: distinct addresses, eager path
        for x, y in zip(inputs, outs):
            b.apply_device(x[:, :11], y[:, :11], prec, stream=st)
        st.synchronize()
        X = torch.empty(n, 16, dtype=dt, device="cuda")
        Y = torch.empty(n, 16, dtype=dt, device="cuda")
        with torch.cuda.stream(st):
            for it in range(10):
                i = it % len(inputs)
                X.copy_(inputs[i])
                Y.fill_(-1.0)
                b.apply_device(X[:, :11], Y[:, :11], prec, stream=st)
                st.synchronize()
                assert torch.equal(Y[:, :11], outs[i][:, :11]), (prec, it)
                assert torch.all(Y[:, 11:] == -1.0)


def test_factored_apply_device_padded(gpu, golden):
    """Device-resident vectors with a padded power-of-two pitch; the padding columns hold
    NaN and must neither leak into the result nor be overwritten."""
    import torch
    g = golden("sd4p4n_s0")
    s = gpu.System.from_xml(g.xml)
    a = gpu.Operator.build(s)
    b = gpu.Operator.build(s, gpu.OP_FACTORED)
    n = a.n
    for prec, dt, tol in ((gpu.PREC_FP64, torch.float64, 1e-12), (gpu.PREC_FP32, torch.float32, 2e-5)):
        X = torch.full((n, 16), float("nan"), dtype=dt, device="cuda")
        X[:, :11] = torch.randn(n, 11, dtype=dt, device="cuda")
        Y = torch.full((n, 16), -7.0, dtype=dt, device="cuda")
        b.apply_device(X[:, :11], Y[:, :11], prec, stream=torch.cuda.current_stream())
        torch.cuda.synchronize()
        ref = a.apply(X[:, :11].double().cpu().numpy(), gpu.PREC_FP64)
        got = Y[:, :11].double().cpu().numpy()
        assert np.all(np.isfinite(got))
        assert np.max(np.abs(got - ref)) < tol * np.abs(ref).max()
        assert torch.all(Y[:, 11:] == -7.0)


def test_factored_davidson_matches_stored(gpu, golden):
    g = golden("sd4p4n_s0")
    s = gpu.System.from_xml(g.xml)
    a = gpu.Operator.build(s)
    b = gpu.Operator.build(s, gpu.OP_FACTORED)
    # unit vectors at the lowest diagonal entries (the reference's start) are symmetric
    # under the n <-> p exchange of this N=Z, M=0 system, so an exact solve from them
    # never reaches the antisymmetric states; a small random admixture opens both sectors
    n = a.n
    idx = np.argsort(a.diagonal(), kind="stable")[:8]
    init = np.zeros((n, 8))
    init[idx, np.arange(8)] = 1.0
    init += 1e-2 * np.random.default_rng(5).standard_normal((n, 8))
    kw = dict(want_vectors=False, lim=48, init=init)
    ea, _, ra, _, sa = gpu.davidson(a, 8, critr=1e-7, **kw)
    eb, _, rb, _, sb = gpu.davidson(b, 8, critr=1e-7, **kw)
    assert sa["converged"] and sb["converged"]
    assert np.max(np.abs(ea - eb)) < 1e-9
    em, _, rm, _, sm = gpu.davidson(b, 8, critr=1e-6, precision=gpu.PREC_FP32, **kw)
    assert sm["converged"]
    assert np.max(np.abs(em - ea)) <= 1e-3
    assert np.max(rm) <= 1e-6


def test_factored_rejects_single_species(gpu, tmp_path):
    from cosmo_amd import synth
    xml = synth.write_system(str(tmp_path), "sd", "p", 0, 3, 1, "+", 2)
    s = gpu.System.from_xml(xml)
    with pytest.raises(gpu.CosmoError, match="valence neutrons and protons"):
        gpu.Operator.build(s, gpu.OP_FACTORED)
