"""The allocation-free entry point (sbpx_log_posterior_noalloc) against sbpx_log_posterior on the
same block matrix, held by the caller in device memory: bit-identical results on two calls in a
row, the second with fewer cells, and the size limit rejected without touching the device."""
import numpy as np
import pytest

from pkg import sbpx

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _noalloc(c, rows, cols, vals, dout, din, stream):
    dev = [torch.from_numpy(np.ascontiguousarray(x, np.int32)).cuda() for x in (rows, cols, vals, dout, din)]
    out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    st = c.L.sbpx_log_posterior_noalloc(c.h, len(rows), *[t.data_ptr() for t in dev], out.data_ptr(), stream)
    assert st == 0, c.L.sbpx_last_error(c.h)
    torch.cuda.synchronize()
    return float(out.cpu()[0])


def test_noalloc_log_posterior_equals_the_context_entry_point():
    m = sbpx()
    cfg = dict(m.CONFIGS["C2"], num_vertices=4000, num_blocks=20)
    op, oi, ip, ii, truth = m.generate_dcsbm(**cfg, seed=5)
    n = len(op) - 1
    c = m.Context(0)
    c.upload_graph(op, oi, ip, ii)
    rng = np.random.default_rng(3)
    stream = torch.cuda.current_stream().cuda_stream
    for call, B in enumerate((n, 300)):  # the second call has far fewer cells
        a = np.arange(n, dtype=np.int32) if B == n else rng.integers(0, B, n).astype(np.int32)
        c.build(B, a)
        rows, cols, vals = c.coo()
        _, dout, din, _ = c.state()
        want = c.log_posterior()
        got = _noalloc(c, rows, cols, vals, dout, din, stream if call == 0 else None)
        print("cells", len(rows), "log posterior", want, got)
        assert got == want
    with pytest.raises(m.SbpxError):
        c._chk(c.L.sbpx_log_posterior_noalloc(c.h, (1 << 28) + 1, None, None, None, None, None, None, None))
    c.close()
