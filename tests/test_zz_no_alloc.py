"""psum_hip_compact_noalloc (static device scratch, no allocation) against
psum_hip_compact (caller workspace): identical survivors, order and count on
two calls in a row, the second smaller; sizes beyond the scratch's limit are
rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _arrays(n, dim, dtype, seed, holes):
    g = torch.Generator(device="cuda").manual_seed(seed)
    pos = [torch.rand(n, generator=g, device="cuda", dtype=dtype) * 100 for _ in range(dim)]
    vel = [torch.randn(n, generator=g, device="cuda", dtype=dtype) for _ in range(dim)]
    dead = torch.rand(n, generator=g, device="cuda") < holes
    pos[0][dead] = float("nan")
    return pos, vel


def _run(fn_noalloc, pos, vel):
    from psum_amd._lib import lib, check
    from psum_amd.engine import _dtype_code, _ptrs, _p, _stream
    dim, n = len(pos), pos[0].numel()
    op = [torch.full_like(a, -7.0) for a in pos]
    ov = [torch.full_like(a, -7.0) for a in vel]
    cnt = C.c_int64(-1)
    if fn_noalloc:
        check(lib().psum_hip_compact_noalloc(_dtype_code(pos[0]), dim, _ptrs(pos), _ptrs(vel), n,
                                             _ptrs(op), _ptrs(ov), C.byref(cnt), _stream()),
              "psum_hip_compact_noalloc")
    else:
        wb = lib().psum_hip_compact_workspace(n)
        work = torch.empty(max(1, wb), dtype=torch.uint8, device="cuda")
        check(lib().psum_hip_compact(_dtype_code(pos[0]), dim, _ptrs(pos), _ptrs(vel), n, _ptrs(op),
                                     _ptrs(ov), C.byref(cnt), _p(work), wb, _stream()),
              "psum_hip_compact")
    torch.cuda.synchronize()
    m = cnt.value
    return m, [a.cpu().numpy() for a in op], [a.cpu().numpy() for a in ov]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_compact_noalloc_matches_workspace_entry_point(dtype):
    # second call smaller, other dimension: the scratch keeps the first
    # call's contents, which the second must not depend on
    for n, dim, seed, holes in ((3_000_017, 3, 1, 0.3), (100_003, 2, 2, 0.05)):
        print(f"compact_noalloc: n={n} dim={dim} dtype={dtype}", flush=True)
        pos, vel = _arrays(n, dim, dtype, seed, holes)
        m_ref, p_ref, v_ref = _run(False, pos, vel)
        m, p, v = _run(True, pos, vel)
        assert m == m_ref
        x0 = pos[0].cpu().numpy()
        assert m == int(np.isfinite(x0).sum())
        for a, b in zip(p + v, p_ref + v_ref):
            assert np.array_equal(a[:m], b[:m])
            assert np.all(a[m:] == -7.0)


def test_compact_noalloc_rejects_sizes_beyond_its_scratch():
    from psum_amd._lib import lib
    PSUM_ERR_ARG = 1
    from psum_amd.engine import _ptrs, _stream
    pos, vel = _arrays(16, 2, torch.float32, 3, 0.0)
    cnt = C.c_int64(-1)
    st = lib().psum_hip_compact_noalloc(1, 2, _ptrs(pos), _ptrs(vel), (1 << 31) + 1, _ptrs(pos),
                                        _ptrs(vel), C.byref(cnt), _stream())
    assert st == PSUM_ERR_ARG
    assert cnt.value == -1
    assert b"2^31" in lib().psum_hip_last_error()
