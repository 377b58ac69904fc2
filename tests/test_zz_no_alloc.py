"""GPU test of the no-allocation dense matvec (ed_dense_matvec_f64_noalloc)
through the C ABI, against the allocating entry point ed_dense_matvec_f64.

Two calls in a row on the same GPU (so the second reuses the shared scratch
array the first one filled), the second with a smaller input:
- 81,920 panels, 2 right-hand sides (run in passes of one: the scratch holds
  split x n = 16 x 81,920 doubles), then 1,280 panels, 3 right-hand sides:
  both entry points run the split form, so the results are bit-identical;
- one right-hand side (the allocating entry point then runs the symmetric
  form, a different summation order): max relative difference <= 1e-13;
- the small case also against the oracle's FP64 dense matvec (1e-12);
- n beyond the limit is rejected on the host (ED_ERR_ARG) before any launch.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import dense as odense

pytestmark = pytest.mark.gpu


def _call(fn, C, A, S, si):
    from electrodrive_amd import _native as nat
    out = torch.empty_like(S)
    rc = fn(nat.ptr(C), nat.ptr(A), nat.ptr(S), nat.ptr(si), nat.ptr(out), C.shape[0],
            S.shape[0], nat.stream_ptr(S.device))
    return rc, out


def _case(subdiv, nrhs, seed):
    from electrodrive_amd.core import bem_mesh
    m = bem_mesh.icosphere(subdiv)
    rng = np.random.default_rng(seed)
    C = torch.as_tensor(m.centroids, device="cuda").contiguous()
    A = torch.as_tensor(m.areas, device="cuda").contiguous()
    S = torch.as_tensor(rng.normal(size=(nrhs, m.n_panels)), device="cuda").contiguous()
    si = torch.as_tensor(rng.uniform(1e10, 2e10, size=m.n_panels), device="cuda").contiguous()
    return m, C, A, S, si


def test_noalloc_dense_matches_allocating_entry_point():
    from electrodrive_amd import _native as nat
    L = nat.lib()
    for subdiv, nrhs, seed in ((6, 2, 1), (3, 3, 2)):
        print(f"noalloc dense: icosphere({subdiv}) nrhs={nrhs} seed={seed}", flush=True)
        m, C, A, S, si = _case(subdiv, nrhs, seed)
        rc0, ref = _call(L.ed_dense_matvec_f64, C, A, S, si)
        rc1, got = _call(L.ed_dense_matvec_f64_noalloc, C, A, S, si)
        torch.cuda.synchronize()
        assert rc0 == 0 and rc1 == 0
        assert torch.equal(got, ref), float((got - ref).abs().max())
        # one right-hand side: symmetric form vs split form
        rc0, ref1 = _call(L.ed_dense_matvec_f64, C, A, S[:1].contiguous(), si)
        rc1, got1 = _call(L.ed_dense_matvec_f64_noalloc, C, A, S[:1].contiguous(), si)
        torch.cuda.synchronize()
        assert rc0 == 0 and rc1 == 0
        d = float((got1 - ref1).abs().max() / ref1.abs().max())
        print(f"  one RHS: max rel diff vs symmetric form {d:.2e}", flush=True)
        assert d <= 1e-13
        if subdiv == 3:
            o = odense.dense_matvec(m.centroids, m.areas, S[0].cpu().numpy(), si.cpu().numpy())
            g = got1[0].cpu().numpy()
            assert np.max(np.abs(g - o)) / np.max(np.abs(o)) < 1e-12


def test_noalloc_dense_rejects_oversize_n():
    from electrodrive_amd import _native as nat
    L = nat.lib()
    _, C, A, S, si = _case(2, 1, 3)
    out = torch.empty_like(S)
    rc = L.ed_dense_matvec_f64_noalloc(nat.ptr(C), nat.ptr(A), nat.ptr(S), nat.ptr(si),
                                       nat.ptr(out), 131073, 1, nat.stream_ptr(S.device))
    assert rc == nat.ED_ERR_ARG
    assert b"scratch limit" in L.ed_last_error()
    _ = ctypes
