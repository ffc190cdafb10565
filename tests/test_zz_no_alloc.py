"""The allocation-free spMV_ntimes entry point (dq_spmv_ntimes_device_noalloc) against the allocating one
(dq_spmv_ntimes_device), through the C ABI: two calls in a row on untuned qubit Hamiltonians, the second smaller, so
the zero-block map is made in the library's static scratch each time (the first call's contents are still there when
the second runs); FP64 results bit for bit equal and within 1e-12 of the oracle. Then the same call captured into a
graph (the map is made inside the graph) and replayed three times with eager work between the replays. A map larger
than the scratch (over 2^22 words) needs a matrix of more than 10^5 diagonals of more than 2^17 rows, tens of GB, so
that error path is not exercised here."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402
from diaq_amd import hamiltonians as Hm  # noqa: E402


@pytest.fixture(scope="module")
def d():
    import diaq_amd  # noqa: F401
    import diaq_gpu
    return diaq_gpu


def rel(a, b):
    b = np.asarray(b)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - b)) / (den if den > 0 else 1.0))


def run(capi, fn, D, x, n_times, stream=0, in_place=False):
    import torch
    xr = torch.tensor(x.real, device="cuda")
    xi = torch.tensor(x.imag, device="cuda")
    wr, wi = torch.empty_like(xr), torch.empty_like(xi)
    yr, yi = (xr, xi) if in_place else (torch.empty_like(xr), torch.empty_like(xi))
    capi.check(getattr(capi.lib(), fn)(D.handle, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(), yi.data_ptr(), n_times,
                                       wr.data_ptr(), wi.data_ptr(), stream))
    torch.cuda.synchronize()
    return yr.cpu().numpy() + 1j * yi.cpu().numpy()


def test_noalloc_ntimes_matches_the_allocating_entry(d):
    from diaq_amd import capi
    for q, model, n_times in ((14, "tfim", 3), (11, "xxz", 4)):  # the second call smaller than the first
        rows, cols, offs, vals = Hm.tfim(q, 1.0, 0.7) if model == "tfim" else Hm.xxz(q, seed=9)
        D = d.to_device(d.diaQMatrix.from_diagonals([rows, cols], offs, vals))
        rng = np.random.default_rng(q)
        x = rng.standard_normal(rows) + 1j * rng.standard_normal(rows)
        x /= np.linalg.norm(x)
        want = run(capi, "dq_spmv_ntimes_device", D, x, n_times)
        got = run(capi, "dq_spmv_ntimes_device_noalloc", D, x, n_times)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(run(capi, "dq_spmv_ntimes_device_noalloc", D, x, n_times, in_place=True), want)
        assert rel(got, O.spmv_ntimes(O.Dia(rows, cols, offs, vals), x, n_times)) < 1e-12


def test_noalloc_ntimes_in_a_captured_graph(d):
    import torch
    from diaq_amd import capi
    rows, cols, offs, vals = Hm.tfim(13, 1.0, 0.7)
    D = d.to_device(d.diaQMatrix.from_diagonals([rows, cols], offs, vals))
    rng = np.random.default_rng(5)
    x = rng.standard_normal(rows) + 1j * rng.standard_normal(rows)
    want = run(capi, "dq_spmv_ntimes_device", D, x, 3)
    xr = torch.tensor(x.real, device="cuda")
    xi = torch.tensor(x.imag, device="cuda")
    yr, yi = torch.empty_like(xr), torch.empty_like(xi)
    wr, wi = torch.empty_like(xr), torch.empty_like(xi)
    cs = torch.cuda.Stream()
    cs.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=cs):
        capi.check(capi.lib().dq_spmv_ntimes_device_noalloc(D.handle, xr.data_ptr(), xi.data_ptr(), yr.data_ptr(),
                                                            yi.data_ptr(), 3, wr.data_ptr(), wi.data_ptr(),
                                                            cs.cuda_stream))
    other = Hm.tfim(10, 1.0, 0.7)
    Dother = d.to_device(d.diaQMatrix.from_diagonals([other[0], other[1]], other[2], other[3]))
    xo = rng.standard_normal(other[0]) + 0j
    for _ in range(3):
        yr.fill_(7.0)
        yi.fill_(7.0)
        # eager work between the replays, including another allocation-free call that overwrites the scratch map
        run(capi, "dq_spmv_ntimes_device_noalloc", Dother, xo, 2)
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(yr.cpu().numpy() + 1j * yi.cpu().numpy(), want)
