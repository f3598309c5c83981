"""aeo_vector_curl_div_noalloc (the entry point that allocates nothing) against
aeo_vector_curl_div with a caller workspace, through the C ABI.

Bar: bitwise equal (the same kernels on the same inputs), on two calls in a
row, the second with fewer fields, on the Gauss and geodesic product grids;
a request larger than the fixed scratch is refused on the host.
"""
import ctypes
import math
import warnings

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R = 6.371e6
OMEGA = 2.0 * math.pi / 86400.0


def _planet(kind):
    from aeolus_amd.planet import Planet
    kw = (dict(grid_type="latlon", nlat=64, nlon=128, l_max=42) if kind == "latlon"
          else dict(grid_type="geodesic", grid_resolution=3, l_max=10))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        return Planet.with_radius(R, OMEGA, product_quadrature="fine", **kw)


def _fields(n, npts, seed):
    import torch
    rng = np.random.default_rng(seed)
    fe = torch.as_tensor(rng.standard_normal((n, npts)), device="cuda")
    fn = torch.as_tensor(rng.standard_normal((n, npts)), device="cuda")
    return fe, fn


def _with_work(lib, nat, ext, fe, fn, nf, L):
    import torch
    nw = ctypes.c_longlong()
    nat.check(lib.aeo_vector_curl_div_work(ext._plan, nf, ctypes.byref(nw)), "work")
    work = torch.empty(nw.value, dtype=torch.float64, device="cuda")
    curl = torch.empty((nf, L + 1, L + 1), dtype=torch.complex128, device="cuda")
    div = torch.empty_like(curl)
    nat.check(lib.aeo_vector_curl_div(ext._plan, R, fe.data_ptr(), fn.data_ptr(), nf, 1,
                                      curl.data_ptr(), div.data_ptr(), work.data_ptr(),
                                      nat.stream()), "vector_curl_div")
    return curl, div


def _no_alloc(lib, nat, ext, fe, fn, nf, L):
    import torch
    curl = torch.full((nf, L + 1, L + 1), float("nan"), dtype=torch.complex128, device="cuda")
    div = torch.full_like(curl, float("nan"))
    nat.check(lib.aeo_vector_curl_div_noalloc(ext._plan, R, fe.data_ptr(), fn.data_ptr(), nf, 1,
                                              curl.data_ptr(), div.data_ptr(), nat.stream()),
              "vector_curl_div_noalloc")
    return curl, div


@pytest.mark.parametrize("kind", ["latlon", "geodesic"])
def test_no_alloc_equals_workspace_entry_point(kind):
    import torch
    from aeolus_amd import _native as nat
    print(f"case no-alloc vector curl/div {kind}", flush=True)
    lib = nat.lib()
    pl = _planet(kind)
    ext = pl.so.extended_product_transform()
    L = pl.so.l_max
    N = ext.n_points
    for call, nf in enumerate((6, 3)):      # second call: a smaller input
        fe, fn = _fields(nf, N, 100 + call)
        got = _no_alloc(lib, nat, ext, fe, fn, nf, L)
        ref = _with_work(lib, nat, ext, fe, fn, nf, L)
        torch.cuda.synchronize()
        for g, r in zip(got, ref):
            assert torch.isfinite(g.real).all() and torch.isfinite(g.imag).all()
            assert torch.equal(g, r), (kind, call, nf)


def test_no_alloc_refuses_oversize_requests():
    from aeolus_amd import _native as nat
    print("case no-alloc oversize", flush=True)
    lib = nat.lib()
    pl = _planet("latlon")
    ext = pl.so.extended_product_transform()
    need = ctypes.c_longlong()
    nat.check(lib.aeo_vector_curl_div_work(ext._plan, 1, ctypes.byref(need)), "work")
    too_many = (2 << 20) // need.value + 1
    rc = lib.aeo_vector_curl_div_noalloc(ext._plan, R, 16, 16, too_many, 1, 16, 16, None)
    assert rc == nat.AEO_ERR_UNSUPPORTED and b"16 MiB" in lib.aeo_last_error()
