"""The allocation-free single-task integrate
(gpw_integrate_pgf_product_rspace_noalloc) against the allocating device
entry point it varies (gpw_integrate_pgf_product_rspace_dev), both called
through the C ABI: two calls in a row, the second with a smaller task, so
the second runs on the scratch the first left behind. hab and forces agree
to 1e-10 relative (the moment sums use atomics, so the two are not bitwise
equal) and match the oracle. A task whose scratch would exceed
GPW_NOALLOC_SCRATCH_BYTES is refused before any device work."""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import random_task, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-10


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _integrate(gpw, torch, name, t, W, pab):
    hab = torch.zeros((t["maxco"], t["maxco"]), dtype=torch.float64, device="cuda")
    f = torch.zeros(6, dtype=torch.float64, device="cuda")
    Wd, pd = _dev(torch, W), _dev(torch, pab)
    fn = getattr(gpw.lib(), name)
    code = fn(*gpw._single_args(t), C.c_void_p(Wd.data_ptr()), C.c_void_p(pd.data_ptr()),
              C.c_void_p(hab.data_ptr()), C.c_void_p(f.data_ptr()),
              C.c_void_p(torch.cuda.current_stream().cuda_stream))
    gpw._check(code)
    torch.cuda.synchronize()
    return hab.cpu().numpy(), f.cpu().numpy().reshape(2, 3)


@pytest.mark.parametrize("ortho", [True, False])
def test_noalloc_matches_allocating_entry_point(gpw, oracle, torch_cuda, ortho):
    torch = torch_cuda
    rng = np.random.default_rng(1234 + ortho)
    tasks = [
        random_task(rng, oracle, la=(0, 3), lb=(1, 3), ortho=ortho, shear=0.0 if ortho else 0.2,
                    npts=(48, 50, 52), cell=(12.0, 12.5, 13.0), zeta=0.8, zetb=1.1),
        random_task(rng, oracle, la=(0, 1), lb=(0, 1), ortho=ortho, shear=0.0 if ortho else 0.2,
                    npts=(20, 22, 24), cell=(6.0, 6.5, 7.0), zeta=3.0, zetb=2.0),
    ]
    for k, (t, pab) in enumerate(tasks):
        say("noalloc", "ortho" if ortho else "general", "call", k, t["npts"], t["la_max"], t["lb_max"])
        W = rng.uniform(-1.0, 1.0, (t["ngrid"][2], t["ngrid"][1], t["ngrid"][0]))
        h_a, f_a = _integrate(gpw, torch, "gpw_integrate_pgf_product_rspace_dev", t, W, pab)
        h_b, f_b = _integrate(gpw, torch, "gpw_integrate_pgf_product_rspace_noalloc", t, W, pab)
        assert np.abs(h_a).max() > 0.0
        assert rel_err(h_b, h_a) < TOL, k
        assert rel_err(f_b, f_a) < TOL, k
        h_ref, f_ref = oracle.integrate(t, W, pab=pab, forces=True)
        assert rel_err(h_b, h_ref) < TOL, k
        assert rel_err(f_b, f_ref) < TOL, k


def test_noalloc_refuses_oversized_task(gpw, oracle, torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(77)
    # a 25 bohr lattice box on a 1 bohr, 4-point periodic cell: 51 images per
    # axis, about 132,000 sub-boxes of geometry (more than 8 MiB)
    t, pab = random_task(rng, oracle, la=(0, 0), lb=(0, 0), ortho=False, npts=(4, 4, 4),
                         cell=(1.0, 1.0, 1.0), radius=25.0)
    W = np.zeros((4, 4, 4))
    say("noalloc oversized task")
    with pytest.raises(gpw.GpwError) as e:
        _integrate(gpw, torch, "gpw_integrate_pgf_product_rspace_noalloc", t, W, pab)
    assert e.value.code == 1 and "GPW_NOALLOC_SCRATCH_BYTES" in str(e.value)
