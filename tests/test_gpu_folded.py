"""GPU tests of the libgpw folded-table path against the oracle (whose
restatement tests/test_oracle_folded.py pins by a brute-force periodic image
sum): periodic folding and support masks (bit for bit: integer masks, and
the folds add the same entries in the same order), collocate<T> and
integrate<T> on local windows with partly clear masks (1e-12 relative), the
host-array entry points behind the C++ templates, and the whole chain
tables -> fold -> collocate -> integrate on the device."""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


def _tables(rng, L, lengths, ldu):
    U = np.zeros((3, L, ldu))
    for d in range(3):
        x = np.linspace(-3.0, 3.0, lengths[d]) + rng.uniform(-0.1, 0.1)
        e = np.exp(-rng.uniform(0.2, 1.0) * x * x)
        for a in range(L):
            U[d, a, :lengths[d]] = x ** a * e
    return U


@pytest.mark.parametrize("case", [0, 1, 2])
def test_fold_tables_bitwise(gpw, oracle, torch_cuda, case):
    torch = torch_cuda
    rng = np.random.default_rng(300 + case)
    L = [3, 6, 11][case]
    period = [(12, 14, 16), (20, 9, 31), (7, 8, 64)][case]
    lengths = [(41, 5, 16), (13, 9, 100), (60, 61, 3)][case]
    centers = [int(rng.integers(-20, 80)) for _ in range(3)]
    ldu, ldp = max(lengths) + 3, max(period) + 5
    U = _tables(rng, L, lengths, ldu)
    say("fold", L, period, lengths, centers)
    f_ref, m_ref = oracle.fold_tables(U, lengths, centers, period, ldp)
    f, m = gpw.fold_tables(torch.from_numpy(U).cuda(), lengths, centers, period, ldp)
    torch.cuda.synchronize()
    assert np.array_equal(f.cpu().numpy(), f_ref)
    assert np.array_equal(m.cpu().numpy(), m_ref)
    # host entry point
    fh = np.zeros((3, L, ldp)); mh = np.zeros((3, ldp), dtype=np.int32)
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    gpw._check(gpw.lib().gpw_fold_tables_host(L, gpw._ci(lengths), gpw._ci(centers), gpw._ci(period),
                                              U.ctypes.data_as(dp), L, ldu, fh.ctypes.data_as(dp), ldp,
                                              mh.ctypes.data_as(ip), ldp))
    assert np.array_equal(fh, f_ref) and np.array_equal(mh, m_ref)


@pytest.mark.parametrize("L,gs,lc", [(1, (12, 13, 14), (0, 0, 0)), (4, (7, 9, 20), (3, 1, 5)),
                                      (9, (16, 16, 16), (2, 0, 4)), (17, (5, 6, 7), (1, 2, 3))])
def test_collocate_integrate_folded(gpw, oracle, torch_cuda, L, gs, lc):
    torch = torch_cuda
    rng = np.random.default_rng(L * 31 + gs[0])
    ldp = max(gs[d] + lc[d] for d in range(3)) + 2
    P = rng.uniform(-1, 1, (3, L, ldp))
    mask = (rng.uniform(0, 1, (3, ldp)) < 0.8).astype(np.int32)
    co = rng.uniform(-1, 1, (L, L, L))
    base = rng.uniform(-1, 1, gs)
    W = rng.uniform(-1, 1, gs)
    say("folded collocate/integrate", L, gs, lc)
    ref = oracle.collocate_folded(co, P, mask, gs, lc, density=base.copy())
    ref_co = oracle.integrate_folded(P, mask, gs, lc, W)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    d = dev(base)
    gpw.collocate_folded(dev(co), dev(P), dev(mask), gs, lc, d)
    got_co = gpw.integrate_folded(dev(P), dev(mask), gs, lc, dev(W))
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    m3 = (mask[0, lc[0]:lc[0] + gs[0]][:, None, None] * mask[1, lc[1]:lc[1] + gs[1]][None, :, None] *
          mask[2, lc[2]:lc[2] + gs[2]][None, None, :]).astype(bool)
    assert np.array_equal(got[~m3], base[~m3])
    assert rel_err(got - base, ref - base) < TOL
    assert rel_err(got_co.cpu().numpy(), ref_co) < TOL
    # host entry points (behind collocate<T> / integrate<T> of libgpw_amd.hpp)
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    h = base.copy()
    gpw._check(gpw.lib().gpw_collocate_folded_host(L, co.ctypes.data_as(dp), P.ctypes.data_as(dp), ldp,
                                                   mask.ctypes.data_as(ip), ldp, gpw._ci(gs), gpw._ci(lc),
                                                   h.ctypes.data_as(dp)))
    assert np.array_equal(h[~m3], base[~m3])
    assert rel_err(h - base, ref - base) < TOL
    hc = np.zeros((L, L, L))
    gpw._check(gpw.lib().gpw_integrate_folded_host(L, P.ctypes.data_as(dp), ldp, mask.ctypes.data_as(ip),
                                                   ldp, gpw._ci(gs), gpw._ci(lc), W.ctypes.data_as(dp),
                                                   hc.ctypes.data_as(dp)))
    assert rel_err(hc, ref_co) < TOL


def test_folded_chain_on_device(gpw, oracle, torch_cuda):
    """unfolded tables wider than the period -> fold -> collocate -> integrate,
    all on the device, against the same chain in the oracle; the adjoint
    identity between the device collocate and integrate."""
    torch = torch_cuda
    rng = np.random.default_rng(77)
    L, period, lengths = 5, (18, 20, 22), (45, 33, 21)
    centers = [3, 17, 9]
    U = _tables(rng, L, lengths, max(lengths))
    co = rng.uniform(-1, 1, (L, L, L))
    W = rng.uniform(-1, 1, period)
    say("folded chain", period, lengths)
    f_ref, m_ref = oracle.fold_tables(U, lengths, centers, period, max(period))
    g_ref = oracle.collocate_folded(co, f_ref, m_ref, period, (0, 0, 0))
    c_ref = oracle.integrate_folded(f_ref, m_ref, period, (0, 0, 0), W)
    f, m = gpw.fold_tables(torch.from_numpy(U).cuda(), lengths, centers, period, max(period))
    g = torch.zeros(period, dtype=torch.float64, device="cuda")
    gpw.collocate_folded(torch.from_numpy(co).cuda(), f, m, period, (0, 0, 0), g)
    c = gpw.integrate_folded(f, m, period, (0, 0, 0), torch.from_numpy(W).cuda())
    torch.cuda.synchronize()
    assert rel_err(g.cpu().numpy(), g_ref) < TOL
    assert rel_err(c.cpu().numpy(), c_ref) < TOL
    lhs = float((g.cpu().numpy() * W).sum())
    rhs = float(np.einsum("agb,gba->", co, c.cpu().numpy()))
    assert abs(lhs - rhs) < 1e-11 * max(1.0, abs(lhs))
