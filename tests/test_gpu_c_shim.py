"""The reference-compatible ``_C`` shim: the reference wrappers' exact calling
patterns (reference grey.py:92-100, binary.py:23-25, distance_transform.py:89-94,
:125-130, optimal_transport.py:75-113) through torchmorph_amd._C."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def C():
    from torchmorph_amd import _C

    return _C


def test_shim_grey(C, oracle_mod):
    rng = np.random.default_rng(0)
    x = rng.random((2, 1, 9, 11)).astype(np.float32)
    st = torch.zeros((3, 3))
    out = C.grey_dilation_cuda(torch.tensor(x).cuda().contiguous().float(), st,
                               torch.empty(0, dtype=torch.bool), [0, 0], 1, 0.0)
    want = oracle_mod.grey(x, None, None, (3, 3), (0, 0), "reflect", 0.0, True)
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_shim_binary(C, oracle_mod):
    rng = np.random.default_rng(1)
    x = rng.random((2, 1, 9, 11)) < 0.6
    st = torch.tensor([[0, 1, 0], [1, 1, 1], [0, 1, 0]], dtype=torch.bool)
    out = C.binary_erosion_cuda(torch.tensor(x).cuda(), st, [0, 0], False)
    want = oracle_mod.binary(x, st.numpy(), (0, 0), False, False)
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_shim_edt_cdt(C, oracle_mod):
    rng = np.random.default_rng(2)
    m = rng.random((2, 1, 20, 30)) < 0.8
    x = torch.tensor(m).cuda().float().contiguous()
    d, i = C.edt_cuda(x, [1.0, 1.0], True, True)
    od, oi = oracle_mod.edt(m)
    assert torch.equal(d.cpu(), torch.from_numpy(od.astype(np.float32)))
    assert np.array_equal(i.cpu().numpy(), oi)
    d, none = C.edt_cuda(x, [1.0, 1.0], True, False)
    assert none is None
    d, i = C.cdt_cuda(x, "taxicab", True, True)
    od, oi = oracle_mod.cdt(m, "taxicab")
    assert np.array_equal(d.cpu().numpy(), od.astype(np.float32))
    assert np.array_equal(i.cpu().numpy(), oi)


def test_shim_sinkhorn_logiter_in_place(C, oracle_mod):
    g = torch.Generator().manual_seed(3)
    a = torch.rand((3, 6), generator=g) + 0.1
    b = torch.rand((3, 6), generator=g) + 0.1
    a, b = a / a.sum(1, keepdim=True), b / b.sum(1, keepdim=True)
    cost = torch.cdist(torch.arange(6.0)[:, None], torch.arange(6.0)[:, None])
    la, lb = a.log().cuda(), b.log().cuda()
    lu = torch.zeros_like(la)
    lv = torch.zeros_like(lb)
    r = C.sinkhorn_logiter(la, lb, cost.cuda(), cost.cuda(), lu, lv, 200, 1.5)
    assert r[0] is lu and r[1] is lv
    olu, olv = oracle_mod.sinkhorn_log_hist(a.numpy(), b.numpy(), cost.numpy(), 1.5, 200)
    np.testing.assert_allclose(lu.cpu().numpy(), olu, atol=1e-4)
    np.testing.assert_allclose(lv.cpu().numpy(), olv, atol=1e-4)
