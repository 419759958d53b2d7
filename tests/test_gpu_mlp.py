"""BF16 MFMA GEMM of the warm-start network against torch (fp32 math on the
same bf16-rounded operands): checks the MFMA lane layout, strided operands
(A.B, A.B^T, A^T.B) and the split-K reduction."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu


def _gemm(A, B, transA=False, transB=False):
    from torchmorph_amd import _lib

    L = _lib.lib()
    Ab = A.to(torch.bfloat16).contiguous()
    Bb = B.to(torch.bfloat16).contiguous()
    if transA:  # logical A = Ab^T
        M, K = Ab.shape[1], Ab.shape[0]
        sam, sak = 1, Ab.shape[1]
    else:
        M, K = Ab.shape
        sam, sak = Ab.shape[1], 1
    if transB:
        N = Bb.shape[0]
        sbk, sbn = 1, Bb.shape[1]
    else:
        N = Bb.shape[1]
        sbk, sbn = Bb.shape[1], 1
    C = torch.empty((M, N), device="cuda")
    ws = torch.empty(64 * M * N, device="cuda")
    _lib.check(L.tm_mlp_gemm_bf16(_lib.ptr(Ab), sam, sak, _lib.ptr(Bb), sbk, sbn, _lib.ptr(C),
                                  M, N, K, _lib.ptr(ws), ws.numel(),
                                  _lib.stream_of(A.device)), "gemm")
    Af = Ab.float().T if transA else Ab.float()
    Bf = Bb.float().T if transB else Bb.float()
    return C, Af @ Bf


@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (100, 70, 50), (1000, 256, 128), (7, 3, 5)])
def test_gemm_nn(M, N, K):
    g = torch.Generator(device="cuda").manual_seed(M + N + K)
    A = torch.randn((M, K), device="cuda", generator=g)
    B = torch.randn((K, N), device="cuda", generator=g)
    C, ref = _gemm(A, B)
    torch.testing.assert_close(C, ref, rtol=1e-4, atol=1e-3)


def test_gemm_transposed_operands_and_splitk():
    g = torch.Generator(device="cuda").manual_seed(3)
    A = torch.randn((300, 128), device="cuda", generator=g)
    B = torch.randn((256, 128), device="cuda", generator=g)
    C, ref = _gemm(A, B, transB=True)  # A . B^T
    torch.testing.assert_close(C, ref, rtol=1e-4, atol=1e-3)
    X = torch.randn((20000, 128), device="cuda", generator=g)
    D = torch.randn((20000, 256), device="cuda", generator=g)
    C, ref = _gemm(X, D, transA=True)  # X^T . D, long reduction -> split-K
    torch.testing.assert_close(C, ref, rtol=1e-4, atol=2e-2)
