"""Surrogate kernels on the GPU: the BF16 MFMA product and its epilogues
against numpy, backward gradients against the CPU FP64 MLP (tests/mlp_ref.py,
itself checked by finite differences), training that lowers the loss, and,
when a trained model is present, the error gate and the speed-up over the
FP64 kernel."""

import os

import numpy as np
import pytest

import mlp_ref

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')


def _bits(a):
    """float array -> bf16 bit patterns (int16 tensor on the GPU) and the
    rounded values as float64."""
    r = mlp_ref.bf16(a)
    b = (r.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16).view(np.int16)
    return torch.from_numpy(b.copy()).cuda(), r


@pytest.mark.parametrize('M,N,K', [(128, 64, 32), (256, 128, 96), (384, 192, 160)])
def test_gemm_matches_numpy(M, N, K):
    from pyjac.surrogate import _lib, _p, _stream
    rng = np.random.default_rng(M + N + K)
    A, Ar = _bits(rng.normal(size=(M, K)))
    B, Br = _bits(rng.normal(size=(N, K)))
    C = torch.empty((M, N), dtype=torch.float32, device='cuda')
    assert _lib().pyjac_sur_gemm(3, _p(A), K, _p(B), K, M, N, K, None, _p(C), N, None, 0, None,
                                 0, None, 0, _stream()) == 0
    torch.cuda.synchronize()
    ref = Ar @ Br.T
    err = np.abs(C.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < 1e-5, err


def test_gemm_epilogues():
    from pyjac.surrogate import _lib, _p, _stream
    M, N, K = 256, 128, 64
    rng = np.random.default_rng(3)
    A, Ar = _bits(rng.normal(size=(M, K)))
    B, Br = _bits(rng.normal(size=(N, K)) * 0.3)
    bias = rng.normal(size=N).astype(np.float32)
    tb = torch.from_numpy(bias).cuda()
    z = torch.empty((M, N), dtype=torch.float32, device='cuda')
    a = torch.empty((M, N), dtype=torch.int16, device='cuda')
    at = torch.empty((N, M), dtype=torch.int16, device='cuda')
    L = _lib()
    assert L.pyjac_sur_gemm(0, _p(A), K, _p(B), K, M, N, K, _p(tb), _p(z), N, _p(a), N, _p(at), M,
                            None, 0, _stream()) == 0
    torch.cuda.synchronize()
    zr = Ar @ Br.T + bias
    np.testing.assert_allclose(z.cpu().numpy(), zr, rtol=1e-5, atol=1e-5)
    av = a.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    av = av.view(np.float32).astype(np.float64)
    np.testing.assert_allclose(av, mlp_ref.silu(zr), rtol=1e-2, atol=1e-2)
    assert np.array_equal(a.cpu().numpy(), at.cpu().numpy().T)
    # derivative epilogue
    d = torch.empty((M, N), dtype=torch.int16, device='cuda')
    dt = torch.empty((N, M), dtype=torch.int16, device='cuda')
    assert L.pyjac_sur_gemm(2, _p(A), K, _p(B), K, M, N, K, None, None, 0, _p(d), N, _p(dt), M,
                            _p(z), N, _stream()) == 0
    torch.cuda.synchronize()
    dv = (d.cpu().numpy().view(np.uint16).astype(np.uint32) << 16).view(np.float32)
    np.testing.assert_allclose(dv, (Ar @ Br.T) * mlp_ref.dsilu(z.cpu().numpy().astype(np.float64)),
                               rtol=1e-2, atol=1e-2)
    assert np.array_equal(d.cpu().numpy(), dt.cpu().numpy().T)


def _tiny_net():
    from pyjac.surrogate import MLP
    return MLP((55, 100, 120, 54), device='cuda:0', seed=4)


def test_backward_matches_cpu_fp64():
    """Gradients of the HIP backward pass vs the CPU FP64 MLP with the same
    bf16-rounded weights and inputs (BF16 activations and gradients on the
    GPU: 2 % relative Frobenius error per tensor)."""
    from pyjac.surrogate import _lib, _p, _stream
    net = _tiny_net()
    n, M, nout = 100, 128, 54
    rng = np.random.default_rng(7)
    xv = np.zeros((M, net.Kp[0]))
    xv[:n, :55] = rng.normal(size=(n, 55))
    x, xr = _bits(xv)
    xt = x.t().contiguous()
    yv = np.zeros((M, net.Np[-1]), dtype=np.float32)
    yv[:n, :nout] = rng.normal(size=(n, nout))
    y = torch.from_numpy(yv).cuda()
    out = net.forward(x, M, keep=True, xt=xt)
    dz = torch.empty((M, net.Np[-1]), dtype=torch.int16, device='cuda')
    dzt = torch.empty((net.Np[-1], M), dtype=torch.int16, device='cuda')
    loss = torch.zeros(1, dtype=torch.float32, device='cuda')
    assert _lib().pyjac_sur_mse_grad(_p(out), _p(y), n, M, net.Np[-1], nout, 1.0 / (n * nout),
                                     _p(dz), _p(dzt), _p(loss), _stream()) == 0
    dW, db = net.backward(dz, dzt, M)
    torch.cuda.synchronize()
    Wr = [mlp_ref.bf16(w.cpu().numpy()) for w in net.W]
    br = [b.cpu().numpy().astype(np.float64) for b in net.b]
    lref, gW, gb = mlp_ref.loss_and_grads(Wr, br, xr, yv.astype(np.float64), n, nout)
    assert abs(loss.item() - lref) <= 2e-2 * lref
    for l in range(len(Wr)):
        g = dW[l].cpu().numpy()
        e = np.linalg.norm(g - gW[l]) / np.linalg.norm(gW[l])
        assert e < 2e-2, (l, e)
        e = np.linalg.norm(db[l].cpu().numpy() - gb[l]) / np.linalg.norm(gb[l])
        assert e < 2e-2, (l, e)


def test_adam_matches_formula():
    from pyjac.surrogate import _lib, _p, _stream
    rng = np.random.default_rng(9)
    w0 = rng.normal(size=(4, 64)).astype(np.float32)
    g = rng.normal(size=(4, 64)).astype(np.float32)
    w = torch.from_numpy(w0.copy()).cuda()
    tg = torch.from_numpy(g).cuda()
    m1 = torch.zeros_like(w)
    m2 = torch.zeros_like(w)
    wb = torch.empty((4, 64), dtype=torch.int16, device='cuda')
    wbt = torch.empty((64, 4), dtype=torch.int16, device='cuda')
    lr, b1, b2, eps, wd = 1e-2, 0.9, 0.999, 1e-8, 1e-3
    assert _lib().pyjac_sur_adam(_p(w), _p(tg), _p(m1), _p(m2), w.numel(), 4, 64, lr, b1, b2, eps,
                                 1 / (1 - b1), 1 / (1 - b2), wd, _p(wb), _p(wbt), _stream()) == 0
    torch.cuda.synchronize()
    a = (1 - b1) * g
    v = (1 - b2) * g * g
    ref = w0 - lr * (wd * w0 + (a / (1 - b1)) / (np.sqrt(v / (1 - b2)) + eps))
    np.testing.assert_allclose(w.cpu().numpy(), ref, rtol=1e-5, atol=1e-6)
    assert np.array_equal(wb.cpu().numpy(), wbt.cpu().numpy().T)


def test_training_lowers_the_loss():
    """A few epochs on GRI-3.0 states cut the training loss several-fold."""
    from pyjac._paths import mechanism_files
    from pyjac.core.mech_tables import MechTables
    from pyjac.surrogate import Surrogate
    from pyjac.surrogate.train import train
    mech = MechTables.from_file(*mechanism_files('gri30'))
    sur = Surrogate(mech, conp=False, order='F', seed=1)
    hist = train(sur, 131072, epochs=8, batch=4096, lr=2e-3, seed=5, log=None)
    assert all(np.isfinite(hist))
    assert hist[-1] < 0.3 * hist[0], hist


MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                     'pyjac-v2_amd', 'pyjac', 'surrogate', 'gri30_conv.safetensors')


@pytest.mark.skipif(not os.path.exists(MODEL), reason='no trained model in the tree')
def test_trained_model_gate():
    """The committed model at the 1e-3 tolerance, the network always run:
    every state the gate lets through is within 1e-3 of the FP64 kernel and
    the rest are the FP64 kernel's own result; the accepted fraction and the
    speed-up over the FP64 kernel alone are reported."""
    from pyjac._paths import mechanism_files
    from pyjac.core.mech_tables import MechTables
    from pyjac.surrogate import Surrogate
    from pyjac.surrogate.train import evaluate
    mech = MechTables.from_file(*mechanism_files('gri30'))
    sur = Surrogate(mech, conp=False, order='F').load(MODEL)
    res = evaluate(sur, 262144, seed=4242, tol=1e-3)
    print('surrogate at 1e-3: accepted {:.4f}, speed-up {:.3f}'.format(
        res['accepted_fraction'], res['speedup']))
    assert res['max_error_all'] <= 1e-3, res


@pytest.mark.parametrize('order', ['F', 'C'])
def test_gate_mixes_network_and_fp64(order):
    """The gated path with about half of the states accepted: the gate's
    threshold set to the median of the error head's outputs.  Accepted
    states carry the network's production rates (sur.predict), rejected
    states the FP64 kernel's (gathered, evaluated and scattered back,
    bath gas by mass conservation), bit for bit."""
    from pyjac._paths import mechanism_files
    from pyjac.core.mech_tables import MechTables
    from pyjac.surrogate import Surrogate
    from pyjac.surrogate.train import (_head, fp64_rates, host_states, production_rates,
                                       rates_with_fallback, train)
    mech = MechTables.from_file(*mechanism_files('gri30'))
    sur = Surrogate(mech, conp=False, order=order, seed=2)
    if os.path.exists(MODEL):
        sur.load(MODEL)
    else:
        train(sur, 65536, epochs=2, batch=4096, lr=2e-3, seed=6, log=None)
    n = 5000
    colmajor = order == 'F'
    param_h, phi_h = host_states(mech, n, 77, False)
    param = torch.from_numpy(param_h).cuda()
    phi = torch.from_numpy(np.ascontiguousarray(phi_h.T) if colmajor else phi_h).cuda()
    head = _head(sur, param, phi, n, colmajor)
    sur.log10_tol = float(torch.median(head).item())
    w_net, acc_net = sur.predict(param, phi, n)
    w, accept = rates_with_fallback(sur, param, phi)
    assert torch.equal(accept, acc_net)
    frac = float(accept.float().mean().item())
    assert 0.3 < frac < 0.7, frac
    ref = production_rates(sur, param, phi, fp64_rates(mech, False, param, phi, order),
                           colmajor)
    if colmajor:
        w, w_net, ref = w.t(), w_net.t(), ref.t()
    a = accept.bool()
    assert torch.equal(w[a], w_net[a])
    assert torch.equal(w[~a], ref[~a])
    # mass conservation of the rejected states' rates (bath gas included)
    mw = torch.tensor(mech.mw, dtype=torch.float64, device='cuda')
    bal = (w[~a] * mw).sum(dim=1).abs() / (w[~a].abs() * mw).sum(dim=1)
    assert float(bal.max()) < 1e-12
