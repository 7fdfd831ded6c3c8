"""Species-rate surrogate: an MLP trained and evaluated by the project's own
HIP kernels (``csrc/surrogate_kernels.hip``, C ABI in
``include/pyjac_surrogate.h``), with a per-state error gate that falls back
to the FP64 species-rate kernel.

There is no counterpart in the reference.  Shapes follow the workload the
surrogate is built for (GRI-Mech 3.0): inputs log T, log P and log mole
fractions of the Ns species (55 for GRI), four hidden layers of 512 SiLU
units, outputs normalised per state -- wdot_k / S for the Ns species with
S = max_k |wdot_k|, an error head predicting log10 of the state's relative
error ``max_k |wdot_k - wdot_ref_k| / max_k |wdot_ref_k|``, and the
standardised log10 S.

BF16 operands on the MFMA units, FP32 accumulation, FP32 master weights and
Adam moments.  PyTorch provides device memory, the per-epoch permutation and
the gather / scatter of rejected states; every product, activation, loss,
gradient and optimizer update is one of the kernels above.
"""

import ctypes
import math
import os

import numpy as np

from ..core.exceptions import HipRuntimeError
from ..pywrap import load_library

RU = 8314.4621
HIDDEN = 512
N_HIDDEN = 4
OUT_PAD = 128          # outputs padded to the GEMM tile
EPI_HIDDEN, EPI_LINEAR, EPI_DSILU, EPI_F32 = 0, 1, 2, 3


def _torch():
    import torch
    return torch


def _lib():
    L = load_library()
    if getattr(L, '_sur_ready', False):
        return L
    P, I, Lg, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_float
    L.pyjac_sur_gemm.argtypes = [I, P, I, P, I, I, I, I, P, P, I, P, I, P, I, P, I, P]
    L.pyjac_sur_featurize.argtypes = [P, P, Lg, I, I, I, P, P, I, Lg, P, P, P]
    L.pyjac_sur_targets.argtypes = [P, P, P, Lg, I, I, I, P, P, I, Lg, P, P]
    L.pyjac_sur_err_target.argtypes = [P, P, Lg, I, I, P, P]
    L.pyjac_sur_mse_grad.argtypes = [P, P, Lg, Lg, I, I, F, P, P, P, P]
    L.pyjac_sur_colsum.argtypes = [P, Lg, I, P, P]
    L.pyjac_sur_adam.argtypes = [P, P, P, P, Lg, I, I, F, F, F, F, F, F, F, P, P, P]
    L.pyjac_sur_to_bf16.argtypes = [P, Lg, I, I, P, P, P]
    L.pyjac_sur_decode.argtypes = [P, I, Lg, I, I, I, P, P, P, F, P, P, P]
    L.pyjac_sur_gather.argtypes = [P, P, Lg, I, I, P, Lg, P, P, P]
    L.pyjac_sur_scatter.argtypes = [P, P, P, Lg, I, I, I, P, P, P, Lg, P]
    L._sur_ready = True
    return L


def _p(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _ok(st, what):
    if st != 0:
        raise HipRuntimeError(st, '{} failed ({})'.format(what, st))


def pad_to(x, m):
    return (x + m - 1) // m * m


class MLP(object):
    """Weights, optimizer state and the forward / backward passes.

    ``dims`` = (inputs, hidden..., outputs) of the real network; storage is
    padded (inputs to 32, hidden to 64, outputs to 128) with zeros that stay
    zero through training.
    """

    def __init__(self, dims, device='cuda:0', seed=0):
        torch = _torch()
        self.dims = tuple(int(d) for d in dims)
        self.device = torch.device(device)
        n_in = pad_to(self.dims[0], 64)
        hid = [pad_to(d, 128) for d in self.dims[1:-1]]
        self.Kp = [n_in] + hid
        self.Np = hid + [pad_to(self.dims[-1], OUT_PAD)]
        g = torch.Generator().manual_seed(seed)
        self.W, self.b = [], []
        for l, (K, N) in enumerate(zip(self.Kp, self.Np)):
            k_real, n_real = self.dims[l], self.dims[l + 1]
            w = torch.zeros(N, K, dtype=torch.float32)
            bound = math.sqrt(6.0 / k_real) if l < len(self.Np) - 1 else math.sqrt(3.0 / k_real)
            w[:n_real, :k_real] = (torch.rand(n_real, k_real, generator=g) * 2 - 1) * bound
            self.W.append(w.to(self.device))
            self.b.append(torch.zeros(N, dtype=torch.float32, device=self.device))
        self._alloc_state()

    def _alloc_state(self):
        torch = _torch()
        self.m1 = [torch.zeros_like(w) for w in self.W] + [torch.zeros_like(b) for b in self.b]
        self.m2 = [torch.zeros_like(w) for w in self.W] + [torch.zeros_like(b) for b in self.b]
        self.Wb = [torch.empty(w.shape, dtype=torch.int16, device=self.device) for w in self.W]
        self.Wbt = [torch.empty((w.shape[1], w.shape[0]), dtype=torch.int16, device=self.device)
                    for w in self.W]
        self.t = 0
        self.refresh_bf16()

    def refresh_bf16(self):
        L = _lib()
        for w, wb, wbt in zip(self.W, self.Wb, self.Wbt):
            _ok(L.pyjac_sur_to_bf16(_p(w), w.numel(), w.shape[0], w.shape[1], _p(wb), _p(wbt),
                                    _stream()), 'to_bf16')

    # --------------------------------------------------------------- forward
    def forward(self, x, M, keep=False, xt=None):
        """x: bf16 [M][Kp0] (int16 storage).  Returns fp32 [M][Np_last].
        With keep=True the activations (and their transposes; xt is the
        transposed input) are kept for backward()."""
        torch = _torch()
        L = _lib()
        s = _stream()
        self._acts = [(x, xt)]
        self._z = []
        a = x
        nl = len(self.W)
        for l in range(nl):
            K, N = self.Kp[l], self.Np[l]
            if l < nl - 1:
                an = torch.empty((M, N), dtype=torch.int16, device=self.device)
                at = torch.empty((N, M), dtype=torch.int16, device=self.device) if keep else None
                z = torch.empty((M, N), dtype=torch.float32, device=self.device) if keep else None
                _ok(L.pyjac_sur_gemm(EPI_HIDDEN, _p(a), K, _p(self.Wb[l]), K, M, N, K,
                                     _p(self.b[l]), _p(z), N, _p(an), N, _p(at), M, None, 0, s),
                    'gemm')
                if keep:
                    self._acts.append((an, at))
                    self._z.append(z)
                a = an
            else:
                out = torch.empty((M, N), dtype=torch.float32, device=self.device)
                _ok(L.pyjac_sur_gemm(EPI_LINEAR, _p(a), K, _p(self.Wb[l]), K, M, N, K,
                                     _p(self.b[l]), _p(out), N, None, 0, None, 0, None, 0, s),
                    'gemm')
        return out

    # -------------------------------------------------------------- backward
    def backward(self, dz, dzt, M):
        """dz, dzt: bf16 gradient of the outputs [M][Np_last] / [Np_last][M].
        Returns the fp32 gradients (dW list, db list)."""
        torch = _torch()
        L = _lib()
        s = _stream()
        nl = len(self.W)
        dW = [None] * nl
        db = [None] * nl
        for l in range(nl - 1, -1, -1):
            K, N = self.Kp[l], self.Np[l]
            a_prev, a_prev_t = self._acts[l]
            g = torch.empty((N, K), dtype=torch.float32, device=self.device)
            _ok(L.pyjac_sur_gemm(EPI_F32, _p(dzt), M, _p(a_prev_t), M, N, K, M, None, _p(g), K,
                                 None, 0, None, 0, None, 0, s), 'gemm dW')
            dW[l] = g
            gb = torch.empty(N, dtype=torch.float32, device=self.device)
            _ok(L.pyjac_sur_colsum(_p(dzt), M, N, _p(gb), s), 'colsum')
            db[l] = gb
            if l > 0:
                dzp = torch.empty((M, K), dtype=torch.int16, device=self.device)
                dzpt = torch.empty((K, M), dtype=torch.int16, device=self.device)
                _ok(L.pyjac_sur_gemm(EPI_DSILU, _p(dz), N, _p(self.Wbt[l]), N, M, K, N, None,
                                     None, 0, _p(dzp), K, _p(dzpt), M, _p(self._z[l - 1]), K, s),
                    'gemm dA')
                dz, dzt = dzp, dzpt
        return dW, db

    def adam_step(self, dW, db, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        L = _lib()
        s = _stream()
        self.t += 1
        c1 = 1.0 / (1.0 - b1 ** self.t)
        c2 = 1.0 / (1.0 - b2 ** self.t)
        nl = len(self.W)
        for l in range(nl):
            w, g = self.W[l], dW[l]
            _ok(L.pyjac_sur_adam(_p(w), _p(g), _p(self.m1[l]), _p(self.m2[l]), w.numel(),
                                 w.shape[0], w.shape[1], lr, b1, b2, eps, c1, c2, wd,
                                 _p(self.Wb[l]), _p(self.Wbt[l]), s), 'adam')
            b = self.b[l]
            _ok(L.pyjac_sur_adam(_p(b), _p(db[l]), _p(self.m1[nl + l]), _p(self.m2[nl + l]),
                                 b.numel(), 1, b.numel(), lr, b1, b2, eps, c1, c2, 0.0, None,
                                 None, s), 'adam')


class Surrogate(object):
    """The trained network for one mechanism and option set, with feature
    normalisation, output scales and the calibrated error gate."""

    def __init__(self, mech, conp=False, order='F', device='cuda:0', seed=0):
        self.mech = mech
        self.conp = bool(conp)
        self.order = order
        self.Ns = mech.Ns
        self.n_in = mech.Ns + 2
        self.n_out = mech.Ns + 2          # species, error head, log10 S
        self.net = MLP((self.n_in,) + (HIDDEN,) * N_HIDDEN + (self.n_out,), device, seed)
        self.device = self.net.device
        self.mu = self.sd = self.wscale = None
        self.log10_tol = -3.0
        mw = np.asarray(mech.mw, dtype=np.float64)
        self.mw_ratio = _torch().tensor(mw[:-1] / mw[-1], dtype=torch_f64(), device=self.device)

    # ------------------------------------------------------------ statistics
    def fit_scales(self, param, phi, dphi):
        """Feature mean / scale and per-species output scales from a host
        sample (C order arrays) of states and their FP64 rates."""
        conp, Ns = self.conp, self.Ns
        T = phi[:, 0]
        P = param if conp else phi[:, 1]
        V = phi[:, 1] if conp else param
        n_k = np.maximum(phi[:, 2:], 1e-300)
        ntot = P * V / (RU * T)
        n_ns = np.maximum(0.0, ntot - n_k.sum(axis=1))
        X = np.column_stack([n_k, n_ns]) / ntot[:, None]
        feats = np.column_stack([np.log(T), np.log(P), np.log(X + 1e-12)])
        torch = _torch()
        ld = self.net.Kp[0]
        mu = np.zeros(ld, dtype=np.float32)
        sd = np.ones(ld, dtype=np.float32)
        mu[:feats.shape[1]] = feats.mean(axis=0)
        sd[:feats.shape[1]] = np.maximum(feats.std(axis=0), 1e-6)
        w = dphi[:, 2:] / V[:, None]
        mw = np.asarray(self.mech.mw)
        w_ns = -(w * (mw[:-1] / mw[-1])).sum(axis=1)
        S = np.maximum(np.abs(np.column_stack([w, w_ns])).max(axis=1), 1e-300)
        lS = np.log10(S)
        stat = np.zeros(OUT_PAD, np.float32)
        stat[0], stat[1] = lS.mean(), max(lS.std(), 1e-3)
        self.mu = torch.tensor(mu, device=self.device)
        self.sd = torch.tensor(sd, device=self.device)
        self.wscale = torch.tensor(stat, device=self.device)

    # --------------------------------------------------------------- kernels
    def features(self, param, phi, n, colmajor, transposed=False):
        torch = _torch()
        M = pad_to(max(n, 1), 128)
        ld = self.net.Kp[0]
        x = torch.empty((M, ld), dtype=torch.int16, device=self.device)
        xt = torch.empty((ld, M), dtype=torch.int16, device=self.device) if transposed else None
        _ok(_lib().pyjac_sur_featurize(_p(param), _p(phi), n, self.mech.NN, int(self.conp),
                                       int(colmajor), _p(self.mu), _p(self.sd), ld, M, _p(x),
                                       _p(xt), _stream()), 'featurize')
        return x, xt, M

    def targets(self, param, phi, dphi, n, colmajor):
        torch = _torch()
        M = pad_to(max(n, 1), 128)
        y = torch.empty((M, OUT_PAD), dtype=torch.float32, device=self.device)
        _ok(_lib().pyjac_sur_targets(_p(param), _p(phi), _p(dphi), n, self.mech.NN,
                                     int(self.conp), int(colmajor), _p(self.wscale),
                                     _p(self.mw_ratio), OUT_PAD, M, _p(y), _stream()),
            'targets')
        return y

    def train_step(self, param, phi, y, n, lr, wd=0.0):
        """One step on a batch (C order, n states, targets y [M][128]).
        Returns the loss as a device scalar."""
        torch = _torch()
        L = _lib()
        x, xt, M = self.features(param, phi, n, False, transposed=True)
        out = self.net.forward(x, M, keep=True, xt=xt)
        _ok(L.pyjac_sur_err_target(_p(out), _p(y), n, OUT_PAD, self.Ns, _p(self.wscale),
                                   _stream()), 'err_target')
        dz = torch.empty((M, OUT_PAD), dtype=torch.int16, device=self.device)
        dzt = torch.empty((OUT_PAD, M), dtype=torch.int16, device=self.device)
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        _ok(L.pyjac_sur_mse_grad(_p(out), _p(y), n, M, OUT_PAD, self.n_out,
                                 1.0 / (n * self.n_out), _p(dz), _p(dzt), _p(loss), _stream()),
            'mse_grad')
        dW, db = self.net.backward(dz, dzt, M)
        self.net.adam_step(dW, db, lr, wd=wd)
        return loss

    def predict(self, param, phi, n=None):
        """Surrogate production rates of n states already on the device, in
        the handle's order: returns (wdot [n][Ns] or [Ns][n], accept [n])."""
        torch = _torch()
        colmajor = self.order == 'F'
        n = int(param.numel()) if n is None else n
        x, _, M = self.features(param, phi, n, colmajor)
        out = self.net.forward(x, M)
        shape = (self.Ns, n) if colmajor else (n, self.Ns)
        wdot = torch.empty(shape, dtype=torch.float64, device=self.device)
        accept = torch.empty(n, dtype=torch.int32, device=self.device)
        _ok(_lib().pyjac_sur_decode(_p(out), OUT_PAD, n, self.mech.NN, int(self.conp),
                                    int(colmajor), _p(param), _p(phi), _p(self.wscale),
                                    float(self.log10_tol), _p(wdot), _p(accept), _stream()),
            'decode')
        return wdot, accept

    # ------------------------------------------------------------ save / load
    def state_dict(self):
        d = {'mu': self.mu, 'sd': self.sd, 'wscale': self.wscale,
             'log10_tol': _torch().tensor([self.log10_tol])}
        for l, (w, b) in enumerate(zip(self.net.W, self.net.b)):
            d['W%d' % l] = w
            d['b%d' % l] = b
        return d

    def save(self, path):
        from safetensors.torch import save_file
        save_file({k: v.detach().cpu().contiguous() for k, v in self.state_dict().items()}, path)

    def load(self, path):
        from safetensors.torch import load_file
        d = load_file(path)
        self.mu = d['mu'].to(self.device)
        self.sd = d['sd'].to(self.device)
        self.wscale = d['wscale'].to(self.device)
        self.log10_tol = float(d['log10_tol'][0])
        for l in range(len(self.net.W)):
            self.net.W[l] = d['W%d' % l].to(self.device)
            self.net.b[l] = d['b%d' % l].to(self.device)
        self.net._alloc_state()
        return self


def torch_f64():
    return _torch().float64


def gather_states(param, phi, n, NN, colmajor, idx):
    """Compact copies (same layout) of the states idx (int64 device tensor)."""
    torch = _torch()
    m = int(idx.numel())
    p2 = torch.empty(m, dtype=torch.float64, device=param.device)
    f2 = torch.empty((NN, m) if colmajor else (m, NN), dtype=torch.float64, device=phi.device)
    _ok(_lib().pyjac_sur_gather(_p(param), _p(phi), n, NN, int(colmajor), _p(idx), m, _p(p2),
                                _p(f2), _stream()), 'gather')
    return p2, f2


def scatter_rates(param_c, phi_c, dphi_c, m, NN, conp, colmajor, mw_ratio, idx, wdot, n):
    """wdot of the Ns species of the compact states (from the FP64 kernel's
    dphi; bath gas by mass conservation) into rows idx of wdot."""
    _ok(_lib().pyjac_sur_scatter(_p(param_c), _p(phi_c), _p(dphi_c), m, NN, int(conp),
                                 int(colmajor), _p(mw_ratio), _p(idx), _p(wdot), n, _stream()),
        'scatter')
