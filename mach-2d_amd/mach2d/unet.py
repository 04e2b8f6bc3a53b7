"""Learned initial guess for the p' solve: a 4-level U-Net (16-32-64-128
channels, 3x3 convolutions, BF16 weights and activations, FP32
accumulation) whose every layer runs in the library's gfx950 MFMA kernels
(csrc/unet.hip, m2d_unet_*).  torch tensors only hold device memory.

Input (8 channels, NHWC bf16): the Jacobi-scaled system
p - sum(a_nb / a_P) p_nb = b / a_P, i.e. channel 0 = b / (a_P s) with
s = RMS(b / a_P), channels 1..4 = a_W, a_E, a_S, a_N over a_P, channel 5 =
the corner couplings over a_P, 6..7 zero.  Output: p' / s on the grid.  The
network is scale-free: the guess is s times its output.

Layers (name, first input, its mode, skip input, output channels, ReLU):
down path conv pairs at full, 1/2, 1/4 resolution, the 1/8 bottleneck, up
path pairs reading the nearest-upsampled coarser result concatenated with
the skip tensor of the same level, and a 3x3 head with 8 output channels of
which channel 0 is the guess.  Pooling (2x2 average) and upsampling are
fused into the convolution's input reads.

No reference counterpart: the reference always starts MSI from var as given.
"""
from __future__ import annotations

import ctypes as C
import struct

import numpy as np

from . import _lib

POOL, DIRECT, UP = 1, 0, 2

# name, first input (a layer name or "x"), mode, skip input or None, cout, relu
LAYERS = [
    ("e0a", "x", DIRECT, None, 16, True),
    ("e0b", "e0a", DIRECT, None, 16, True),
    ("e1a", "e0b", POOL, None, 32, True),
    ("e1b", "e1a", DIRECT, None, 32, True),
    ("e2a", "e1b", POOL, None, 64, True),
    ("e2b", "e2a", DIRECT, None, 64, True),
    ("b3a", "e2b", POOL, None, 128, True),
    ("b3b", "b3a", DIRECT, None, 128, True),
    ("d2a", "b3b", UP, "e2b", 64, True),
    ("d2b", "d2a", DIRECT, None, 64, True),
    ("d1a", "d2b", UP, "e1b", 32, True),
    ("d1b", "d1a", DIRECT, None, 32, True),
    ("d0a", "d1b", UP, "e0b", 16, True),
    ("d0b", "d0a", DIRECT, None, 16, True),
    ("head", "d0b", DIRECT, None, 8, False),
]
IN_CH = 8
LEVEL = {"e0a": 0, "e0b": 0, "e1a": 1, "e1b": 1, "e2a": 2, "e2b": 2, "b3a": 3, "b3b": 3,
         "d2a": 2, "d2b": 2, "d1a": 1, "d1b": 1, "d0a": 0, "d0b": 0, "head": 0}
MAGIC = b"M2DW"
VERSION = 1


def _r(x, m):
    return (x + m - 1) // m * m


def _check(rc, what):
    if rc != 0:
        raise _lib.SolverError(rc, what)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _bind():
    L = _lib.load()
    vp, i, f, d, lg = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_long
    L.m2d_unet_conv3x3.argtypes = [vp, i, i, vp, i, vp, vp, i, i, i, i, i, i, i, vp, vp, i, vp]
    L.m2d_unet_conv3x3_wgrad.argtypes = [vp, i, i, vp, i, vp, i, i, i, i, i, vp, vp, vp]
    L.m2d_unet_relu_bwd.argtypes = [vp, vp, lg, vp]
    L.m2d_unet_mode_bwd.argtypes = [vp, i, i, i, i, i, i, i, vp, vp]
    L.m2d_unet_to_bf16.argtypes = [vp, vp, lg, vp]
    L.m2d_unet_adam.argtypes = [vp, vp, vp, vp, lg, f, f, f, f, i, vp]
    L.m2d_unet_pack_weights.argtypes = [vp, i, i, i, vp, vp, i, i, vp]
    L.m2d_unet_mse.argtypes = [vp, i, vp, vp, lg, f, vp, i, vp, vp]
    L.m2d_unet_pack_input.argtypes = [vp, vp, i, i, i, i, d, vp, i, vp]
    L.m2d_unet_unpack_output.argtypes = [vp, i, i, i, i, d, vp, vp]
    for name in ("m2d_unet_conv3x3", "m2d_unet_conv3x3_wgrad", "m2d_unet_relu_bwd",
                 "m2d_unet_mode_bwd", "m2d_unet_to_bf16", "m2d_unet_adam",
                 "m2d_unet_pack_weights", "m2d_unet_mse", "m2d_unet_pack_input",
                 "m2d_unet_unpack_output"):
        getattr(L, name).restype = C.c_int
    return L


class Layer:
    def __init__(self, name, src, mode, skip, cin1, cin2, cout, relu):
        self.name, self.src, self.mode, self.skip = name, src, mode, skip
        self.c1, self.c2, self.cout, self.relu = cin1, cin2, cout, relu
        self.cin = cin1 + cin2
        self.kpad = _r(9 * self.cin, 32)
        self.cpad = _r(cout, 16)
        self.kpad_r = _r(9 * cout, 32)       # input-gradient weights
        self.cin_pad = _r(self.cin, 16)


class UNet:
    """The network, its FP32 master weights and Adam state on the device."""

    def __init__(self, seed=0, device="cuda"):
        import torch
        self.torch = torch
        self.dev = device
        self.L = _bind()
        chans = {"x": IN_CH}
        self.layers = []
        for name, src, mode, skip, cout, relu in LAYERS:
            lay = Layer(name, src, mode, skip, chans[src], chans[skip] if skip else 0, cout, relu)
            chans[name] = cout
            self.layers.append(lay)
        g = torch.Generator().manual_seed(seed)
        for lay in self.layers:       # He initialisation on the real (tap, channel) entries
            w = torch.zeros(lay.cpad, lay.kpad)
            std = (2.0 / (9 * lay.cin)) ** 0.5 if lay.relu else (1.0 / (9 * lay.cin)) ** 0.5
            w[:lay.cout, :9 * lay.cin] = torch.randn(lay.cout, 9 * lay.cin, generator=g) * std
            lay.w = w.to(device)
            lay.b = torch.zeros(lay.cout, device=device)
        self._alloc_train_state()
        self.step_count = 0
        self.pack()

    def _alloc_train_state(self):
        t = self.torch
        for lay in self.layers:
            lay.mw, lay.vw = t.zeros_like(lay.w), t.zeros_like(lay.w)
            lay.mb, lay.vb = t.zeros_like(lay.b), t.zeros_like(lay.b)
            lay.gw, lay.gb = t.zeros_like(lay.w), t.zeros_like(lay.b)
            lay.wb = t.zeros(lay.cpad, lay.kpad, dtype=t.bfloat16, device=self.dev)
            lay.wr = t.zeros(lay.cin_pad, lay.kpad_r, dtype=t.bfloat16, device=self.dev)

    def pack(self, stream=None):
        """bf16 copies of the weights (forward and input-gradient forms)."""
        for lay in self.layers:
            _check(self.L.m2d_unet_pack_weights(_ptr(lay.w), lay.cout, lay.cin, lay.kpad,
                                                _ptr(lay.wb), _ptr(lay.wr), lay.kpad_r,
                                                lay.cin_pad, stream), "pack_weights")

    # -- inputs --------------------------------------------------------------
    @staticmethod
    def padded(nx, ny):
        return _r(ny, 8), _r(nx, 8)

    def pack_input(self, a, b, nx, ny, scales, stream=None):
        """Network input for nb systems: a (nb*9*N FP64, reference layout per
        system, consecutive), b (nb*N), scales s (one per system, all equal
        here: the kernel takes one).  Returns the (nb, H, W, 8) bf16 tensor."""
        t = self.torch
        nb = b.numel() // (nx * ny)
        h, w = self.padded(nx, ny)
        x = t.empty(nb, h, w, IN_CH, dtype=t.bfloat16, device=self.dev)
        _check(self.L.m2d_unet_pack_input(_ptr(a), _ptr(b), nx, ny, h, w, 1.0 / scales, _ptr(x),
                                          nb, stream), "pack_input")
        return x

    # -- forward / backward ----------------------------------------------------
    def forward(self, x, stream=None, keep=False):
        """x: (n, H, W, 8) bf16 with H, W multiples of 8.  Returns the head's
        FP32 output (n, H, W, 8); with keep, the activations for backward."""
        t = self.torch
        n, h, w, _ = x.shape
        acts = {"x": x}
        out32 = None
        for lay in self.layers:
            lv = LEVEL[lay.name]
            hh, ww = h >> lv, w >> lv
            src, skip = acts[lay.src], acts[lay.skip] if lay.skip else None
            if lay.name == "head":
                out32 = t.empty(n, hh, ww, lay.cout, dtype=t.float32, device=self.dev)
                out = None
            else:
                out = t.empty(n, hh, ww, lay.cout, dtype=t.bfloat16, device=self.dev)
            _check(self.L.m2d_unet_conv3x3(
                _ptr(src), lay.c1, lay.mode, _ptr(skip), lay.c2, _ptr(lay.wb), _ptr(lay.b),
                lay.cout, lay.cout, lay.kpad, n, hh, ww, 1 if lay.relu else 0, _ptr(out),
                _ptr(out32) if out is None else None, 0, stream), "conv " + lay.name)
            acts[lay.name] = out if out is not None else out32
        self._acts = acts if keep else None
        return out32

    def loss_and_backward(self, x, target, mask, stream=None):
        """One forward and backward pass: loss = 0.5 mean over the mask of
        (out[..., 0] - target)^2; weight gradients into lay.gw / lay.gb."""
        t = self.torch
        n, h, w, _ = x.shape
        out = self.forward(x, stream, keep=True)
        acts = self._acts
        npix = n * h * w
        cnt = float(mask.sum().item())
        head = self.layers[-1]
        g = t.empty(n, h, w, head.cout, dtype=t.bfloat16, device=self.dev)
        partial = t.empty((npix + 255) // 256, dtype=t.float32, device=self.dev)
        _check(self.L.m2d_unet_mse(_ptr(out), head.cout, _ptr(target), _ptr(mask), npix,
                                   1.0 / cnt, _ptr(g), head.cout, _ptr(partial), stream), "mse")
        loss = partial.sum()
        grads32 = {}   # FP32 gradient accumulators of the activations
        gpre = {"head": g}
        for lay in reversed(self.layers):
            lv = LEVEL[lay.name]
            hh, ww = h >> lv, w >> lv
            if lay.name != "head":
                gb = t.empty_like(acts[lay.name])
                _check(self.L.m2d_unet_to_bf16(_ptr(grads32[lay.name]), _ptr(gb), gb.numel(),
                                               stream), "to_bf16")
                if lay.relu:
                    _check(self.L.m2d_unet_relu_bwd(_ptr(gb), _ptr(acts[lay.name]), gb.numel(),
                                                    stream), "relu_bwd")
                gpre[lay.name] = gb
            gl = gpre[lay.name]
            src, skip = acts[lay.src], acts[lay.skip] if lay.skip else None
            lay.gw.zero_()
            lay.gb.zero_()
            _check(self.L.m2d_unet_conv3x3_wgrad(_ptr(src), lay.c1, lay.mode, _ptr(skip), lay.c2,
                                                 _ptr(gl), lay.cout, lay.kpad, n, hh, ww,
                                                 _ptr(lay.gw), _ptr(lay.gb), stream),
                   "wgrad " + lay.name)
            if lay.src == "x":
                continue
            gi = t.empty(n, hh, ww, lay.cin, dtype=t.float32, device=self.dev)
            _check(self.L.m2d_unet_conv3x3(_ptr(gl), lay.cout, DIRECT, None, 0, _ptr(lay.wr),
                                           None, lay.cin, lay.cin, lay.kpad_r, n, hh, ww, 0,
                                           None, _ptr(gi), 0, stream), "dgrad " + lay.name)
            for name, c0, c, mode in ((lay.src, 0, lay.c1, lay.mode),
                                      (lay.skip, lay.c1, lay.c2, DIRECT)):
                if not name:
                    continue
                if name not in grads32:
                    grads32[name] = t.zeros(acts[name].shape, dtype=t.float32, device=self.dev)
                _check(self.L.m2d_unet_mode_bwd(_ptr(gi), lay.cin, c0, c, mode, n, hh, ww,
                                                _ptr(grads32[name]), stream), "mode_bwd")
        self._acts = None
        return loss

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, stream=None):
        self.step_count += 1
        for lay in self.layers:
            for p, m, v, gr in ((lay.w, lay.mw, lay.vw, lay.gw), (lay.b, lay.mb, lay.vb, lay.gb)):
                _check(self.L.m2d_unet_adam(_ptr(p), _ptr(m), _ptr(v), _ptr(gr), p.numel(), lr,
                                            b1, b2, eps, self.step_count, stream), "adam")
        self.pack(stream)

    # -- guesses ---------------------------------------------------------------
    def guess(self, a, b, nx, ny, var, stream=None):
        """Initial guess var (FP64 device vector) for A var = b (device FP64,
        reference layout).  s = RMS of b / a_P over the grid."""
        t = self.torch
        ap = a.view(9, -1)[4]
        q = t.where(ap != 0, b / t.where(ap != 0, ap, t.ones_like(ap)), t.zeros_like(b))
        s = float(t.sqrt((q * q).mean()).item()) or 1.0
        x = self.pack_input(a, b, nx, ny, s, stream)
        out = self.forward(x, stream)
        _check(self.L.m2d_unet_unpack_output(_ptr(out), out.shape[-1], nx, ny, out.shape[2], s,
                                             _ptr(var), stream), "unpack_output")
        return s

    def guess_many(self, a_list, b_list, nx, ny, var_list, stream=None):
        """guess() for several systems of one grid size with one forward pass
        over all of them (the images of a batch are independent, so each
        guess equals guess() of its system alone).  Returns the scales."""
        t = self.torch
        h, w = self.padded(nx, ny)
        n = len(b_list)
        x = t.empty(n, h, w, IN_CH, dtype=t.bfloat16, device=self.dev)
        # the scales of all systems (guess()'s expression, system by system,
        # so the values are the same) with one read back
        sq = []
        for a, b in zip(a_list, b_list):
            ap = a.view(9, -1)[4]
            r = t.where(ap != 0, b / t.where(ap != 0, ap, t.ones_like(ap)), t.zeros_like(b))
            sq.append(t.sqrt((r * r).mean()))
        scales = [float(v) or 1.0 for v in t.stack(sq).cpu()]
        for q, (a, b) in enumerate(zip(a_list, b_list)):
            _check(self.L.m2d_unet_pack_input(_ptr(a), _ptr(b), nx, ny, h, w, 1.0 / scales[q],
                                              _ptr(x[q]), 1, stream), "pack_input")
        out = self.forward(x, stream)
        for q, var in enumerate(var_list):
            _check(self.L.m2d_unet_unpack_output(_ptr(out[q]), out.shape[-1], nx, ny,
                                                 out.shape[2], scales[q], _ptr(var), stream),
                   "unpack_output")
        return scales

    # -- weights file ----------------------------------------------------------
    def save(self, path):
        """Versioned binary weights file: b'M2DW', version, layer count, then
        per layer (cout, cin, kpad) as int32 and the FP32 weights
        [cout][9*cin] (k = tap*cin + ci) and biases, little-endian."""
        with open(path, "wb") as f:
            f.write(MAGIC + struct.pack("<ii", VERSION, len(self.layers)))
            for lay in self.layers:
                f.write(struct.pack("<iii", lay.cout, lay.cin, lay.kpad))
                w = lay.w[:lay.cout, :9 * lay.cin].float().cpu().numpy().astype("<f4")
                f.write(w.tobytes())
                f.write(lay.b.float().cpu().numpy().astype("<f4").tobytes())

    def load(self, path):
        t = self.torch
        with open(path, "rb") as f:
            data = f.read()
        if data[:4] != MAGIC:
            raise ValueError("%s: not a weights file" % path)
        ver, nl = struct.unpack_from("<ii", data, 4)
        if ver != VERSION or nl != len(self.layers):
            raise ValueError("%s: version %d with %d layers, expected %d / %d"
                             % (path, ver, nl, VERSION, len(self.layers)))
        off = 12
        for lay in self.layers:
            cout, cin, _ = struct.unpack_from("<iii", data, off)
            off += 12
            if (cout, cin) != (lay.cout, lay.cin):
                raise ValueError("%s: layer %s shape mismatch" % (path, lay.name))
            nw = cout * 9 * cin
            w = np.frombuffer(data, "<f4", nw, off).reshape(cout, 9 * cin)
            off += 4 * nw
            b = np.frombuffer(data, "<f4", cout, off)
            off += 4 * cout
            lay.w.zero_()
            lay.w[:cout, :9 * cin] = t.from_numpy(w.copy()).to(self.dev)
            lay.b.copy_(t.from_numpy(b.copy()))
        self.pack()
