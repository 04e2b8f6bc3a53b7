"""GPU numerics of the U-Net kernels (csrc/unet.hip) against a plain
PyTorch FP32 reference of the same operations, computed on the CPU from the
same bf16-valued inputs and weights.

Bar: the kernels accumulate in FP32 on the matrix cores, in a different
order than the reference, so results agree to FP32 rounding of the sums
(relative 1e-4 of the output scale for one layer); whole-network results,
whose activations are rounded to bf16 between layers, agree to 2e-2 of the
output scale (a bf16 rounding flip of one activation moves the rest)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bf(t):
    import torch
    return t.to(torch.bfloat16).to(torch.float32)


def _ref_conv(x_nhwc, w_ck, b, cin, cout, relu):
    """x (n, h, w, cin) fp32, w [cout][9*cin] (k = tap*cin + ci) -> NHWC."""
    import torch
    import torch.nn.functional as F
    wt = w_ck.reshape(cout, 3, 3, cin).permute(0, 3, 1, 2).contiguous()
    y = F.conv2d(x_nhwc.permute(0, 3, 1, 2), wt, b, padding=1).permute(0, 2, 3, 1)
    return torch.relu(y) if relu else y


def _effective(x1, mode, x2):
    """The convolution's input: x1 through its mode, then x2's channels."""
    import torch
    import torch.nn.functional as F
    if mode == 1:
        x1 = _pool(x1)
    elif mode == 2:
        x1 = F.interpolate(x1.permute(0, 3, 1, 2), scale_factor=2,
                           mode="nearest").permute(0, 2, 3, 1)
    return torch.cat([x1, x2], dim=3) if x2 is not None else x1


def _pool(x):
    n, h, w, c = x.shape
    t = x.reshape(n, h // 2, 2, w // 2, 2, c)
    s = (t[:, :, 0, :, 0] + t[:, :, 0, :, 1]) + (t[:, :, 1, :, 0] + t[:, :, 1, :, 1])
    return _bf(s * 0.25)


def _dev(t, dtype=None):
    import torch
    return t.to(dtype or torch.bfloat16).to("cuda").contiguous()


def _conv_gpu(L, x1, c1, mode, x2, c2, wb, bias, cout, kpad, n, h, w, relu):
    import ctypes as C

    import torch
    out = torch.empty(n, h, w, cout, dtype=torch.float32, device="cuda")
    rc = L.m2d_unet_conv3x3(C.c_void_p(x1.data_ptr()), c1, mode,
                            C.c_void_p(x2.data_ptr()) if x2 is not None else None, c2,
                            C.c_void_p(wb.data_ptr()),
                            C.c_void_p(bias.data_ptr()) if bias is not None else None,
                            cout, cout, kpad, n, h, w, relu, None, C.c_void_p(out.data_ptr()), 0,
                            None)
    assert rc == 0
    torch.cuda.synchronize()
    return out.cpu()


CONV_CASES = [  # n, h, w, c1, mode, c2, cout, relu
    (2, 16, 24, 8, 0, 0, 16, 1), (1, 8, 8, 16, 1, 0, 32, 1), (2, 8, 16, 32, 2, 16, 16, 1),
    (1, 12, 20, 64, 0, 0, 128, 0), (3, 4, 8, 16, 0, 0, 8, 0)]


@pytest.mark.parametrize("case", CONV_CASES)
def test_conv3x3_matches_torch(gpu, case):
    import torch

    from mach2d import unet as U
    L = U._bind()
    n, h, w, c1, mode, c2, cout, relu = case
    g = torch.Generator().manual_seed(sum(case))
    sh1 = {0: (n, h, w, c1), 1: (n, 2 * h, 2 * w, c1), 2: (n, h // 2, w // 2, c1)}[mode]
    x1 = _bf(torch.randn(*sh1, generator=g))
    x2 = _bf(torch.randn(n, h, w, c2, generator=g)) if c2 else None
    cin = c1 + c2
    kpad = (9 * cin + 31) // 32 * 32
    cpad = (cout + 15) // 16 * 16
    wf = torch.zeros(cpad, kpad)
    wf[:cout, :9 * cin] = _bf(torch.randn(cout, 9 * cin, generator=g) * 0.2)
    b = torch.randn(cout, generator=g)
    got = _conv_gpu(L, _dev(x1), c1, mode, _dev(x2) if c2 else None, c2, _dev(wf),
                    _dev(b, torch.float32), cout, kpad, n, h, w, relu)
    xe = _effective(x1, mode, x2 if c2 else None)
    want = _ref_conv(xe, wf[:cout, :9 * cin], b, cin, cout, relu)
    scale = want.abs().max().item()
    assert torch.allclose(got, want, atol=1e-4 * scale, rtol=0), (case, (got - want).abs().max())


@pytest.mark.parametrize("case", CONV_CASES[:4])
def test_conv3x3_gradients_match_autograd(gpu, case):
    """wgrad kernel, input gradient (rotated weights through the forward
    kernel) and the gradient of the fused pool / upsampling against autograd
    of the FP32 reference."""
    import ctypes as C

    import torch

    from mach2d import unet as U
    L = U._bind()
    n, h, w, c1, mode, c2, cout, _ = case
    g = torch.Generator().manual_seed(7 + sum(case))
    sh1 = {0: (n, h, w, c1), 1: (n, 2 * h, 2 * w, c1), 2: (n, h // 2, w // 2, c1)}[mode]
    x1 = _bf(torch.randn(*sh1, generator=g))
    x2 = _bf(torch.randn(n, h, w, c2, generator=g)) if c2 else None
    cin = c1 + c2
    kpad, cpad = (9 * cin + 31) // 32 * 32, (cout + 15) // 16 * 16
    kpad_r, cin_pad = (9 * cout + 31) // 32 * 32, (cin + 15) // 16 * 16
    wf = torch.zeros(cpad, kpad)
    wf[:cout, :9 * cin] = _bf(torch.randn(cout, 9 * cin, generator=g) * 0.2)
    G = _bf(torch.randn(n, h, w, cout, generator=g))
    # reference: loss = sum(conv(x) * G), no bias / relu (the kernels take G
    # as the pre-activation gradient)
    x1r = x1.clone().requires_grad_(True)
    x2r = x2.clone().requires_grad_(True) if c2 else None
    wr_ = wf[:cout, :9 * cin].clone().requires_grad_(True)
    if mode == 1:
        n_, hh, ww, cc = x1r.shape
        t = x1r.reshape(n_, hh // 2, 2, ww // 2, 2, cc)
        pooled = ((t[:, :, 0, :, 0] + t[:, :, 0, :, 1]) + (t[:, :, 1, :, 0] + t[:, :, 1, :, 1])) * 0.25
        # the kernels read the pooled value rounded to bf16; its gradient is the pool's
        xe = _bf(pooled).detach() + (pooled - pooled.detach())
    else:
        xe = _effective(x1r, mode, None)
    if c2:
        xe = torch.cat([xe, x2r], dim=3)
    y = _ref_conv(xe, wr_, None, cin, cout, False)
    (y * G).sum().backward()
    # GPU: weight and bias gradients
    dw = torch.zeros(cpad, kpad, device="cuda")
    db = torch.zeros(cout, device="cuda")
    dx1, dx2, dG = _dev(x1), _dev(x2) if c2 else None, _dev(G)
    rc = L.m2d_unet_conv3x3_wgrad(C.c_void_p(dx1.data_ptr()), c1, mode,
                                  C.c_void_p(dx2.data_ptr()) if c2 else None, c2,
                                  C.c_void_p(dG.data_ptr()), cout, kpad, n, h, w,
                                  C.c_void_p(dw.data_ptr()), C.c_void_p(db.data_ptr()), None)
    assert rc == 0
    # input gradient: forward kernel on G with the rotated weights
    wfd = _dev(wf, torch.float32)
    wb = torch.empty(cpad, kpad, dtype=torch.bfloat16, device="cuda")
    wrot = torch.empty(cin_pad, kpad_r, dtype=torch.bfloat16, device="cuda")
    assert L.m2d_unet_pack_weights(C.c_void_p(wfd.data_ptr()), cout, cin, kpad,
                                   C.c_void_p(wb.data_ptr()), C.c_void_p(wrot.data_ptr()), kpad_r,
                                   cin_pad, None) == 0
    gi = _conv_gpu(L, dG, cout, 0, None, 0, wrot, None, cin, kpad_r, n, h, w, 0)
    gi_d = gi.to("cuda")
    g1 = torch.zeros(sh1, device="cuda")
    assert L.m2d_unet_mode_bwd(C.c_void_p(gi_d.data_ptr()), cin, 0, c1, mode, n, h, w,
                               C.c_void_p(g1.data_ptr()), None) == 0
    torch.cuda.synchronize()
    sw = wr_.grad.abs().max().item()
    assert torch.allclose(dw.cpu()[:cout, :9 * cin], wr_.grad, atol=1e-4 * sw, rtol=0)
    assert not dw.cpu()[:, 9 * cin:].any()
    assert torch.allclose(db.cpu(), G.sum(dim=(0, 1, 2)), atol=1e-4 * G.abs().sum().item() / cout)
    s1 = x1r.grad.abs().max().item()
    assert torch.allclose(g1.cpu(), x1r.grad, atol=1e-4 * s1, rtol=0)
    if c2:
        g2 = torch.zeros(n, h, w, c2, device="cuda")
        assert L.m2d_unet_mode_bwd(C.c_void_p(gi_d.data_ptr()), cin, c1, c2, 0, n, h, w,
                                   C.c_void_p(g2.data_ptr()), None) == 0
        torch.cuda.synchronize()
        s2 = x2r.grad.abs().max().item()
        assert torch.allclose(g2.cpu(), x2r.grad, atol=1e-4 * s2, rtol=0)


def _ref_net(net, x):
    """The whole network in FP32 on the CPU from the same weights, rounding
    activations to bf16 between layers as the kernels do."""
    import torch
    acts = {"x": x}
    n, h, w, _ = x.shape
    out = None
    for lay in net.layers:
        src = acts[lay.src]
        if lay.mode == 1:
            src = _pool(src)
        elif lay.mode == 2:
            src = _effective(src, 2, None)
        if lay.skip:
            src = torch.cat([src, acts[lay.skip]], dim=3)
        wv = _bf(lay.w.cpu()[:lay.cout, :9 * lay.cin])
        y = _ref_conv(src, wv, lay.b.cpu(), lay.cin, lay.cout, lay.relu)
        acts[lay.name] = y if lay.name == "head" else _bf(y)
        out = y
    return out


def test_unet_forward_matches_reference(gpu):
    import torch

    from mach2d import unet as U
    net = U.UNet(seed=3)
    g = torch.Generator().manual_seed(5)
    for lay in net.layers:   # non-zero biases exercise the bias path
        lay.b.copy_(torch.randn(lay.cout, generator=g) * 0.1)
    x = _bf(torch.randn(2, 16, 24, 8, generator=g))
    x[..., 6:] = 0
    got = net.forward(_dev(x)).cpu()
    torch.cuda.synchronize()
    want = _ref_net(net, x)
    scale = want[..., 0].abs().max().item()
    err = (got[..., 0] - want[..., 0]).abs().max().item()
    assert err <= 2e-2 * scale, (err, scale)


def test_unet_training_reduces_loss(gpu):
    """Adam on a fixed small batch: the loss falls well below its start."""
    import torch

    from mach2d import unet as U
    net = U.UNet(seed=1)
    g = torch.Generator().manual_seed(2)
    x = _bf(torch.randn(2, 16, 16, 8, generator=g))
    x[..., 6:] = 0
    target = (x[..., 0] * 0.5 + 0.25 * x[..., 1]).contiguous()
    xd, td = _dev(x), target.to("cuda")
    mask = torch.ones_like(td)
    losses = []
    for _ in range(60):
        loss = net.loss_and_backward(xd, td, mask)
        net.adam_step(lr=3e-3)
        losses.append(float(loss.item()))
    assert all(np.isfinite(losses))
    assert losses[-1] < 0.2 * losses[0], losses[::10]


def test_pack_input_and_guess_scale(gpu):
    """Network input channels from a 9-diagonal system, and the guess path
    (output times s) on the grid."""
    import ctypes as C

    import torch

    from mach2d import problems as P
    from mach2d import unet as U
    nx, ny = 20, 11
    s = P.pcorr_system(nx, ny, seed=4, q=1.05, rhs="gauss")
    net = U.UNet(seed=0)
    a = torch.tensor(s.a.reshape(-1), device="cuda")
    b = torch.tensor(s.rhs, device="cuda")
    x = net.pack_input(a, b, nx, ny, 2.0).cpu().float()
    h, w = U.UNet.padded(nx, ny)
    assert x.shape == (1, h, w, 8)
    A = s.a.reshape(9, ny, nx)
    ap = A[4]
    with np.errstate(divide="ignore", invalid="ignore"):
        want0 = np.where(ap != 0, s.rhs.reshape(ny, nx) / ap / 2.0, 0.0)
        want1 = np.where(ap != 0, A[3] / ap, 0.0)
    got = x[0, :ny, :nx].numpy()
    assert np.allclose(got[..., 0], want0, rtol=1e-2, atol=1e-6)
    assert np.allclose(got[..., 1], want1, rtol=1e-2, atol=1e-6)
    assert not x[0, ny:].any() and not x[0, :, nx:].any() and not x[..., 6:].any()
    var = torch.zeros(nx * ny, dtype=torch.float64, device="cuda")
    sc = net.guess(a, b, nx, ny, var)
    out = net.forward(net.pack_input(a, b, nx, ny, sc)).cpu()
    assert np.allclose(var.cpu().numpy().reshape(ny, nx), sc * out[0, :ny, :nx, 0].double().numpy())
    assert C.sizeof(C.c_void_p) == 8


def test_guess_scale_minimises_residual(gpu):
    """m2d_guess_scale: alpha = (Ag.b)/(Ag.Ag); the exact solution scales by
    1, a doubled one by 0.5, and a scaled arbitrary guess is never worse than
    the zero guess (first-order optimality: (A(alpha g) - b).Ag = 0)."""
    import torch

    from mach2d import problems as P
    s = P.pcorr_system(34, 18, seed=2, q=1.05, rhs="gauss")
    dev = gpu.DeviceMSI(s.nx, s.ny, "fp64")
    a = torch.tensor(s.a.reshape(-1), device="cuda")
    b = torch.tensor(s.rhs, device="cuda")
    A = np.zeros((s.N, s.N))
    nx = s.nx
    off = [-nx - 1, -nx, -nx + 1, -1, 0, 1, nx - 1, nx, nx + 1]
    for k, o in enumerate(off):
        for p in range(s.N):
            if 0 <= p + o < s.N and s.a[k, p] != 0.0:
                A[p, p + o] = s.a[k, p]
    x = np.linalg.solve(A, s.rhs)
    for mult, want in ((1.0, 1.0), (2.0, 0.5)):
        g = torch.tensor(mult * x, device="cuda")
        alpha = dev.guess_scale(a, b, g)
        assert abs(alpha - want) < 1e-10, (mult, alpha)
    rng = np.random.default_rng(3)
    g0 = rng.standard_normal(s.N)
    g = torch.tensor(g0, device="cuda")
    alpha = dev.guess_scale(a, b, g)
    Ag = A @ g0
    assert abs(alpha - (Ag @ s.rhs) / (Ag @ Ag)) < 1e-12 * abs(alpha) + 1e-300
    r = A @ (alpha * g0) - s.rhs
    assert abs(r @ Ag) < 1e-9 * np.linalg.norm(Ag) * np.linalg.norm(s.rhs)
    assert dev.residual_norm(a, b, g) <= dev.residual_norm(a, b, torch.zeros_like(g)) * (1 + 1e-12)
    dev.close()


def test_guess_scale_energy(gpu):
    """m2d_guess_scale_energy: alpha = (g.b)/(g.Ag); the exact solution
    scales by 1, a doubled one by 0.5; for an arbitrary guess the scaled
    one satisfies the first-order condition g.(A(alpha g) - b) = 0, and a
    guess with g.Ag <= 0 becomes zero."""
    import torch

    from mach2d import problems as P
    s = P.pcorr_system(34, 18, seed=2, q=1.05, rhs="gauss")
    dev = gpu.DeviceMSI(s.nx, s.ny, "fp64")
    a = torch.tensor(s.a.reshape(-1), device="cuda")
    b = torch.tensor(s.rhs, device="cuda")
    A = np.zeros((s.N, s.N))
    nx = s.nx
    off = [-nx - 1, -nx, -nx + 1, -1, 0, 1, nx - 1, nx, nx + 1]
    for k, o in enumerate(off):
        for p in range(s.N):
            if 0 <= p + o < s.N and s.a[k, p] != 0.0:
                A[p, p + o] = s.a[k, p]
    x = np.linalg.solve(A, s.rhs)
    for mult, want in ((1.0, 1.0), (2.0, 0.5)):
        g = torch.tensor(mult * x, device="cuda")
        alpha = dev.guess_scale(a, b, g, energy=True)
        assert abs(alpha - want) < 1e-10, (mult, alpha)
    g0 = np.random.default_rng(4).standard_normal(s.N)
    alpha = dev.guess_scale(a, b, torch.tensor(g0, device="cuda"), energy=True)
    assert abs(alpha - (g0 @ s.rhs) / (g0 @ (A @ g0))) < 1e-12 * abs(alpha)
    assert abs(g0 @ (A @ (alpha * g0) - s.rhs)) < 1e-9 * np.linalg.norm(g0) * np.linalg.norm(s.rhs)
    g = torch.zeros(s.N, dtype=torch.float64, device="cuda")
    assert dev.guess_scale(a, b, g, energy=True) == 0.0
    dev.close()


def test_guessed_solve_matches_zero_guess_solution(gpu):
    """Guess-then-MSI with an untrained network: the guess is scaled by the
    energy-minimising amplitude, the solve reaches gamma * ||b||, and the
    solution agrees with the zero-guess solution to 1e-8 relative."""
    import torch

    from mach2d import learned as Lr
    from mach2d import problems as P
    from mach2d import unet as U
    s = P.pcorr_system(66, 34, seed=5, q=1.05, rhs="gauss")
    net = U.UNet(seed=11)
    dev = gpu.DeviceMSI(s.nx, s.ny, "fp64")
    a = torch.tensor(s.a.reshape(-1), device="cuda")
    b = torch.tensor(s.rhs, device="cuda")
    dev.factor(a)
    v0 = torch.zeros(s.N, dtype=torch.float64, device="cuda")
    st0 = dev.solve(a, b, v0, 1e-10, 5000)
    v1 = torch.zeros_like(v0)
    st1, used, rg, rb = Lr.guessed_solve(dev, net, a, b, v1, s.nx, s.ny)
    assert st0.status == 0 and st1.status == 0
    assert st1.rn < 1e-10 * rb * (1 + 1e-9)
    rel = float(torch.linalg.norm(v1 - v0) / torch.linalg.norm(v0))
    assert rel < 1e-8, rel
    dev.close()


def test_guess_many_equals_single_guesses(gpu):
    """One network pass over several systems of one grid gives each system
    the guess it gets alone, bit for bit (the images are independent)."""
    import torch

    from mach2d import learned as Lr
    from mach2d import problems as P
    from mach2d import unet as U
    net = U.UNet()
    net.load(Lr.default_weights())
    nx, ny = 130, 66
    sys_ = [P.pcorr_system(nx, ny, seed=40 + q, rhs="gauss") for q in range(3)]
    a = [torch.tensor(s.a.reshape(-1), device="cuda") for s in sys_]
    b = [torch.tensor(s.rhs, device="cuda") for s in sys_]
    one = [torch.zeros(s.N, dtype=torch.float64, device="cuda") for s in sys_]
    many = [torch.zeros(s.N, dtype=torch.float64, device="cuda") for s in sys_]
    s1 = [net.guess(a[q], b[q], nx, ny, one[q]) for q in range(3)]
    s2 = net.guess_many(a, b, nx, ny, many)
    torch.cuda.synchronize()
    assert s1 == s2
    for q in range(3):
        assert torch.equal(one[q], many[q]), q
        assert float(one[q].abs().max()) > 0


def test_guess_prepare_many_matches_single_calls(gpu):
    """The batched scaling gives each system guess_scale's alpha (same sums
    in the same order) and residual_norm's ||rhs - A g|| to 1e-12."""
    import torch

    from mach2d import problems as P
    nx, ny = 97, 64
    sys_ = [P.pcorr_system(nx, ny, seed=60 + q, rhs="gauss") for q in range(3)]
    dev = gpu.DeviceMSI(nx, ny)
    a = [torch.tensor(s.a.reshape(-1), device="cuda") for s in sys_]
    b = [torch.tensor(s.rhs, device="cuda") for s in sys_]
    g0 = [torch.tensor(np.random.default_rng(q).uniform(-1, 1, s.N), device="cuda")
          for q, s in enumerate(sys_)]
    g1 = [x.clone() for x in g0]
    g2 = [x.clone() for x in g0]
    alphas, rn = dev.guess_prepare_many(a, b, g2, energy=True)
    for q in range(3):
        al = dev.guess_scale(a[q], b[q], g1[q], energy=True)
        assert al == alphas[q]
        assert torch.equal(g1[q], g2[q])
        assert rn[q] == pytest.approx(dev.residual_norm(a[q], b[q], g1[q]), rel=1e-12)
    dev.close()


def test_simplec_pairs_match_reference_fixture(gpu, tmp_path):
    """Training pairs generated on the GPU (device SIMPLEC time step, batched
    FP64 MSI targets) equal the pairs of the same time steps composed from
    the compiled reference's routines and solved by its class_solver_msi9d
    (tests/golden/simplec_pairs_34x18.m2dt): systems, right-hand sides and
    targets bit for bit (as stored, float32), the same step and itp
    metadata, the residual drop to 1e-3 relative (a residual 1e-13 of ||b||
    summed in another order and, on the host, from float64 values
    recomputed by the oracle)."""
    import os

    from mach2d import dataset as D
    from mach2d import learned as Lr
    path = str(tmp_path / "p.m2dt")
    n = Lr.simplec_pairs(path, variants=1, steps=12, nx=34, ny=18, seed=0)
    assert n == 12
    got = D.Dataset(path)
    want = D.Dataset(os.path.join(os.path.dirname(__file__), "golden",
                                  "simplec_pairs_34x18.m2dt"))
    assert (got.nx, got.ny, got.ndiag, len(got)) == (want.nx, want.ny, want.ndiag, len(want))
    for q in range(len(want)):
        for u, v in zip(got[q], want[q]):
            assert np.array_equal(u, v), q
        assert list(got.meta[q][:3]) == list(want.meta[q][:3])
        assert got.meta[q][3] == pytest.approx(want.meta[q][3], rel=1e-3)
