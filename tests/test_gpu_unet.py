"""U-Net kernels against the CPU float64 oracle (torch.nn.functional on the
CPU in float64, on the same bfloat16-rounded inputs)."""
import numpy as np
import pytest
import torch
import torch.nn.functional as Fn

from natrix_amd.learned import unet as U

pytestmark = pytest.mark.gpu

BF = torch.bfloat16


def rel(a, b):
    a = a.double().cpu()
    b = b.double().cpu()
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-30))


def ref_conv(x, w, b, relu):
    """x [N,H,W,C] -> [N,H,W,Co] float64 CPU, w [Co,9,Ci]."""
    co, _, ci = w.shape
    k = w.double().view(co, 3, 3, ci).permute(0, 3, 1, 2)
    y = Fn.conv2d(x.double().permute(0, 3, 1, 2), k, b.double(), padding=1)
    if relu:
        y = torch.relu(y)
    return y.permute(0, 2, 3, 1)


@pytest.mark.parametrize("n,h,w,ci,co", [(2, 16, 24, 8, 32), (1, 32, 32, 32, 64),
                                         (3, 8, 8, 96, 32), (1, 16, 16, 384, 128)])
def test_conv3x3_forward(gpu, n, h, w, ci, co):
    g = torch.Generator().manual_seed(ci + co)
    x = torch.randn(n, h, w, ci, generator=g).to(BF)
    wt = (torch.randn(co, 9, ci, generator=g) * 0.2).to(BF).float()
    b = torch.randn(co, generator=g)
    for relu in (False, True):
        y = U.Conv3x3Fn.apply(x.cuda(), wt.cuda(), b.cuda(), relu)
        r = ref_conv(x.float(), wt, b, relu)
        assert rel(y.float(), r) < 8e-3, (relu, rel(y.float(), r))


@pytest.mark.parametrize("n,h,w,ci,co,splits", [(1, 16, 16, 256, 256, 36),
                                                (1, 32, 32, 384, 128, 54),
                                                (1, 8, 8, 64, 64, 5),
                                                (2, 12, 20, 32, 96, 18),
                                                (1, 16, 16, 256, 256, 16),
                                                (1, 32, 32, 128, 128, 9),
                                                (1, 64, 64, 192, 64, 13),
                                                (1, 10, 14, 64, 64, 4)])
def test_conv3x3_forward_split_k(gpu, n, h, w, ci, co, splits):
    """K split with float32 partials summed in split order: over launches
    and a finish pass, or, when every split fits one wave (at most 16
    splits of at most 9 K steps: 5, 16, 9, 13 here, 140 pixels in the last
    case, not a multiple of the 32-pixel tile), over the waves of one
    workgroup summed in LDS.  The same float64 oracle as the single launch,
    and within bf16 rounding of it; repeated calls are bitwise equal (the
    sum's order is fixed)."""
    import ctypes as C
    from natrix_amd import _lib
    g = torch.Generator().manual_seed(ci * co + splits)
    x = torch.randn(n, h, w, ci, generator=g).to(BF)
    wt = (torch.randn(co, 9, ci, generator=g) * 0.1).to(BF)
    b = torch.randn(co, generator=g)
    xd, wd, bd = x.cuda(), wt.cuda(), b.cuda()
    y1 = torch.empty(n, h, w, co, dtype=BF, device="cuda")
    y2 = torch.empty_like(y1)
    work = torch.empty(splits * n * h * w * co, device="cuda")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: C.c_void_p(t.data_ptr())
    _lib.call("nx_conv3x3_fwd", p(xd), p(wd), p(bd), p(y1), n, h, w, ci, co, 1, s)
    outs = []
    for _ in range(5):
        _lib.call("nx_conv3x3_fwd_split", p(xd), p(wd), p(bd), p(y2), p(work), splits,
                  n, h, w, ci, co, 1, s)
        outs.append(y2.clone())
    for o in outs[1:]:
        assert torch.equal(o, outs[0])
    r = ref_conv(x.float(), wt.float(), b, True)
    assert rel(y1.float(), r) < 8e-3
    assert rel(y2.float(), r) < 8e-3
    assert rel(y2.float(), y1.float()) < 8e-3
    with pytest.raises(ValueError):
        _lib.call("nx_conv3x3_fwd_split", p(xd), p(wd), p(bd), p(y2), p(work),
                  9 * ci // 16 + 1, n, h, w, ci, co, 1, s)
    with pytest.raises(ValueError):
        _lib.call("nx_conv3x3_fwd_split", p(xd), p(wd), p(bd), p(y2), None,
                  splits, n, h, w, ci, co, 1, s)


@pytest.mark.parametrize("n,h,w,ci,co", [(2, 16, 16, 32, 32), (1, 16, 24, 64, 128),
                                         (2, 8, 8, 96, 32)])
def test_conv3x3_backward(gpu, n, h, w, ci, co):
    g = torch.Generator().manual_seed(7 * ci + co)
    x = torch.randn(n, h, w, ci, generator=g).to(BF)
    wt = (torch.randn(co, 9, ci, generator=g) * 0.2).to(BF).float()
    b = torch.randn(co, generator=g)
    dy = torch.randn(n, h, w, co, generator=g).to(BF)
    xg = x.cuda().requires_grad_(True)
    wg = wt.cuda().requires_grad_(True)
    bg = b.cuda().requires_grad_(True)
    y = U.Conv3x3Fn.apply(xg, wg, bg, True)
    y.backward(dy.cuda())
    # oracle: float64 autograd on the same bf16 values; the ReLU mask is the
    # kernel's (rounded) output > 0
    xr = x.double().requires_grad_(True)
    wr = wt.double().requires_grad_(True)
    br = b.double().requires_grad_(True)
    k = wr.view(co, 3, 3, ci).permute(0, 3, 1, 2)
    z = Fn.conv2d(xr.permute(0, 3, 1, 2), k, br, padding=1).permute(0, 2, 3, 1)
    mask = (y.detach().float().cpu() > 0).double()
    (z * mask).backward(dy.double())
    assert rel(xg.grad.float(), xr.grad) < 1e-2
    assert rel(wg.grad, wr.grad) < 2e-3
    assert rel(bg.grad, br.grad) < 2e-3


def test_pool_upsample_concat_head(gpu):
    g = torch.Generator().manual_seed(3)
    x = torch.randn(2, 8, 12, 32, generator=g).to(BF)
    xg = x.cuda().requires_grad_(True)
    p = U.Pool2Fn.apply(xg)
    u = U.Up2Fn.apply(p)
    c = U.ConcatFn.apply(u, xg)
    wh = torch.randn(64, generator=g)
    bh = torch.randn(1, generator=g)
    whg, bhg = wh.cuda().requires_grad_(True), bh.cuda().requires_grad_(True)
    out = U.HeadFn.apply(c, whg, bhg)
    dout = torch.randn(2, 8, 12, generator=g)
    out.backward(dout.cuda())
    xr = x.double().requires_grad_(True)
    pr = Fn.avg_pool2d(xr.permute(0, 3, 1, 2), 2)
    ur = Fn.interpolate(pr, scale_factor=2, mode="nearest")
    cr = torch.cat([ur, xr.permute(0, 3, 1, 2)], 1).permute(0, 2, 3, 1)
    whr, bhr = wh.double().requires_grad_(True), bh.double().requires_grad_(True)
    orf = (cr * whr).sum(-1) + bhr
    orf.backward(dout.double())
    assert rel(out, orf) < 1e-2
    assert rel(xg.grad.float(), xr.grad) < 2e-2
    assert rel(whg.grad, whr.grad) < 1e-2 and rel(bhg.grad, bhr.grad) < 1e-4


def test_unet_gradients_match_float64_network(gpu):
    torch.manual_seed(0)
    net = U.UNet().cuda()
    g = torch.Generator().manual_seed(1)
    rhs = torch.randn(2, 32, 32, generator=g)
    solid = torch.rand(2, 32, 32, generator=g) < 0.1
    x = U.pack_input(rhs.cuda(), solid.cuda())
    out = net(x)
    tgt = torch.randn(2, 32, 32, generator=g)
    loss = ((out - tgt.cuda()) ** 2).mean()
    loss.backward()
    # float64 CPU twin with the same (bf16-rounded) weights
    def rconv(t, conv, relu=True):
        w = conv.w.detach().to(BF).double().cpu()
        co, _, ci = w.shape
        y = Fn.conv2d(t, w.view(co, 3, 3, ci).permute(0, 3, 1, 2),
                      conv.b.detach().double().cpu(), padding=1)
        return torch.relu(y) if relu else y
    params = {}
    for name, conv in net.named_modules():
        if isinstance(conv, U.Conv):
            params[name] = conv
    xr = x.double().cpu().permute(0, 3, 1, 2)
    def two(t, ms):
        return rconv(rconv(t, ms[0]), ms[1])
    e0 = two(xr, net.e0)
    e1 = two(Fn.avg_pool2d(e0, 2), net.e1)
    e2 = two(Fn.avg_pool2d(e1, 2), net.e2)
    e3 = two(Fn.avg_pool2d(e2, 2), net.e3)
    up = lambda t: Fn.interpolate(t, scale_factor=2, mode="nearest")
    d2 = two(torch.cat([up(e3), e2], 1), net.d2)
    d1 = two(torch.cat([up(d2), e1], 1), net.d1)
    d0 = two(torch.cat([up(d1), e0], 1), net.d0)
    o = (d0 * net.hw.detach().double().cpu()[None, :, None, None]).sum(1) + \
        net.hb.detach().double().cpu()
    assert rel(out.detach(), o) < 5e-2
    # gradient direction of a deep-layer weight against float64 autograd
    w_ref = net.e3[1].w.detach().double().cpu().requires_grad_(True)
    co, _, ci = w_ref.shape
    def rconv_w(t, conv, w):
        y = Fn.conv2d(t, w.view(co, 3, 3, ci).permute(0, 3, 1, 2),
                      conv.b.detach().double().cpu(), padding=1)
        return torch.relu(y)
    e3r = rconv_w(rconv(Fn.avg_pool2d(e2, 2), net.e3[0]), net.e3[1], w_ref)
    d2r = two(torch.cat([up(e3r), e2], 1), net.d2)
    d1r = two(torch.cat([up(d2r), e1], 1), net.d1)
    d0r = two(torch.cat([up(d1r), e0], 1), net.d0)
    orr = (d0r * net.hw.detach().double().cpu()[None, :, None, None]).sum(1) + \
        net.hb.detach().double().cpu()
    ((orr - tgt.double()) ** 2).mean().backward()
    gg = net.e3[1].w.grad.double().cpu().flatten()
    gr = w_ref.grad.flatten()
    cos = float((gg @ gr) / (gg.norm() * gr.norm()))
    assert cos > 0.95, cos
