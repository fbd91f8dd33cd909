"""U-Net for the learned pressure guess, on the project's HIP kernels.

Every forward and backward operation (3x3 conv with bias/ReLU on the matrix
cores, its input and weight gradients, average pooling, nearest upsampling,
channel concat, the 1x1 head) is a kernel of libnatrix_hip.so called through
the C-ABI; PyTorch supplies only tensor storage, autograd bookkeeping and
the optimizer.  Activations are NHWC bfloat16 ([N, H, W, C] tensors),
conv weights are float32 master copies [Cout, 9, Cin] cast to bfloat16 per
call, accumulation is float32.

Architecture (4 levels, base 32 channels, two 3x3 conv + ReLU per level):
  enc 8->32->32 | pool | 32->64->64 | pool | 64->128->128 | pool | 128->256->256
  dec up, concat skip: 384->128->128 | up, 192->64->64 | up, 96->32->32
  head 1x1 32->1 (float32).  Input channels: 0 = right-hand side, 1 = solid
  mask, 2..7 zero.  H and W must be multiples of 8.
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from .. import _lib

BF16 = torch.bfloat16
IN_CH = 8


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _s():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _call(name, *args):
    _lib.check(getattr(_lib.lib(), name)(*args))


def conv_splits(m, cin, cout):
    """K splits of a forward convolution over m pixels: about one workgroup
    per compute unit (~256) for small images with many channels, at least 8
    K steps of 16 channels each (fewer, longer splits: every split writes
    and the finish pass re-reads a float32 partial of the whole output); 1
    (no split) when the image alone fills the chip.  A count whose splits do
    not fit one workgroup's waves (at most 16 splits of at most 9 K steps,
    one load phase of the in-workgroup split-K convolution, no finish pass)
    is replaced by the smallest one that does, else by 16 splits when those
    take at most 18 steps (two load phases)."""
    if cin % 16:
        return 1
    base = -(-m // 256) * (cout // (64 if cout % 64 == 0 else 32))
    if base >= 256:
        return 1
    steps = 9 * cin // 16
    s = max(1, min(steps // 8, -(-256 // base)))
    if s == 1:
        return 1
    per = -(-steps // s)
    used = -(-steps // per)
    if used > 16 or per > 9:
        s2 = -(-steps // 9)
        if s2 <= 16:
            s = s2
        elif -(-steps // 16) <= 18:
            s = 16
    return s


class Conv3x3Fn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w, b, relu):
        n, h, wd, cin = x.shape
        cout = w.shape[0]
        wb = w.to(BF16).contiguous()
        y = torch.empty((n, h, wd, cout), dtype=BF16, device=x.device)
        splits = conv_splits(n * h * wd, cin, cout)
        if splits > 1:
            work = torch.empty(splits * n * h * wd * cout, dtype=torch.float32,
                               device=x.device)
            _call("nx_conv3x3_fwd_split", _p(x), _p(wb), _p(b), _p(y), _p(work), splits,
                  n, h, wd, cin, cout, 1 if relu else 0, _s())
        else:
            _call("nx_conv3x3_fwd", _p(x), _p(wb), _p(b), _p(y), n, h, wd, cin, cout,
                  1 if relu else 0, _s())
        ctx.save_for_backward(x, wb, y)
        ctx.relu = relu
        return y

    @staticmethod
    def backward(ctx, dy):
        x, wb, y = ctx.saved_tensors
        n, h, wd, cin = x.shape
        cout = wb.shape[0]
        dy = dy.contiguous().to(BF16)
        if ctx.relu:
            dz = torch.empty_like(dy)
            _call("nx_relu_bwd", _p(y), _p(dy), _p(dz), dy.numel(), _s())
        else:
            dz = dy
        dx = None
        if ctx.needs_input_grad[0]:
            dx = torch.empty_like(x)
            wt = torch.empty(9 * cin * cout, dtype=BF16, device=x.device)
            _call("nx_conv3x3_dgrad", _p(dz), _p(wb), _p(dx), _p(wt), n, h, wd, cin,
                  cout, _s())
        m = n * h * wd
        splits = max(1, min(64, m // 4096))
        dw = torch.empty((cout, 9, cin), dtype=torch.float32, device=x.device)
        work = torch.empty(splits * cout * 9 * cin, dtype=torch.float32, device=x.device)
        _call("nx_conv3x3_wgrad", _p(x), _p(dz), _p(dw), _p(work), splits, n, h, wd,
              cin, cout, _s())
        db = torch.empty(cout, dtype=torch.float32, device=x.device)
        cw = torch.empty(1024 * cout, dtype=torch.float32, device=x.device)
        _call("nx_colsum", _p(dz), _p(db), _p(cw), m, cout, _s())
        return dx, dw, db, None


class Pool2Fn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        n, h, w, c = x.shape
        y = torch.empty((n, h // 2, w // 2, c), dtype=BF16, device=x.device)
        _call("nx_avgpool2", _p(x), _p(y), n, h, w, c, 0, _s())
        ctx.shape = (n, h, w, c)
        return y

    @staticmethod
    def backward(ctx, dy):
        n, h, w, c = ctx.shape
        dx = torch.empty((n, h, w, c), dtype=BF16, device=dy.device)
        _call("nx_avgpool2", _p(dy.contiguous().to(BF16)), _p(dx), n, h, w, c, 1, _s())
        return dx


class Up2Fn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        n, h, w, c = x.shape
        y = torch.empty((n, 2 * h, 2 * w, c), dtype=BF16, device=x.device)
        _call("nx_upsample2", _p(x), _p(y), n, 2 * h, 2 * w, c, 0, _s())
        ctx.shape = (n, h, w, c)
        return y

    @staticmethod
    def backward(ctx, dy):
        n, h, w, c = ctx.shape
        dx = torch.empty((n, h, w, c), dtype=BF16, device=dy.device)
        _call("nx_upsample2", _p(dy.contiguous().to(BF16)), _p(dx), n, 2 * h, 2 * w, c,
              1, _s())
        return dx


class ConcatFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, b):
        n, h, w, ca = a.shape
        cb = b.shape[3]
        y = torch.empty((n, h, w, ca + cb), dtype=BF16, device=a.device)
        _call("nx_concat", _p(y), _p(a), _p(b), n * h * w, ca, cb, 0, _s())
        ctx.dims = (n, h, w, ca, cb)
        return y

    @staticmethod
    def backward(ctx, dy):
        n, h, w, ca, cb = ctx.dims
        da = torch.empty((n, h, w, ca), dtype=BF16, device=dy.device)
        db = torch.empty((n, h, w, cb), dtype=BF16, device=dy.device)
        _call("nx_concat", _p(dy.contiguous().to(BF16)), _p(da), _p(db), n * h * w, ca,
              cb, 1, _s())
        return da, db


class HeadFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w, b):
        n, h, wd, c = x.shape
        out = torch.empty((n, h, wd), dtype=torch.float32, device=x.device)
        _call("nx_head_fwd", _p(x), _p(w), _p(b), _p(out), n * h * wd, c, _s())
        ctx.save_for_backward(x, w)
        return out

    @staticmethod
    def backward(ctx, dout):
        x, w = ctx.saved_tensors
        n, h, wd, c = x.shape
        dout = dout.contiguous().float()
        dx = torch.empty_like(x)
        dwb = torch.empty(c + 1, dtype=torch.float32, device=x.device)
        work = torch.empty(1024 * (c + 1), dtype=torch.float32, device=x.device)
        _call("nx_head_bwd", _p(x), _p(w), _p(dout), _p(dx), _p(dwb), _p(work),
              n * h * wd, c, _s())
        return dx, dwb[:c].clone(), dwb[c:].clone()


class Conv(torch.nn.Module):
    def __init__(self, cin, cout, relu=True):
        super().__init__()
        std = math.sqrt(2.0 / (9 * cin))  # He initialisation
        self.w = torch.nn.Parameter(torch.randn(cout, 9, cin) * std)
        self.b = torch.nn.Parameter(torch.zeros(cout))
        self.relu = relu

    def forward(self, x):
        return Conv3x3Fn.apply(x, self.w, self.b, self.relu)


class UNet(torch.nn.Module):
    def __init__(self, base=32):
        super().__init__()
        c = base
        self.e0 = torch.nn.ModuleList([Conv(IN_CH, c), Conv(c, c)])
        self.e1 = torch.nn.ModuleList([Conv(c, 2 * c), Conv(2 * c, 2 * c)])
        self.e2 = torch.nn.ModuleList([Conv(2 * c, 4 * c), Conv(4 * c, 4 * c)])
        self.e3 = torch.nn.ModuleList([Conv(4 * c, 8 * c), Conv(8 * c, 8 * c)])
        self.d2 = torch.nn.ModuleList([Conv(12 * c, 4 * c), Conv(4 * c, 4 * c)])
        self.d1 = torch.nn.ModuleList([Conv(6 * c, 2 * c), Conv(2 * c, 2 * c)])
        self.d0 = torch.nn.ModuleList([Conv(3 * c, c), Conv(c, c)])
        self.hw = torch.nn.Parameter(torch.randn(c) * math.sqrt(1.0 / c))
        self.hb = torch.nn.Parameter(torch.zeros(1))

    @staticmethod
    def _two(convs, x):
        return convs[1](convs[0](x))

    def forward(self, x):
        """x: [N, H, W, 8] bfloat16 -> [N, H, W] float32."""
        e0 = self._two(self.e0, x)
        e1 = self._two(self.e1, Pool2Fn.apply(e0))
        e2 = self._two(self.e2, Pool2Fn.apply(e1))
        e3 = self._two(self.e3, Pool2Fn.apply(e2))
        d2 = self._two(self.d2, ConcatFn.apply(Up2Fn.apply(e3), e2))
        d1 = self._two(self.d1, ConcatFn.apply(Up2Fn.apply(d2), e1))
        d0 = self._two(self.d0, ConcatFn.apply(Up2Fn.apply(d1), e0))
        return HeadFn.apply(d0, self.hw, self.hb)

    def n_params(self):
        return sum(p.numel() for p in self.parameters())


def pack_input(rhs, solid):
    """[N, H, W] float rhs (already normalised) and solid mask -> [N,H,W,8]
    bfloat16 network input."""
    n, h, w = rhs.shape
    x = torch.zeros((n, h, w, IN_CH), dtype=BF16, device=rhs.device)
    x[..., 0] = rhs.to(BF16)
    x[..., 1] = solid.to(BF16)
    return x
