"""Temperature softmax / softmin layers around the ops (reference corenet.py:97-113) and the
feature U-Net of the semantic front-end (corenet.py:115-431)."""
import torch
import torch.nn as nn
import torch.nn.functional as F


class TemperatureSoftmax(nn.Module):
    """softmax(x / T) over ``dim`` with a learnable temperature T (corenet.py:97-104)."""

    def __init__(self, dim, init_temp=1.0, device=None):
        nn.Module.__init__(self)
        self.T = nn.Parameter(torch.ones(1, dtype=torch.float32, device=device) * init_temp, requires_grad=True)
        self.dim = dim

    def forward(self, x):
        return F.softmax(x / self.T, dim=self.dim)


class TemperatureSoftmin(nn.Module):
    """softmin(x / T) over ``dim`` with a learnable temperature T (corenet.py:106-113)."""

    def __init__(self, dim, init_temp=1.0, device=None):
        nn.Module.__init__(self)
        self.T = nn.Parameter(torch.ones(1, dtype=torch.float32, device=device) * init_temp, requires_grad=True)
        self.dim = dim

    def forward(self, x):
        return F.softmin(x / self.T, dim=self.dim)


# ---------------------------------------------------------------------------------------------
# Feature U-Net of the semantic front-end (reference corenet.py:115-431).  Module and parameter
# names follow the reference's state_dict keys (conv1.conv.conv.weight, conv1.bn.weight, ...,
# conv12.conv.weight), so the shipped checkpoints load strictly.  These run on PyTorch-ROCm
# (MIOpen convolutions); they are not part of the LBP hot path.
# ---------------------------------------------------------------------------------------------

def _activation(name, leaky_alpha=1e-2):
    name = name.lower()
    if name == 'relu':
        return nn.ReLU(inplace=True)
    if name == 'leakyrelu':
        return nn.LeakyReLU(negative_slope=leaky_alpha, inplace=True)
    if name == 'elu':
        return nn.ELU(inplace=True)
    raise NotImplementedError("Activation " + name + " is currently not implemented!")


def pad_to_multiple(x, divisor, extra_pad=(0, 0)):
    """Reflect-pad an [N,C,H,W] tensor so that H and W become multiples of ``divisor``, plus
    ``extra_pad`` = (rows, cols) on every side; of an odd number of added rows (columns) the
    extra one goes to the top (left) (corenet.py:127-147).  Returns (padded, (l, r, t, b))."""
    H, W = x.shape[-2:]
    dh, dw = (-H) % divisor, (-W) % divisor
    t, b = extra_pad[0] + (dh + 1) // 2, extra_pad[0] + dh // 2
    l, r = extra_pad[1] + (dw + 1) // 2, extra_pad[1] + dw // 2
    return F.pad(x, (l, r, t, b), mode='reflect'), (l, r, t, b)


def crop_padding(x, lrtb):
    """Inverse of pad_to_multiple on the last two dimensions (corenet.py:156-187)."""
    l, r, t, b = lrtb
    H, W = x.shape[-2:]
    return x[..., t:H - b, l:W - r].contiguous()


class PaddedConv2d(nn.Module):
    """Convolution whose input is reflect-padded by half the (dilated) kernel, so a stride-1
    convolution keeps the size (corenet.py:215-228); the nn.Conv2d is ``.conv``."""

    def __init__(self, in_channels, out_channels, kernel_size, stride=1, dilation=1, bias=False):
        super().__init__()
        self.conv = nn.Conv2d(in_channels, out_channels, kernel_size, stride=stride, padding=0,
                              dilation=dilation, bias=bias)
        self.dilation = dilation

    def forward(self, x):
        p = (self.conv.kernel_size[0] // 2) * self.dilation
        return self.conv(F.pad(x, (p, p, p, p), mode='reflect'))


class ConvBatchNormAct(nn.Module):
    """conv -> per-channel GroupNorm (one group per channel) -> activation (corenet.py:265-303).
    Without normalisation the convolution carries a bias; ``transposed`` uses an unpadded
    ConvTranspose2d."""

    def __init__(self, in_channels, out_channels, kernel_size, stride=1, dilation=1, padding=0,
                 activation='ReLU', transposed=False, with_bn=True):
        super().__init__()
        if transposed:
            self.conv = nn.ConvTranspose2d(in_channels, out_channels, kernel_size, stride=stride, padding=padding,
                                           dilation=dilation)
        else:
            self.conv = PaddedConv2d(in_channels, out_channels, kernel_size, stride=stride, dilation=dilation,
                                     bias=not with_bn)
        self.bn = nn.GroupNorm(out_channels, out_channels, affine=True) if with_bn else None
        self.act = _activation(activation)

    def forward(self, x):
        x = self.conv(x)
        if self.bn is not None:
            x = self.bn(x)
        return self.act(x)


class StereoUnaryUnetDyn(nn.Module):
    """Three-level U-Net, 3 -> 32 channels at full resolution (corenet.py:329-431): two
    convolutions per level (16, 32, 64 channels) with 2x max-pooling between levels; the decoder
    upsamples (bilinear, or a stride-2 transposed convolution with ``with_upconv``), concatenates
    [skip, upsampled] and applies two convolutions per level; a last 3x3 convolution (GroupNorm
    with ``with_output_bn``).  Inputs must be multiples of ``divisor`` = 8 in H and W.  Returns a
    list: [full-resolution features] or, with ``multi_level_output``, also 32-channel maps of
    levels 1 and 2."""

    def __init__(self, multi_level_output=False, activation='relu', with_bn=True, with_upconv=False,
                 with_output_bn=True):
        super().__init__()
        self.multi_level_output = multi_level_output
        self.with_upconv = with_upconv
        self._divisor = 8
        self._num_output_levels = 3 if multi_level_output else 1
        self._out_channels = [32] * self._num_output_levels
        cba = lambda i, o: ConvBatchNormAct(i, o, 3, padding=1, activation=activation, with_bn=with_bn)
        for idx, (i, o) in enumerate([(3, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64)], start=1):
            setattr(self, f"conv{idx}", cba(i, o))
        self.pool1 = nn.MaxPool2d(2)
        self.pool2 = nn.MaxPool2d(2)
        if with_upconv:
            self.upconv7 = ConvBatchNormAct(64, 32, 3, stride=2, activation=activation, transposed=True,
                                            with_bn=with_bn)
            self.upconv10 = ConvBatchNormAct(32, 16, 3, stride=2, activation=activation, transposed=True,
                                             with_bn=with_bn)
        self.conv8 = cba(64 if with_upconv else 96, 32)
        self.conv9 = cba(32, 32)
        self.conv11 = cba(32 if with_upconv else 48, 32)
        self.conv12 = PaddedConv2d(32, 32, 3)
        self.bn12 = nn.GroupNorm(32, 32, affine=True) if (with_bn and with_output_bn) else None
        if multi_level_output:
            self.conv_lvl1 = PaddedConv2d(32, 32, 3)
            self.conv_lvl2 = PaddedConv2d(64, 32, 3)

    @property
    def divisor(self):
        return self._divisor

    @property
    def out_channels(self):
        return self._out_channels

    @property
    def num_output_levels(self):
        return self._num_output_levels

    def _up(self, x, skip, upconv, trim):
        if self.with_upconv:
            x = upconv(x)
            x = x[:, :, 1:, 1:] if trim == 'lead' else x[:, :, :-1, :-1]
        else:
            x = F.interpolate(x, skip.shape[2:], mode='bilinear', align_corners=False)
        return torch.cat([skip, x], dim=1)

    def forward(self, x):
        lvl0 = self.conv2(self.conv1(x))
        lvl1 = self.conv4(self.conv3(self.pool1(lvl0)))
        x = self.conv6(self.conv5(self.pool2(lvl1)))
        lvl2_out = self.conv_lvl2(x) if self.multi_level_output else None
        x = self.conv9(self.conv8(self._up(x, lvl1, getattr(self, 'upconv7', None), 'lead')))
        lvl1_out = self.conv_lvl1(x) if self.multi_level_output else None
        x = self.conv12(self.conv11(self._up(x, lvl0, getattr(self, 'upconv10', None), 'tail')))
        if self.bn12 is not None:
            x = self.bn12(x)
        return [x, lvl1_out, lvl2_out] if self.multi_level_output else [x]
