"""Temperature softmax / softmin layers around the ops (reference corenet.py:97-113)."""
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
