"""MI355X-native BATS: exact event-based LIF simulation and BATS backprop on gfx950.

Drop-in for the reference package ``bats`` (Florian-BACHO/bats): the same Network /
layers / losses / optimizers API over torch device tensors, backed by hand-written HIP
kernels in libbats_amd.so (see include/bats_amd.h).
"""
from .AbstractLayer import AbstractLayer
from .AbstractOptimizer import AbstractOptimizer
from .AbstractLoss import AbstractLoss
from .AbstractConvLayer import AbstractConvLayer
from .Network import Network
from .Losses import *

__version__ = "0.1.0"
