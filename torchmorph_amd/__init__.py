"""torchmorph_amd -- MI355X-native (gfx950 HIP) implementation of torchmorph's
batched operator path: grey/binary morphology, exact Euclidean and chamfer
distance transforms, and log-domain Sinkhorn optimal transport.

The public names mirror the reference package ``torchmorph`` (reference
torchmorph/__init__.py:1-50); ``import torchmorph_amd as torchmorph`` is the
intended drop-in.
"""
from .distance_transform import (
    brute_force_distance_transform,
    chamfer_distance_transform,
    distance_transform_cdt,
    distance_transform_edt,
    euclidean_distance_transform,
)

__all__ = [
    "euclidean_distance_transform",
    "chamfer_distance_transform",
    "brute_force_distance_transform",
    "distance_transform_edt",
    "distance_transform_cdt",
]

__version__ = "0.1.0"
