"""Learned warm start for the batched bootstrap (mlp.py): an MLP trained on the GPU on draws the
method has already certified."""
from .mlp import MLP, WarmStart

__all__ = ["MLP", "WarmStart"]
