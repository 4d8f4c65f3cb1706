"""Inference on the RUM distance: Hong and Li's bootstrap p-value and confidence interval.

    boot = rum.bootstrap(panel=panel)        # respondents resampled
    boot = rum.bootstrap(replies=replies)    # each menu's replies resampled
    boot.p_value(), boot.confidence_interval(), boot.distance_interval()
"""
from . import hong_li
from .bootstrap import Bootstrap, DistanceInterval, Interval, PValue
from .draws import WEIGHTS, Draws, MenuDraws, cells_from_panel, read_panel

__all__ = ["hong_li", "Bootstrap", "DistanceInterval", "Interval", "PValue", "WEIGHTS", "Draws", "MenuDraws",
           "cells_from_panel", "read_panel"]
