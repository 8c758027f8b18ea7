"""The learned prolongation of the two-level preconditioner: the trained graph-network parameters and their
layout (13 input features, hidden width 32, three message-passing layers, 4 parent-slot outputs)."""
from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "prolongation_gnn.json")
F_IN, HIDDEN, LAYERS = 13, 32, 3


def param_count(hidden: int = HIDDEN) -> int:
    return hidden * F_IN + hidden + LAYERS * (2 * hidden * hidden + hidden) + 4 * hidden + 4


def load_model(path: str = MODEL_PATH) -> np.ndarray:
    """The trained parameters as a flat float32 array (W0, b0, then Ws, Wn, b per layer, then Wo, bo)."""
    with open(path) as f:
        meta = json.load(f)
    if meta["hidden"] != HIDDEN or meta["layers"] != LAYERS or meta["features"] != F_IN:
        raise ValueError("prolongation model does not match the library's network shape")
    theta = np.asarray(meta["params"], np.float32)
    if theta.size != param_count():
        raise ValueError("prolongation model has the wrong parameter count")
    return theta
