"""The learned prolongation of the two-level preconditioner: the trained graph-network parameters and their
layout (13 input features, hidden width 32, three message-passing layers, 4 parent-slot outputs, and with a
coarse space of rotation unknowns 4 rotation-scale outputs), and the coarse-space rules it was trained for."""
from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PATH = os.path.join(HERE, "prolongation_gnn.json")
F_IN, HIDDEN, LAYERS = 13, 32, 3


# the coarse space of a model trained with rotation unknowns: three rotations per coarse node, and the
# factorisation's pivot floor that keeps its near-dependent directions positive definite
ROTATION_RULES = {"coarse_rotations": 1, "coarse_pivot_floor": 1e-10}


def param_count(hidden: int = HIDDEN, rotation_head: bool = False) -> int:
    base = hidden * F_IN + hidden + LAYERS * (2 * hidden * hidden + hidden) + 4 * hidden + 4
    return base + (4 * hidden + 4 if rotation_head else 0)


def _meta(path: str) -> dict:
    with open(path) as f:
        meta = json.load(f)
    if meta["hidden"] != HIDDEN or meta["layers"] != LAYERS or meta["features"] != F_IN:
        raise ValueError("prolongation model does not match the library's network shape")
    return meta


def load_model(path: str = MODEL_PATH) -> np.ndarray:
    """The trained parameters as a flat float32 array (W0, b0, then Ws, Wn, b per layer, then Wo, bo, and Wr, br
    for a model with the rotation head)."""
    meta = _meta(path)
    theta = np.asarray(meta["params"], np.float32)
    if theta.size != param_count(rotation_head=bool(meta.get("rotations", False))):
        raise ValueError("prolongation model has the wrong parameter count")
    return theta


def model_rules(path: str = MODEL_PATH) -> dict:
    """The rules of the coarse space the model was trained for (to pass to default_rules with it)."""
    return dict(ROTATION_RULES) if _meta(path).get("rotations", False) else {}
