"""Writes tests/golden/dynamic_trajectories.npz: the oracle's f64 backward-Euler trajectories of the two
dynamic workloads (its simulate, drape.rs:426-439 restated), for the GPU dynamic-loop tests.

  C1: 0.5 m sheet, n33, isotropic E = 1e6 Pa, nu = 0.3, t = 0.5 mm, 0.2 kg/m^2, hinge 1e-5 N m, top corners
      pinned, h = 1/60, 60 steps from rest.
  C2: n65, the same material, no pins, released 0.05 m above the top of a sphere of radius 0.15 m (dhat = 1 mm),
      h = 1/60, the first 30 of its 120 steps (falling, first contact, draping).
Newton to a relative residual of 1e-10, regularised, as the oracle's dynamic runs. Takes a few minutes.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.dirname(HERE)]

import drape_amd as D  # noqa: E402  (scene description only: the builder's parameters)
import oracle_lib as O  # noqa: E402


def oracle_of(scene):
    s = O.Scene()
    for name, _ in O.Scene._fields_:
        setattr(s, name, getattr(scene, name))
    return O.Drape(s)


def main():
    out = {}
    for name, scene, steps in (
            ("c1", D.make_scene(33, material=D.isotropic_material(), bend=(1e-5, 1e-5), density=0.2,
                                time_step=1.0 / 60.0, pinned=True), 60),
            ("c2", D.drape_config_scene(65, clearance=0.05), 30)):
        od = oracle_of(scene)
        iterations, _, x = od.simulate(steps, 1e-10, 200, True, 2.0)
        out[f"{name}_positions"] = x
        out[f"{name}_iterations"] = np.asarray(iterations, np.int32)
        out[f"{name}_rest"] = od.rest_positions()
        print(name, steps, "steps, Newton", int(np.sum(iterations)), flush=True)
    np.savez_compressed(os.path.join(HERE, "dynamic_trajectories.npz"), **out)


if __name__ == "__main__":
    main()
