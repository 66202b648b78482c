"""Writes the config-1 golden fixture with the CPU oracle.

Config 1: N=1024 Plummer sphere (generator seed 1, scale radius 1, total mass
1, G=1), softening 0.01; t=0 accelerations, then positions after 100
velocity-Verlet steps of dt=1e-3.  The reference holds no expected outputs
for this path, so these vectors come from the oracle (itself pinned by
tests/test_oracle.py) and serve as a fixed regression anchor that travels
with the repository.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "n-body-sim_amd")]

import nbody_amd as nb  # noqa: E402  (host-side generator only)
import oracle  # noqa: E402


def main():
    b = nb.plummer(1024, 1)
    ax, ay, az = oracle.newton(b.masses, b.pos_x, b.pos_y, b.pos_z, 1.0, 0.01)
    x, y, z, vx, vy, vz = oracle.integrate("velocity-verlet", *b.arrays(), 1.0,
                                           0.01, 1e-3, 100)
    path = os.path.join(HERE, "config1_plummer1024.npz")
    np.savez(path, m=b.masses, x=b.pos_x, y=b.pos_y, z=b.pos_z, vx=b.vel_x,
             vy=b.vel_y, vz=b.vel_z, ax0=ax, ay0=ay, az0=az, x100=x, y100=y,
             z100=z, vx100=vx, vy100=vy, vz100=vz)
    digest = hashlib.sha256(open(path, "rb").read()).hexdigest()
    with open(os.path.join(HERE, "config1_plummer1024.json"), "w") as f:
        json.dump({"file": "config1_plummer1024.npz", "sha256": digest,
                   "n": 1024, "seed": 1, "softening": 0.01, "g": 1.0,
                   "dt": 1e-3, "steps": 100, "integrator": "velocity-verlet"},
                  f, indent=1)
    print(path, digest)


if __name__ == "__main__":
    main()
