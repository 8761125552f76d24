"""Write the reference C++ solver's output on its own parity scene as a
fixture, so the GPU parity test needs no compiled reference code.

Run in the build container only, after ``make -C oracle`` has built
oracle/_ref/libref_growth.so from the reference sources:

    python tests/golden/make_cpp_fixture.py

The scene is tests/test_growth.cpp:337-357 (27-particle Hencky free swell,
Fg = 1.05 I, N = 12, dx = 0.05, 30 steps of 1e-3), run by the reference's
FP32 forward solver.  Output: cpp_parity_scene.npz (data only: the inputs and
the solver's final positions and deformation gradients).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)


def main():
    from test_ref_cpp import cpp_run, parity_scene
    pts, Fg, _ = parity_scene()
    x, F = cpp_run(pts, Fg, 12, 0.05, 30, 1e-3)
    np.savez(os.path.join(HERE, "cpp_parity_scene.npz"), pts=pts, Fg=Fg, steps=30, x=x, F=F)
    print("wrote cpp_parity_scene.npz")


if __name__ == "__main__":
    main()
