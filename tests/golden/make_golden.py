"""Writes the committed golden fixture of the whole-pyramid oracle.

The reference holds no golden vectors for this path, so the fixture records
the oracle's own FP32 output on a seeded 64x64 input (2 levels, so each level samples >= 16 patches;
radius 3, luma threshold 3): it
pins the oracle against silent regressions, while the oracle itself is pinned
by tests/test_oracle.py.  Run from the repository root:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))

import oracle  # noqa: E402
from tests.helpers import default_params, nlf_for, noisy_ycc  # noqa: E402


def main():
    noisy, _ = noisy_ycc(64, 64, seed=7)
    nlf = nlf_for(2)
    res = oracle.pyramid_denoise(noisy, default_params(2, luma=3.0), nlf, levels=2, radius=3)
    out = {"input": noisy, "nlf": nlf}
    for i in range(2):
        out[f"denoised{i}"] = res.denoised[i]
        out[f"projector{i}"] = res.bases[i] @ res.bases[i].T
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pyramid_64x64_l2_r3.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
