"""The reference-shaped C++ surface on the device (GPU).

tests/cpp/facade_denoise (built by __graft_entry__.build()) runs
gcam::PyramidProcessor<5>::denoise with calibrateFromImage set, reads
denoisedImagePyramid(3) and calls MeasureYCbCrNLF on the input, the sequence
RawConverter::denoise runs (MetalPipeline/raw_converter.cpp:80-90) against
PyramidProcessor<levels> (MetalPipeline/pyramid_processor.hpp:57-75).  Its
outputs are compared with the oracle: levels 0 and 3 within max-abs 1e-4, the
fitted noise model of every level and the measured one within 1e-4 (B
relative, A against the model's variance at mid-grey).
"""
import os
import subprocess

import numpy as np
import pytest

import glasscamerapublic_amd as gcam
import oracle
from tests.helpers import preset_params

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "facade_denoise")
FIELDS = ("luma", "chroma", "chromaBoost", "gradientBoost", "gradientThreshold", "sharpening")


def check_nlf(got, exp, what):
    """(A, B) of var = A + B Y: B within 1e-4 relative; A, the intercept, within 1e-4 of the model's variance
    at mid-grey (A + 0.5 B): A is a small difference of large regression sums, so its own relative error
    is amplified by the summation order of the FP32 statistics."""
    np.testing.assert_allclose(got[1], exp[1], rtol=1e-4, err_msg=f"{what} B")
    np.testing.assert_allclose(got[0], exp[0], rtol=0, atol=1e-4 * np.abs(exp[0] + 0.5 * exp[1]).max(),
                               err_msg=f"{what} A")


def test_cpp_facade_denoise_on_device(tmp_path, oracle_threads):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    w, h, radius = 1024, 768, 4
    img = gcam.synth_frame(w, h, 4e-3, 2e-3, scene_seed=71, noise_seed=72).cpu().numpy()
    prm = preset_params(5)
    img.astype(np.float32).tofile(tmp_path / "input.f32")
    np.array([[p[k] for k in FIELDS] for p in prm], np.float32).tofile(tmp_path / "params.f32")
    print(f"running {EXE} {w}x{h} r{radius}", flush=True)
    r = subprocess.run([EXE, str(w), str(h), str(radius), str(tmp_path / "input.f32"), str(tmp_path / "params.f32"),
                        str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    level0 = np.fromfile(tmp_path / "level0.f32", np.float32).reshape(h, w, 4)
    level3 = np.fromfile(tmp_path / "level3.f32", np.float32).reshape(h // 8, w // 8, 4)
    fitted = np.fromfile(tmp_path / "nlf.f32", np.float32).reshape(5, 2, 3)
    measured = np.fromfile(tmp_path / "measured.f32", np.float32).reshape(2, 3)

    ref = oracle.pyramid_denoise(img, prm, np.zeros((5, 2, 3)), levels=5, radius=radius, calibrate=True)
    for i in range(5):
        check_nlf(fitted[i], ref.nlf[i], f"fitted noise model, level {i}")
    for got, exp, name in ((level0, ref.denoised[0], "level 0"), (level3, ref.denoised[3], "level 3")):
        err = float(np.abs(got[..., :3] - exp[..., :3]).max())
        print(f"{name}: max-abs vs oracle {err:.3g}", flush=True)
        assert err <= 1e-4, name
        np.testing.assert_allclose(got[..., 3], exp[..., 3], rtol=1e-4, err_msg=f"{name} .w")
    fit = oracle.fit_ycbcr_nlf(oracle.noise_stats(img))
    check_nlf(measured, np.stack([fit.A, fit.B]), "MeasureYCbCrNLF")
