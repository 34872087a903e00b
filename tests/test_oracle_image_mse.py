"""Pins of the oracle's image MSE (texturemse.cpp:39-66, sqerror.frag) by
closed forms: equal images, a constant per-channel offset, and references
resampled so that nearest sampling at pixel centres must pick known texels."""
import numpy as np


def test_equal_images_give_zero(oracle):
    rng = np.random.default_rng(1)
    acc = rng.random((24, 40, 4), dtype=np.float32) * 8
    ref = acc[..., :3] / np.float32(8)
    assert oracle.image_mse(acc, 8, ref) == 0.0


def test_constant_offset_is_the_sum_of_squares(oracle):
    # diff = c in every pixel, so the error is dot(c, c) = the sum over channels
    # (not the mean over channels)
    acc = np.zeros((16, 16, 4), np.float32)
    acc[..., :3] = 3.0
    ref = np.full((16, 16, 3), 1.0, np.float32) - np.array([0.5, 0.25, 0.125], np.float32)
    got = oracle.image_mse(acc, 3, ref)
    assert abs(got - (0.25 + 0.0625 + 0.015625)) < 1e-7


def test_nearest_sampling_of_larger_and_smaller_references(oracle):
    rng = np.random.default_rng(2)
    acc = rng.random((20, 30, 4), dtype=np.float32)
    img = acc[..., :3]
    up = np.repeat(np.repeat(img, 2, axis=0), 2, axis=1)  # every render texel twice in x and y
    assert oracle.image_mse(acc, 1, up) == 0.0
    down = img[1::2, 1::2]  # centre of reference pixel x lies in render texel 2x + 1
    assert oracle.image_mse(acc, 1, down) == 0.0
    assert oracle.image_mse(acc, 1, img[0::2, 0::2]) > 1e-3  # the even texels are the wrong ones


def test_divisor_scales_the_render_only(oracle):
    acc = np.zeros((4, 4, 4), np.float32)
    acc[..., :3] = 10.0
    ref = np.zeros((4, 4, 3), np.float32)
    assert abs(oracle.image_mse(acc, 5, ref) - 12.0) < 1e-6  # (10/5)^2 * 3
