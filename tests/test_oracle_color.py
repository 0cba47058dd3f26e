"""The oracle's sRGB <-> L*a*b* conversion (io_utility.cpp:268-318, :380-450), pinned on
the CPU:
* bit-exact against the reference's own IOUtility::rgb_to_lab / lab_to_rgb, compiled in
  place from io_utility.cpp (oracle/_ref/libnli_ref_color.so), where that build exists;
* against published CIELAB values of the sRGB primaries, secondaries and greys (D65,
  2 degree observer), which fail on a dropped matrix term, a swapped channel or a wrong
  white point;
* by the round trip rgb -> lab -> rgb, monotonic lightness on the grey ramp, and the
  reference's pass-through of images it does not convert."""
import os

import numpy as np
import pytest

from conftest import bitexact

# published CIELAB (D65) values of 8-bit sRGB colours
PUBLISHED = {
    (255, 255, 255): (100.0, 0.0, 0.0),
    (0, 0, 0): (0.0, 0.0, 0.0),
    (255, 0, 0): (53.2408, 80.0925, 67.2032),
    (0, 255, 0): (87.7347, -86.1827, 83.1793),
    (0, 0, 255): (32.2970, 79.1875, -107.8602),
    (255, 255, 0): (97.1393, -21.5537, 94.4780),
    (0, 255, 255): (91.1132, -48.0875, -14.1312),
    (255, 0, 255): (60.3242, 98.2343, -60.8249),
    (128, 128, 128): (53.5850, 0.0, 0.0),
}


def colour_cases(seed=0):
    """Every 8-bit grey, 19872 random 8-bit colours and 19872 non-integer RGB values
    (40000 pixels, as 200 rows of 200)."""
    rng = np.random.default_rng(seed)
    grey = np.repeat(np.arange(256, dtype=np.float32)[:, None], 3, axis=1)
    rand8 = rng.integers(0, 256, (19872, 3)).astype(np.float32)
    randf = rng.uniform(0, 255, (19872, 3)).astype(np.float32)
    return np.concatenate([grey, rand8, randf]).reshape(-1, 200, 3)


def lab_cases(seed=0):
    """L in [0, 100], a and b in [-128, 128]: in and out of the sRGB gamut."""
    rng = np.random.default_rng(seed)
    n = 40000
    lab = np.stack([rng.uniform(0, 100, n), rng.uniform(-128, 128, n), rng.uniform(-128, 128, n)], axis=1)
    return lab.astype(np.float32).reshape(-1, 200, 3)


def test_oracle_rgb_to_lab_published_values(oracle):
    for rgb, lab in PUBLISHED.items():
        got = oracle.rgb_to_lab(np.array(rgb, np.float32).reshape(1, 1, 3)).ravel()
        assert np.abs(got - np.array(lab)).max() < 0.01, (rgb, got, lab)


def test_oracle_round_trip_and_grey_ramp(oracle):
    img = colour_cases(1)
    back = oracle.lab_to_rgb(oracle.rgb_to_lab(img))
    assert np.abs(back - img).max() < 0.01
    grey = oracle.rgb_to_lab(np.repeat(np.arange(256, dtype=np.float32)[:, None], 3, axis=1).reshape(1, 256, 3))[0]
    assert np.all(np.diff(grey[:, 0]) > 0)            # lightness strictly increasing
    assert np.abs(grey[:, 1:]).max() < 0.01           # greys are achromatic


def test_oracle_passes_through_unconverted_images(oracle):
    rng = np.random.default_rng(2)
    for C in (1, 2, 4):
        img = rng.uniform(0, 255, (7, 5, C)).astype(np.float32)
        assert bitexact(oracle.rgb_to_lab(img), img)
    for C in (1, 2, 5):
        img = rng.uniform(-50, 100, (7, 5, C)).astype(np.float32)
        assert bitexact(oracle.lab_to_rgb(img), img)
    # four channels: the colour planes are converted, the fourth is ignored
    lab4 = np.concatenate([lab_cases(3)[:10], rng.uniform(0, 10, (10, 200, 1)).astype(np.float32)], axis=2)
    assert bitexact(oracle.lab_to_rgb(lab4), oracle.lab_to_rgb(np.ascontiguousarray(lab4[..., :3])))


def reference_color():
    import oracle as O
    if not os.path.exists(O.REF_COLOR_SO):
        pytest.skip("reference colour conversion not built in place (oracle/_ref, needs the reference sources)")
    return O.ReferenceColor()


def test_oracle_equals_reference_colour_conversion(oracle):
    ref = reference_color()
    img = colour_cases(4)
    assert bitexact(oracle.rgb_to_lab(img), ref.rgb_to_lab(img))
    lab = lab_cases(5)
    assert bitexact(oracle.lab_to_rgb(lab), ref.lab_to_rgb(lab))
    lab_real = oracle.rgb_to_lab(img)
    assert bitexact(oracle.lab_to_rgb(lab_real), ref.lab_to_rgb(lab_real))
