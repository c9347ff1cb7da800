"""Host-side logic of the binding that needs no device: the IndexFunc sampling
of PropagateSplitStep (optics/medium.go:22-65), which detects a z-invariant
medium from the sampled planes instead of taking a flag, checked against the
index planes the GPU parity tests pass directly."""
import numpy as np

import wos_hip as W


def graded_index_fn(x, y, z):
    """Config-3 medium as an IndexFunc: n = 1.45 - 0.01 (r/100um)^2 for
    r < 100 um, else 1.44, plus 1e-5 i (z-invariant)."""
    r = np.hypot(x, y)
    return np.where(r < 100e-6, 1.45 - 0.01 * (r / 100e-6) ** 2, 1.44) + 1e-5j


def graded_index_planes(n, dx):
    """The plane tests/test_gpu_parity.py builds for the same medium."""
    x = (np.arange(n) - n / 2) * dx
    r = np.hypot(x[None, :], x[:, None])
    return np.where(r < 100e-6, 1.45 - 0.01 * (r / 100e-6) ** 2, 1.44) + 1e-5j


def test_z_invariant_medium_is_one_plane():
    for n, dx, steps in ((128, 0.5e-6, 40), (1024, 0.5e-6, 40), (100, 0.5e-6, 3)):
        planes = W.sample_index_planes(graded_index_fn, n, dx, steps * 5e-6, steps)
        assert planes.shape == (1, n, n)
        assert np.array_equal(planes[0], graded_index_planes(n, dx))
    uniform = W.sample_index_planes(lambda x, y, z: 1.33 + 0j, 64, 1e-6, 1e-3, 10)   # UniformIndex
    assert planes.dtype == np.complex128 and uniform.shape == (1, 64, 64) and np.all(uniform == 1.33)


def test_stratified_medium_keeps_every_plane():
    """TestSplitStepStratified's medium (optics/medium_test.go:45-66): n = 1.5
    below z/2 and 1.0 above, sampled at every slice centre; the planes equal
    those the stratified GPU parity test passes."""
    wl_z, steps, n, dx = 2e-3, 8, 128, 2e-6
    dz = wl_z / steps
    fn = lambda x, y, zc: np.where(zc < wl_z / 2, 1.5, 1.0) + 2e-6j + 0 * x * y
    planes = W.sample_index_planes(fn, n, dx, wl_z, steps)
    want = np.array([np.full((n, n), (1.5 if (wl_z * s / steps + dz / 2) < wl_z / 2 else 1.0) + 2e-6j)
                     for s in range(steps)])
    assert planes.shape == (steps, n, n)
    assert np.array_equal(planes, want)


def test_a_change_in_the_last_slice_keeps_every_plane():
    z, steps = 1e-3, 6
    fn = lambda x, y, zc: np.where(zc > z * (steps - 1) / steps, 1.2, 1.1) + 0 * x + 0 * y
    planes = W.sample_index_planes(fn, 32, 1e-6, z, steps)
    assert planes.shape == (steps, 32, 32)
    assert np.all(planes[:-1] == 1.1) and np.all(planes[-1] == 1.2)


def test_slice_centres_and_coordinates_follow_the_reference():
    """zc = z*s/steps + dz/2 (optics/medium.go:35) and X(i) = (i - N/2) dx
    (optics/field.go:58-61); steps <= 0 is one slice."""
    seen = []

    def fn(x, y, zc):
        seen.append(zc)
        return x + 1j * y + 0 * zc

    n, dx, z, steps = 8, 1e-6, 2e-3, 4
    planes = W.sample_index_planes(fn, n, dx, z, steps)
    assert seen == [z * s / steps + (z / steps) / 2 for s in range(steps)]
    c = (np.arange(n) - n / 2) * dx
    assert planes.shape == (1, n, n)
    assert np.array_equal(planes[0].real, np.broadcast_to(c[None, :], (n, n)))
    assert np.array_equal(planes[0].imag, np.broadcast_to(c[:, None], (n, n)))
    seen.clear()
    assert W.sample_index_planes(fn, n, dx, z, 0).shape == (1, n, n) and seen == [z / 2]
