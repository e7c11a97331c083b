"""Host-side pieces of the learned boundary (no GPU): the FP32 reference
network's structural properties and the band geometry."""
import numpy as np

from biotfd import learned

FIELDS = ("U", "W", "QX", "QZ", "TXX", "TZZ", "TXZ", "P")


def test_reference_network_is_positively_homogeneous_and_unbiased():
    """No bias and leaky ReLU: the predicted change scales with positive
    factors of the input and vanishes for a zero band (so a boundary at rest
    stays at rest)."""
    rng = np.random.default_rng(0)
    w = learned.init_weights(1)
    sc = np.full(8, 0.5, np.float32)
    bands = rng.standard_normal((4, 40, learned.BAND)).astype(np.float32)
    prev = bands[3].reshape(40, 8, learned.CELLS)[:, :, :8]
    d1 = learned.torch_forward(w[0], sc, bands) - prev
    d2 = learned.torch_forward(w[0], sc, 4.0 * bands) - 4.0 * prev
    np.testing.assert_allclose(d2, 4.0 * d1, rtol=1e-4, atol=1e-6 * np.abs(d2).max())
    z = learned.torch_forward(w[0], sc, np.zeros_like(bands))
    np.testing.assert_array_equal(z, 0.0)
    # a 5-tap convolution three times: position p depends on p-6 .. p+6 only
    b2 = bands.copy()
    b2[:, 30:, :] += 1.0
    d3 = learned.torch_forward(w[0], sc, b2) - b2[3].reshape(40, 8, learned.CELLS)[:, :, :8]
    np.testing.assert_allclose(d3[:24], d1[:24], rtol=1e-5, atol=1e-7)
    assert np.abs(d3[24:] - d1[24:]).max() > 0


def test_band_geometry():
    """band_host lays out the 12-cell bands of the three sides as the device
    extraction does: the value at each band slot encodes the field and the
    grid cell it must come from."""
    mx, mzp, tpml = 20, 14, 10
    nz, nx = mzp + tpml + 1, mx + 2 * tpml
    ll, ii = np.meshgrid(np.arange(nz), np.arange(nx) + 1 - tpml, indexing="ij")
    fields = {f: (k * 1e6 + ll * 1e3 + ii + 500).astype(np.float64) for k, f in enumerate(FIELDS)}
    band = learned.band_host(fields, mx, mzp, tpml).astype(np.float64)
    lz, lx = mzp + 1, mx + 16
    assert band.size == (2 * lz + lx) * learned.BAND
    left = band[:lz * 96].reshape(lz, 8, 12)
    right = band[lz * 96:2 * lz * 96].reshape(lz, 8, 12)
    bottom = band[2 * lz * 96:].reshape(lx, 8, 12)

    def dec(v):
        return int(v // 1e6), int((v % 1e6) // 1e3), int(round(v % 1e3)) - 500

    assert dec(left[5, 2, 0]) == (2, 5, -7) and dec(left[5, 2, 7]) == (2, 5, 0)
    assert dec(left[5, 2, 11]) == (2, 5, 4)
    assert dec(right[0, 7, 0]) == (7, 0, mx + 8) and dec(right[0, 7, 8]) == (7, 0, mx)
    assert dec(bottom[0, 4, 0]) == (4, mzp + 8, -7) and dec(bottom[lx - 1, 4, 11]) == (4, mzp - 3, mx + 8)
