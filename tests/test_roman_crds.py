"""Roman CRDS ePSF reader (src/psf/roman_crds.jl, ext/CrowdPhotASDFExt.jl),
on fixtures written byte by byte to the ASDF 1.x layout, mirroring
test/psf/roman_crds_epsf_tests.jl (CPU; the optical-PSF conversion, which
correlates on the GPU, is in test_gpu_detection-style GPU tests below)."""
import numpy as np
import pytest

from _support import load_pkg
import asdf_fixture as af

pkg = load_pkg()
rc = pkg.roman_crds


def _markers():
    nx, ny, ngrid, nspec, ndef = 7, 5, 4, 2, 1
    data = np.full((ndef, nspec, ngrid, ny, nx), 0.001, np.float32)
    marker_xy = [(2, 4), (6, 1), (1, 1), (5, 3)]   # 1-based (x, y) per node
    for i, (mx, my) in enumerate(marker_xy):
        data[0, 0, i, my - 1, mx - 1] = 100.0
        data[0, 1, i, my - 1, mx - 1] = 200.0
    return data, marker_xy


@pytest.mark.parametrize("compression", [b"\0\0\0\0", b"zlib", b"bzp2"])
def test_read_asdf_round_trip(tmp_path, compression):
    a = np.random.default_rng(1).standard_normal((2, 3, 4, 5, 6)).astype(np.float32)
    p = af.epsf_file(tmp_path / "x.asdf", a, [0.0], [0.0], compression=compression)
    tree = rc.read_asdf(p)
    np.testing.assert_array_equal(tree["roman"]["psf"], a)
    assert tree["roman"]["meta"]["oversample"] == 2
    assert tree["asdf_library"]["name"] == "fixture"


def test_transpose_grid_and_slice_selection(tmp_path):
    data, marker_xy = _markers()
    p = af.epsf_file(tmp_path / "e.asdf", data, [0.0, 10.0, 0.0, 10.0], [0.0, 0.0, 20.0, 20.0])
    model = rc.roman_crds_gridded_epsf(p, defocus=0, spectral_type="G2V")
    assert isinstance(model, pkg.GriddedPSFModel) and len(model.psfs) == 4
    assert sorted(set(model.grid_y.tolist())) == [0.0, 20.0] and sorted(set(model.grid_x.tolist())) == [0.0, 10.0]
    for node, (mx, my) in zip(model.psfs, marker_xy):
        assert node.data.shape == (5, 7)
        assert np.unravel_index(np.argmax(node.data), node.data.shape) == (my - 1, mx - 1)
        assert node.oversampling == (2, 2)
    a0v = rc.roman_crds_gridded_epsf(p, defocus=0, spectral_type="A0V")
    assert a0v.psfs[0].data.max() < model.psfs[0].data.max()


def test_error_handling(tmp_path):
    data = np.full((1, 1, 1, 5, 5), 0.001, np.float32)
    data[0, 0, 0, 2, 2] = 10.0          # already pixel-integrated: no GPU conversion
    p = af.epsf_file(tmp_path / "e.asdf", data, [0.0], [0.0], spectral_type=("G2V",))
    with pytest.raises(pkg.ArgumentError, match="psf_subtype must be"):
        rc.roman_crds_gridded_epsf(p, psf_subtype="extended_psf")
    with pytest.raises(pkg.ArgumentError, match="defocus=99 not found"):
        rc.roman_crds_gridded_epsf(p, defocus=99)
    with pytest.raises(pkg.ArgumentError, match="spectral_type='Z9Z' not found"):
        rc.roman_crds_gridded_epsf(p, spectral_type="Z9Z")
    bad = af.epsf_file(tmp_path / "bad.asdf", data, [0.0], [0.0], spectral_type=("G2V",), roman_key="not_roman")
    with pytest.raises(pkg.ArgumentError, match='top-level "roman" key'):
        rc.roman_crds_gridded_epsf(bad)
    noipc = af.epsf_file(tmp_path / "n.asdf", data, [0.0], [0.0], spectral_type=("G2V",), subtype="psf_noipc")
    with pytest.raises(pkg.ArgumentError, match='does not contain "psf"'):
        rc.roman_crds_gridded_epsf(noipc)
    assert rc.roman_crds_gridded_epsf(noipc, psf_subtype="psf_noipc").psfs[0].data.shape == (5, 5)
    with open(tmp_path / "junk.asdf", "wb") as f:
        f.write(b"not asdf")
    with pytest.raises(pkg.ArgumentError, match="not an ASDF file"):
        rc.read_asdf(tmp_path / "junk.asdf")


def test_normalize_origin_oversample_and_new_format(tmp_path):
    data = np.full((1, 1, 1, 5, 5), 0.001, np.float32)
    data[0, 0, 0, 2, 2] = 10.0          # sums above oversample^2 / 2: used as is
    p = af.epsf_file(tmp_path / "e.asdf", data, [7.0], [3.0], spectral_type=("G2V",), oversample=3)
    model = rc.roman_crds_gridded_epsf(p)
    node = model.psfs[0]
    assert node.oversampling == (3, 3) and node.origin == (3.0, 3.0)
    assert model.grid_y.tolist() == [3.0] and model.grid_x.tolist() == [7.0]
    np.testing.assert_array_equal(node.data, data[0, 0, 0].astype(np.float64))
    assert rc.roman_crds_gridded_epsf(p, origin=(1.0, 1.0)).psfs[0].origin == (1.0, 1.0)
    assert rc.roman_crds_gridded_epsf(p, normalize=True).psfs[0].data.sum() == pytest.approx(9.0)
    assert node.data.sum() != pytest.approx(9.0)


@pytest.mark.gpu
def test_optical_psf_stamps_get_the_pixel_response(tmp_path):
    """Old-format stamps (sum ~ 1) are correlated with the pixel-response
    kernel (zero border) and scaled by oversample^2; total flux becomes
    oversample^2 (roman_crds_epsf_tests.jl:129-160)."""
    from scipy import ndimage
    os_ = 4
    stamp = np.zeros((15, 15), np.float32)
    stamp[7, 7] = 1.0
    p = af.epsf_file(tmp_path / "old.asdf", stamp[None, None, None], [0.0], [0.0], spectral_type=("G2V",),
                     oversample=os_)
    model = rc.roman_crds_gridded_epsf(p)
    k = pkg.pixel_response_kernel(os_)
    want = os_ ** 2 * ndimage.correlate(stamp.astype(np.float64), k, mode="constant")
    np.testing.assert_allclose(model.psfs[0].data, want, rtol=1e-6)
    assert model.psfs[0].data.sum() == pytest.approx(os_ ** 2, rel=1e-6)
