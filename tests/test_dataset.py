"""In-memory Dataset / ImageProp / ImageData and the selection checks of image_synthesis
(no GPU: the checks run before anything reaches the device)."""
import numpy as np
import pytest

import bipp_amd
from bipp_amd import synthesis


def make_dataset(n_ant=4, n_beam=3, n=3):
    r = np.random.default_rng(0)
    ds = bipp_amd.InMemoryDataset.create("test", n_ant, n_beam, 10.0, 20.0)
    for k in range(n):
        v = r.normal(size=(n_ant, n_beam)) + 1j * r.normal(size=(n_ant, n_beam))
        ds.write(float(k), 2.0 + k, 0.25, v, r.normal(size=n_beam), r.normal(size=(n_ant ** 2, 3)))
    return ds


def test_dataset_stores_file_precision():
    """Values are kept as the dataset file stores them (src/dataset_file.cpp:41-73)."""
    ds = make_dataset()
    assert ds.num_samples() == 3 and ds.num_antenna() == 4 and ds.num_beam() == 3
    assert ds.eig_vec(1).dtype == np.complex64 and ds.eig_vec(1).flags.f_contiguous
    assert ds.eig_val(1).dtype == np.float32 and ds.uvw(2).shape == (16, 3)
    assert ds.uvw(2).dtype == np.float32
    assert ds.wl(2) == 4.0 and ds.scale(0) == np.float32(0.25) and ds.time_stamp(1) == 1.0
    assert ds.ra_deg() == 10.0 and ds.dec_deg() == 20.0 and ds.description() == "test"
    with pytest.raises(bipp_amd.InvalidParameterError):
        ds.eig_val(3)
    with pytest.raises(bipp_amd.InvalidParameterError):
        ds.write(0.0, 1.0, 1.0, np.zeros((4, 2)), np.zeros(3), np.zeros((16, 3)))


def test_image_prop_and_data_layout():
    lmn = np.arange(6 * 3, dtype=np.float32).reshape(6, 3)
    prop = bipp_amd.InMemoryImageProp.create(2, 3, 5.0, lmn)
    assert prop.width() == 3 and prop.height() == 2 and prop.fov_deg() == 5.0
    assert np.array_equal(prop.pixel_lmn(), lmn)
    prop.set_meta("n", 3)
    assert prop.meta_data() == {"n": 3}
    data = bipp_amd.InMemoryImageData.create(2, 3, 5.0, 1.0, 2.0)
    data.set("a", np.arange(6, dtype=np.float32))
    img = data.get("a")
    # pixel q is entry (q mod height, q / height), as ImageDataFile's column-major images
    assert img.shape == (2, 3) and img[1, 2] == 5 and img[0, 1] == 2
    assert data.tags() == ["a"] and data.num_tags() == 1
    with pytest.raises(bipp_amd.InvalidParameterError):
        data.set("b", np.zeros(5))


def test_selection_checks():
    """src/image_synthesis.cpp:37-104 and python/pybipp.cpp:181-190."""
    ok = {2: np.ones(3), 0: np.ones(3)}
    names, ids, d = synthesis.check_selection({"img": ok, "b": [(0, [1, 2, 3]), (2, [4, 5, 6])]},
                                              3, 3)
    assert names == ["img", "b"] and ids == [0, 2] and d.shape == (2, 2, 3)
    assert d.dtype == np.float32 and np.array_equal(d[1, 1], [4, 5, 6])
    bad = [
        ({"a/b": ok}, "'/'"),
        ({"a.b": ok}, "'.'"),
        ({"a b": ok}, "spaces"),
        ({"a": {}}, "Empty selection"),
        ({"a": [(1, np.ones(3)), (1, np.ones(3))]}, "duplicate"),
        ({"a": {3: np.ones(3)}}, "exceeds"),
        ({"a": {0: np.ones(2)}}, "does not match number of beams"),
        ({"a": {0: np.ones(3)}, "b": {1: np.ones(3)}}, "same sample ids"),
        ({}, "Empty selection"),
    ]
    for sel, msg in bad:
        with pytest.raises(bipp_amd.InvalidParameterError, match=msg):
            synthesis.check_selection(sel, 3, 3)


def test_image_synthesis_checks_before_device():
    ds = make_dataset()
    prop = bipp_amd.InMemoryImageProp.create(2, 2, 5.0, np.zeros((4, 3)))
    with pytest.raises(bipp_amd.InvalidParameterError, match="duplicate"):
        bipp_amd.image_synthesis(None, bipp_amd.NufftSynthesisOptions(), ds,
                                 {"a": [(0, np.ones(3)), (0, np.ones(3))]}, prop)
