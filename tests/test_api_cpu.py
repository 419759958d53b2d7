"""Validation paths of the public API that run without a GPU (messages are the
reference's, matched by the same regexes as its tests)."""
import pytest
import torch

import torchmorph_amd as tm

DT_OPS = [tm.euclidean_distance_transform, tm.chamfer_distance_transform]


@pytest.mark.parametrize("op", DT_OPS)
def test_distance_transform_requires_cuda(op):
    with pytest.raises(ValueError, match="CUDA"):
        op(torch.zeros((1, 1, 3)))


def test_edt_rejects_removed_algorithm_argument():
    with pytest.raises(TypeError, match="algorithm"):
        tm.euclidean_distance_transform(torch.zeros((1, 1, 3)), algorithm="exact")


def test_sampling_normalisation():
    from torchmorph_amd.distance_transform import _normalize_sampling

    assert _normalize_sampling(None, 2) == [1.0, 1.0]
    assert _normalize_sampling(0.5, 2) == [0.5, 0.5]
    assert _normalize_sampling([0.5], 3) == [0.5] * 3
    for bad in ([1.0, 2.0, 3.0], 0.0, -1.0, float("inf"), float("nan")):
        with pytest.raises(ValueError, match="sampling"):
            _normalize_sampling(bad, 2)


def test_public_names_match_reference():
    ref_names = {
        "euclidean_distance_transform", "chamfer_distance_transform",
        "brute_force_distance_transform",
    }
    assert ref_names <= set(tm.__all__)


def test_ragged_grey_validation():
    from torchmorph_amd.ragged import RaggedGreyMorphology

    with pytest.raises(ValueError, match="every item must have 2 axes"):
        RaggedGreyMorphology([(4, 4), (4, 4, 4)], size=3, device="cuda:0")
    with pytest.raises(ValueError, match="1..3 axes"):
        RaggedGreyMorphology([(2, 2, 2, 2)], size=3, device="cuda:0")
    with pytest.raises(ValueError, match="operation must be"):
        RaggedGreyMorphology([(4, 4)], "gradient", size=3, device="cuda:0")
    with pytest.raises(ValueError, match="Structure dimension 3 must match spatial dimension 2"):
        RaggedGreyMorphology([(4, 4)], structure=torch.zeros(3, 3, 3), device="cuda:0")
    with pytest.raises(ValueError, match="At least one of size"):
        RaggedGreyMorphology([(4, 4)], device="cuda:0")
    with pytest.raises(ValueError, match="CUDA devices only"):
        RaggedGreyMorphology([(4, 4)], size=3, device="cpu")
