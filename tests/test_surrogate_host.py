"""Host-side argument checks of the surrogate wrappers (no GPU needed: they raise before any kernel runs)."""
import pytest
import torch

from xadvcool_amd import surrogate as S


def test_mse_loss_refuses_a_weight_mask_without_channels():
    pred = torch.zeros(1, 4, 4, 4, 16)
    target = torch.zeros(1, 4, 4, 4, S.NOUT)
    with pytest.raises(ValueError, match="mse_loss"):
        S.mse_loss(pred, target, torch.ones(1, 4, 4, 4))  # the kernel reads (V, NOUT) weights


def test_mse_loss_refuses_mismatched_voxels():
    pred = torch.zeros(1, 4, 4, 4, 16)
    with pytest.raises(ValueError, match="mse_loss"):
        S.mse_loss(pred, torch.zeros(1, 4, 4, 3, S.NOUT), torch.ones(1, 4, 4, 3, S.NOUT))
