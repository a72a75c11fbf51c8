"""Host-side validation of packed-ray CSR offsets (no GPU needed: every rejection is
raised before a kernel is launched)."""
import pytest
import torch

from neuralangelo_amd import functional as Fn


@pytest.mark.parametrize("offsets,S,msg", [
    (torch.tensor([0, 3, 5], dtype=torch.int64), 5, "int32"),
    (torch.tensor([[0, 3, 5]], dtype=torch.int32), 5, "1-D"),
    (torch.tensor([1, 3, 5], dtype=torch.int32), 5, "start at 0"),
    (torch.tensor([0, 3, 6], dtype=torch.int32), 5, "start at 0"),   # past the samples
    (torch.tensor([0, 3, 4], dtype=torch.int32), 5, "start at 0"),   # samples left over
    (torch.tensor([0, 4, 2, 5], dtype=torch.int32), 5, "start at 0"),  # decreasing
])
def test_bad_offsets_are_rejected(offsets, S, msg):
    with pytest.raises(RuntimeError, match=msg):
        Fn.check_offsets(offsets, S)
    with pytest.raises(RuntimeError, match=msg):
        Fn.packed_ray_index(offsets, S)


def test_valid_offsets_pass_the_checks():
    """Valid offsets (empty rays included) pass every value check; on the CPU the
    product path then refuses to run without the GPU."""
    with pytest.raises(RuntimeError, match="GPU"):
        Fn.check_offsets(torch.tensor([0, 0, 3, 3, 7], dtype=torch.int32), 7)
