"""`sensitivity.jacobian_rows` in isolation (the reference's tests/test_sensitivity.py:9-40
protocol): one row equals autograd, k rows equal per-row autograd, a leaf an output does
not reach gets an exactly-zero block of the right shape.  Pure torch, so on the CPU."""

import torch


def _sens():
    from sttopt_amd import sensitivity

    return sensitivity


def test_jacobian_rows_single_output_matches_autograd():
    a = torch.tensor([1.0, 2.0, 3.0], requires_grad=True)
    b = torch.tensor([4.0, 5.0], requires_grad=True)
    out = (a**2).sum() + (b**3).sum()
    ja, jb = _sens().jacobian_rows(out[None], (a, b))
    assert ja.shape == (1, 3) and jb.shape == (1, 2)
    torch.testing.assert_close(ja[0], 2 * a.detach())
    torch.testing.assert_close(jb[0], 3 * b.detach() ** 2)


def test_jacobian_rows_several_outputs_match_per_row_autograd():
    a = torch.tensor([1.0, 2.0, 3.0], requires_grad=True)
    outputs = torch.stack([a[0] ** 2, a[1] * a[2], a.sum()])
    (ja,) = _sens().jacobian_rows(outputs, (a,))
    want = torch.tensor([[2.0, 0.0, 0.0], [0.0, 3.0, 2.0], [1.0, 1.0, 1.0]])
    torch.testing.assert_close(ja, want)
    for i in range(3):
        (g,) = torch.autograd.grad(outputs[i], (a,), retain_graph=True)
        torch.testing.assert_close(ja[i], g)


def test_jacobian_rows_unused_leaf_block_is_zero():
    a = torch.tensor([1.0, 2.0], requires_grad=True)
    b = torch.tensor([[3.0, 4.0], [5.0, 6.0]], requires_grad=True)
    out = (a**2).sum()
    ja, jb = _sens().jacobian_rows(out[None], (a, b))
    assert jb.shape == (1, 4) and jb.dtype == b.dtype
    assert torch.all(jb == 0)
    torch.testing.assert_close(ja[0], 2 * a.detach())
