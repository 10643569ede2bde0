"""Structural checks of the netcon_bitpack_ops drop-in (CPU: no kernel runs;
the checks read no tensor values, as the reference's contracts do)."""

import pytest
import torch


def test_structural_errors_come_before_any_launch():
    from dendra_amd import netcon_bitpack_ops as ops

    h = torch.zeros((4, 2), dtype=torch.int64)
    cur = torch.zeros(1, dtype=torch.int64)
    with pytest.raises(TypeError, match="bool"):
        ops.pack_source_spikes(torch.zeros(100), h, cur)
    with pytest.raises(ValueError, match="words"):
        ops.pack_source_spikes(torch.zeros(200, dtype=torch.bool), h, cur)
    with pytest.raises(TypeError, match="one-element"):
        ops.pack_source_spikes(torch.zeros(100, dtype=torch.bool), h, torch.zeros(2).long())
    with pytest.raises(RuntimeError, match="GPU tensor"):
        ops.pack_source_spikes(torch.zeros(100, dtype=torch.bool), h, cur)
    i = torch.zeros(5, dtype=torch.int64)
    w = torch.zeros(5)
    with pytest.raises(TypeError, match="torch.long"):
        ops.build_delivery(h, cur, i.int(), i, i, i, w, torch.zeros(3))
    with pytest.raises(TypeError, match="share a dtype"):
        ops.build_delivery(h, cur, i, i, i, i, w, torch.zeros(3, dtype=torch.float64))
    with pytest.raises(TypeError, match="float32 or float64"):
        ops.build_delivery(h, cur, i, i, i, i, w.half(), torch.zeros(3).half())
    with pytest.raises(ValueError, match="connections"):
        ops.build_delivery(h, cur, i, i[:4], i, i, w, torch.zeros(3))
    with pytest.raises(RuntimeError, match="GPU tensor"):
        ops.build_delivery_uniform(h, cur, 2, i, i, i, w, torch.zeros(3))


def test_availability_reports_why_not():
    from dendra_amd import netcon_bitpack_ops as ops

    if not torch.cuda.is_available():
        assert ops.is_available() is False
        assert isinstance(ops.last_error(), RuntimeError)
