"""Workspace allocation failures (CPU): the cap in OOPAO_AMD_MAX_WORKSPACE_BYTES forces the failure from
the host, and the error is the OopaoError subclass the Pyramid's batch-halving recovery catches."""
import pytest
import torch


def test_alloc_workspace_cap(monkeypatch):
    from oopao_amd._device import alloc_workspace
    from oopao_amd.tools import OopaoError, WorkspaceAllocationError
    monkeypatch.setenv('OOPAO_AMD_MAX_WORKSPACE_BYTES', '4096')
    ws = alloc_workspace(4096, torch.device('cpu'))
    assert ws.dtype == torch.uint8 and ws.numel() == 4096
    with pytest.raises(WorkspaceAllocationError) as e:
        alloc_workspace(4097, torch.device('cpu'))
    assert isinstance(e.value, OopaoError) and '4097' in e.value.string
    monkeypatch.delenv('OOPAO_AMD_MAX_WORKSPACE_BYTES')
    assert alloc_workspace(1 << 20, torch.device('cpu')).numel() == 1 << 20
