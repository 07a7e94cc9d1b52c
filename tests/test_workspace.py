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


def test_engines_share_one_workspace():
    # the four Bi-O Edge engines run one after another on one stream and share one workspace: the bytes
    # held are those of the largest request, not the sum over engines, and a release frees them all
    from oopao_amd.pwfs_engine import PWFSEngine, Workspace
    ws = Workspace(torch.device('cpu'))
    engines = [PWFSEngine(80, 320, 80, device='cpu', workspace=ws) for _ in range(4)]
    sizes = [5000, 7000, 3000, 7000]
    bufs = [e._scratch(nb) for e, nb in zip(engines, sizes)]
    assert ws.nbytes == max(sizes)
    assert bufs[1].data_ptr() == bufs[2].data_ptr() == bufs[3].data_ptr()
    assert all(e.workspace is ws for e in engines)
    ws.release()
    assert ws.nbytes == 0
    # separate engines keep separate workspaces
    a, b = PWFSEngine(80, 320, 80, device='cpu'), PWFSEngine(80, 320, 80, device='cpu')
    a._scratch(100)
    assert b.workspace.nbytes == 0
