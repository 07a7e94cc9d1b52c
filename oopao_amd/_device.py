"""Device plumbing shared by the drop-in classes: default device, host/device conversion.

Arrays live on the GPU as float32 torch tensors; batched maps are stored
batch-major [B, n, n]. The reference's numpy conventions ([n, n] or [n, n, B],
float64) are produced on access.
"""
import os

import numpy as np
import torch

from .tools import OopaoError, WorkspaceAllocationError


def default_device():
    if not torch.cuda.is_available():
        raise OopaoError('oopao_amd needs a HIP device (MI355X); none is visible')
    return torch.device('cuda', torch.cuda.current_device())


def alloc_workspace(nbytes, device):
    """uint8 device workspace of nbytes; raises WorkspaceAllocationError when the allocation fails or when
    nbytes exceeds the cap given by OOPAO_AMD_MAX_WORKSPACE_BYTES (a host-side way to exercise the
    recovery paths without exhausting the device)."""
    cap = os.environ.get('OOPAO_AMD_MAX_WORKSPACE_BYTES')
    if cap is not None and nbytes > int(cap):
        raise WorkspaceAllocationError('workspace of %d bytes exceeds OOPAO_AMD_MAX_WORKSPACE_BYTES=%s'
                                       % (nbytes, cap))
    try:
        return torch.empty(nbytes, dtype=torch.uint8, device=device)
    except torch.OutOfMemoryError as e:
        raise WorkspaceAllocationError('could not allocate a workspace of %d bytes on %s: %s'
                                       % (nbytes, device, e)) from e


def to_dev(val, device, dtype=torch.float32):
    """numpy [n,n] / [n,n,B] or torch -> device tensor ([n,n] / [B,n,n])."""
    if val is None:
        return None
    if isinstance(val, torch.Tensor):
        return val.to(device=device, dtype=dtype)
    a = np.asarray(val)
    if a.ndim == 3:
        a = np.moveaxis(a, -1, 0)
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32 if dtype == torch.float32 else np.float64)).to(device)


def to_host(t, dtype=np.float64):
    """device tensor ([n,n] / [B,n,n]) -> numpy float64 ([n,n] / [n,n,B])."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        return t
    a = t.detach().cpu().numpy().astype(dtype)
    if a.ndim == 3:
        a = np.moveaxis(a, 0, -1)
    return a


def stream_ptr(device):
    import ctypes
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())
