"""Device-tensor plumbing shared by the operator modules (torch is used only for
device memory, streams and copies)."""
import numpy as np
import torch

_NP2T = {
    np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64, np.dtype(np.float16): torch.float16,
    np.dtype(np.uint16): torch.uint16, np.dtype(np.uint8): torch.uint8, np.dtype(np.int32): torch.int32,
    np.dtype(np.int64): torch.int64, np.dtype(np.int16): torch.int16, np.dtype(np.uint32): torch.uint32,
    np.dtype(np.bool_): torch.bool,
}


def default_device():
    if not torch.cuda.is_available():
        raise RuntimeError("mermake_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def as_device(a, device=None):
    """Tensor on the GPU (numpy arrays and CPU tensors are copied)."""
    device = device or (a.device if isinstance(a, torch.Tensor) and a.is_cuda else default_device())
    if isinstance(a, torch.Tensor):
        return a if a.device == device else a.to(device)
    arr = np.ascontiguousarray(np.asarray(a))
    if arr.dtype not in _NP2T:
        raise TypeError(f"unsupported dtype {arr.dtype}")
    return torch.from_numpy(arr).to(device)


def unit_last_stride(t):
    """``t`` itself when its last axis has unit stride, else a contiguous copy."""
    if t.stride(-1) == 1 and all(s >= 0 for s in t.stride()):
        return t
    return t.contiguous()


def span(t):
    """[first, last] byte addresses touched by a strided tensor."""
    if t.numel() == 0:
        return t.data_ptr(), t.data_ptr() - 1
    last = sum((n - 1) * s for n, s in zip(t.shape, t.stride()))
    return t.data_ptr(), t.data_ptr() + (last + 1) * t.element_size() - 1


def shares_memory(a, b):
    if a.device != b.device:
        return False
    a0, a1 = span(a)
    b0, b1 = span(b)
    return not (a1 < b0 or b1 < a0)


def capture_graph(fn, stream):
    """hipGraph of the device work that ``fn`` enqueues on the current stream, captured on ``stream`` (a
    side stream with no work pending).  Unlike ``torch.cuda.graph`` this neither synchronises the device
    nor empties torch's memory cache, so a serving loop can capture while other streams run; the capture
    mode is thread-local, so other host threads may keep allocating and synchronising."""
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(stream):
        g.capture_begin(capture_error_mode="thread_local")
        try:
            fn()
        finally:
            g.capture_end()
    return g
