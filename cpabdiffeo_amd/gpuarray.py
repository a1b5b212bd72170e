"""Host/device array pair, the counterpart of the reference's ``CpuGpuArray``
(from the un-vendored ``of`` package, used throughout cpab/, e.g.
cpab/cpaNd/_CpaSpace.py:190-193).

``.cpu`` is a C-contiguous numpy array, ``.gpu`` a torch tensor on a ROCm
device holding the same shape and dtype; ``cpu2gpu`` / ``gpu2cpu`` copy one
into the other.  As in the reference, computations write ``.gpu`` only and
the caller runs ``gpu2cpu()`` when it wants the host copy.
"""
import numpy as np

from ._lib import CpabError


def torch_mod():
    import torch
    return torch


def require_gpu():
    torch = torch_mod()
    if not torch.cuda.is_available():
        raise CpabError("no ROCm GPU is visible: the HIP path has no CPU fallback")
    return torch


def device_of(device=None):
    torch = require_gpu()
    if device is None:
        return torch.device("cuda", torch.cuda.current_device())
    if isinstance(device, int):
        return torch.device("cuda", device)
    return torch.device(device)


_TORCH_DTYPES = {np.dtype(np.float64): "float64", np.dtype(np.float32): "float32",
                 np.dtype(np.int32): "int32", np.dtype(np.int64): "int64", np.dtype(np.uint8): "uint8"}


class CpuGpuArray:
    """A numpy array and its device copy."""

    def __init__(self, arr, dtype=None, device=None):
        self.cpu = np.array(arr, dtype=dtype, copy=True, order="C")
        if self.cpu.dtype not in _TORCH_DTYPES:
            raise TypeError(self.cpu.dtype)
        torch = require_gpu()
        self.gpu = torch.from_numpy(self.cpu).to(device_of(device))

    @classmethod
    def zeros(cls, shape, dtype=np.float64, device=None):
        return cls(np.zeros(shape, dtype=dtype), device=device)

    @classmethod
    def empty(cls, shape, dtype=np.float64, device=None):
        obj = cls.__new__(cls)
        obj.cpu = np.empty(shape, dtype=dtype)
        torch = require_gpu()
        obj.gpu = torch.empty(obj.cpu.shape, dtype=getattr(torch, _TORCH_DTYPES[obj.cpu.dtype]),
                              device=device_of(device))
        return obj

    @classmethod
    def zeros_like(cls, other):
        return cls.zeros(other.shape, dtype=other.dtype, device=other.gpu.device)

    @classmethod
    def from_gpu(cls, tensor):
        """Wrap an existing device tensor (the host copy is allocated, not filled)."""
        obj = cls.__new__(cls)
        obj.gpu = tensor.contiguous()
        obj.cpu = np.empty(tuple(tensor.shape), dtype=np.dtype(str(tensor.dtype).replace("torch.", "")))
        return obj

    # -------------------------------------------------------------- copies
    def cpu2gpu(self):
        torch = torch_mod()
        self.gpu.copy_(torch.from_numpy(self.cpu))
        return self

    def gpu2cpu(self):
        self.cpu[...] = self.gpu.detach().cpu().numpy()
        return self

    # ------------------------------------------------------------ ndarray-ish
    @property
    def shape(self):
        return self.cpu.shape

    @property
    def dtype(self):
        return self.cpu.dtype

    @property
    def ndim(self):
        return self.cpu.ndim

    @property
    def size(self):
        return self.cpu.size

    def __len__(self):
        return len(self.cpu)

    def ptr(self):
        return self.gpu.data_ptr()

    def __repr__(self):
        return "CpuGpuArray(shape=%s, dtype=%s, device=%s)" % (self.shape, self.dtype, self.gpu.device)


def dev_tensor(x, dtype, shape=None):
    """The device tensor behind x (a CpuGpuArray or a torch cuda tensor),
    checked for dtype, contiguity and optionally shape."""
    torch = torch_mod()
    t = x.gpu if isinstance(x, CpuGpuArray) else x
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("expected a CpuGpuArray or a ROCm tensor, got %s" % type(x))
    want = getattr(torch, dtype)
    if t.dtype != want:
        raise ValueError(t.dtype, want)
    if not t.is_contiguous():
        raise ValueError("Must be 'C'-contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(tuple(t.shape), tuple(shape))
    return t
