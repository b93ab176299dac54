"""The MI355X backend: device arrays in the reference's device layout, on one GPU.

Mirrors the backend registry (``peregrine/backend/__init__.py:8-9``: a backend
is chosen by name) and the allocation interface (``peregrine/backend/allocation.py``:
``shape``, ``get()``, ``set()``). Memory is owned by the host side (torch
tensors on the device) and lent to the HIP library as raw pointers; the
library never allocates or frees it.

Layout: an array of host-visible shape (ni, nj, nk, ncomp) is held in Fortran
order, i fastest and each component a separate plane
(``peregrine/backend/device/backend.py:56-57``). Each i-row is padded to a
multiple of 16 doubles (128 B, one L2 cache line) and the array is placed so
that its first interior index i = lead starts a line: a wave's 64 lanes along
i then read four whole lines rather than touching five. The padding lies
between rows, never inside one; the library takes the strides with the
pointer.
"""

import numpy as np

from peregrine_amd import native
from peregrine_amd.misc import find_subclass

ROW_ALIGN = 16


class BaseBackend:
    name = None

    def __init__(self, cfg):
        self.cfg = cfg


def get_backend(name, cfg):
    return find_subclass(BaseBackend, name=name)(cfg)


def _torch():
    import torch

    return torch


class DeviceArray:
    """A block array on the device: host shape (ni, nj, nk, ncomp), device planes per component."""

    def __init__(self, backend, shape, name=None, lead=0):
        torch = _torch()
        self.backend = backend
        self.shape = tuple(int(n) for n in shape)
        self.name = name
        ni, nj, nk, nc = self.shape
        sj = -(-ni // ROW_ALIGN) * ROW_ALIGN
        self.strides = (sj, sj * nj, sj * nj * nk)
        off = (-int(lead)) % ROW_ALIGN
        # the allocator hands out blocks aligned to at least 128 B
        buf = torch.zeros(off + self.strides[2] * nc, dtype=torch.float64, device=backend.device)
        self.t = buf.as_strided((nc, nk, nj, ni), self.strides[::-1] + (1,), off)

    def __repr__(self):
        return f"<DeviceArray {self.name} {self.shape}>"

    @property
    def ptr(self):
        return self.t.data_ptr()

    def view(self):
        return native.view(self.ptr, *self.shape, strides=self.strides)

    def get(self, component=None):
        self.backend.synchronize()
        host = self.t.cpu().numpy().transpose(3, 2, 1, 0)
        return host if component is None else host[..., component]

    def set(self, values):
        torch = _torch()
        values = np.asarray(values, dtype=np.float64)
        if values.shape != self.shape:
            raise ValueError(f"{self.name}: values of shape {values.shape}, the array is {self.shape}")
        host = torch.from_numpy(np.ascontiguousarray(values.transpose(3, 2, 1, 0)))
        self.t.copy_(host)
        self.backend.synchronize()


class DeviceScalar:
    """A double on the device that kernels read through a pointer; the host keeps the value."""

    def __init__(self, backend, name, value=0.0):
        torch = _torch()
        self.backend, self.name = backend, name
        self.t = torch.zeros(1, dtype=torch.float64, device=backend.device)
        self.value = None
        self.set(value)

    @property
    def ptr(self):
        return self.t.data_ptr()

    def set(self, value):
        value = float(value)
        if value != self.value:
            self.t.fill_(value)
            self.value = value

    def get(self):
        return self.value


class HostScalar:
    """The clock: held on the host in double, as the reference's time is."""

    def __init__(self, value=0.0):
        self.value = float(value)

    def set(self, value):
        self.value = float(value)

    def get(self):
        return self.value


class Mi355xBackend(BaseBackend):
    name = "mi355x"

    def __init__(self, cfg):
        super().__init__(cfg)
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("the mi355x backend needs a GPU; none is visible to this process")
        native.lib()
        sect = cfg["backend-mi355x"]
        self.device = torch.device("cuda", int(sect["device"]))
        torch.cuda.set_device(self.device)
        self.torch = torch

    @property
    def stream(self):
        """The HIP stream every launch of this backend goes on: torch's current one."""
        return self.torch.cuda.current_stream(self.device).cuda_stream

    def alloc(self, shape, name=None, lead=0):
        return DeviceArray(self, shape, name, lead)

    def scalar(self, name, value=0.0):
        return DeviceScalar(self, name, value)

    def tensor(self, values, dtype=None):
        torch = self.torch
        return torch.as_tensor(np.asarray(values), dtype=dtype or torch.float64).to(self.device)

    def synchronize(self):
        self.torch.cuda.synchronize(self.device)

    def stop(self):
        self.synchronize()
