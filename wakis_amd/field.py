"""Device-resident ``Field``: the reference's flat 3N vector-field container.

Mirrors ``wakis.field.Field`` (field.py:19-570): one flat array of 3N values,
x|y|z component blocks, each block x-fastest (``n = i + Nx*j + Nx*Ny*k``).
The array lives in HBM as a torch tensor so that the HIP kernels update it in
place. Reads through ``__getitem__``/``get_abs`` return host numpy copies, as
the reference's GPU path does (``.get()``, field.py:243-249); writes accept
numpy arrays, scalars or torch tensors.
"""

import numpy as np
import torch

_TORCH_DTYPE = {np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32,
                np.dtype(np.int8): torch.int8, np.dtype(np.int32): torch.int32,
                np.dtype(np.int64): torch.int64}


def torch_dtype(dtype):
    return _TORCH_DTYPE[np.dtype(dtype)]


def default_device():
    if not torch.cuda.is_available():
        raise RuntimeError("wakis_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


_CIDX = {0: 0, 1: 1, 2: 2, "x": 0, "y": 1, "z": 2}


class Field:
    """Flat 3N field on the GPU (field.py:19-76)."""

    def __init__(self, Nx, Ny, Nz, dtype=float, use_ones=False, use_gpu=True, device=None):
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        self.N = self.Nx * self.Ny * self.Nz
        self.dtype = np.dtype(dtype).type
        self.on_gpu = True
        self.device = device if device is not None else default_device()
        fill = torch.ones if use_ones else torch.zeros
        self.array = fill(3 * self.N, dtype=torch_dtype(self.dtype), device=self.device)
        self.version = 0  # bumped by every write through this API
        self.xp = torch  # the array module (numpy or cupy in the reference)

    # -- component views (field.py:78-116) --------------------------------
    def _block(self, c):
        c = _CIDX[c]
        return self.array[c * self.N:(c + 1) * self.N]

    @property
    def field_x(self):
        return self._block(0)

    @property
    def field_y(self):
        return self._block(1)

    @property
    def field_z(self):
        return self._block(2)

    def _set_block(self, c, value):
        if hasattr(value, "shape") and len(value.shape) > 1:
            self.from_matrix(value, c)
        else:
            self._block(c).copy_(self._as_tensor(value))
            self.version += 1

    @field_x.setter
    def field_x(self, value):
        self._set_block(0, value)

    @field_y.setter
    def field_y(self, value):
        self._set_block(1, value)

    @field_z.setter
    def field_z(self, value):
        self._set_block(2, value)

    # -- conversions --------------------------------------------------------
    def _as_tensor(self, value):
        if isinstance(value, torch.Tensor):
            return value.to(device=self.device, dtype=self.array.dtype)
        return torch.as_tensor(np.asarray(value), device=self.device).to(self.array.dtype)

    def toarray(self):
        """The flat device array (a torch tensor view, like cupy's on GPU)."""
        return self.array

    def to_numpy(self):
        """Host copy of the flat array."""
        return self.array.detach().cpu().numpy()

    def fromarray(self, array):
        """Copy a flat 3N array in (field.py:130-137)."""
        self.array.copy_(self._as_tensor(array).reshape(-1))
        self.version += 1

    def to_matrix(self, key):
        """(Nx,Ny,Nz) device view of one component (field.py:139-168)."""
        if key not in _CIDX:
            raise IndexError("Component id not valid")
        return self._block(key).view(self.Nz, self.Ny, self.Nx).permute(2, 1, 0)

    def from_matrix(self, mat, key):
        """Set one component from an (Nx,Ny,Nz) matrix (field.py:170-190)."""
        if key not in _CIDX:
            raise IndexError("Component id not valid")
        self.to_matrix(key).copy_(self._as_tensor(mat))
        self.version += 1

    def to_gpu(self):
        """No-op: the array always lives on the GPU."""

    def from_gpu(self):
        raise RuntimeError("wakis_amd Field arrays stay on the GPU; use to_numpy()")

    # -- indexing (field.py:223-348) -----------------------------------------
    def __getitem__(self, key):
        if type(key) is tuple:
            if len(key) != 4:
                raise IndexError("Need 3 indexes and component to access the field")
            comp = key[3]
            if isinstance(comp, str) and comp.lower() == "abs":
                return self.get_abs()[key[0], key[1], key[2]]
            if comp not in _CIDX:
                raise IndexError("Component id not valid")
            out = self.to_matrix(comp)[key[0], key[1], key[2]]
            return out.cpu().numpy()
        if isinstance(key, (int, np.integer)):
            if key <= self.N:
                return self.array[int(key)].cpu().numpy()
            raise IndexError("Lexico-graphic index cannot be higher than product of dimensions")
        if type(key) is slice:
            return self.array[key].cpu().numpy()
        raise ValueError("key must be a 3-tuple or an integer")

    def __setitem__(self, key, value):
        if type(key) is tuple:
            if len(key) != 4:
                raise IndexError("Need 3 indexes and component to access the field")
            m = self.to_matrix(key[3])
            if isinstance(value, np.generic):
                value = value.item()
            v = value if np.isscalar(value) else self._as_tensor(value)
            m[key[0], key[1], key[2]] = v
        elif isinstance(key, (int, np.integer)):
            if key <= self.N:
                self.array[int(key)] = float(value)
            else:
                raise IndexError(
                    "Lexico-graphic index cannot be higher than product of dimensions")
        elif type(key) is slice:
            if isinstance(value, np.generic):
                value = value.item()
            self.array[key] = value if np.isscalar(value) else self._as_tensor(value)
        else:
            raise IndexError("key must be a 3-tuple or an integer")
        self.version += 1

    # -- arithmetic (field.py:350-460) ------------------------------------------
    def _new(self, array, dtype=None):
        f = Field.__new__(Field)
        f.Nx, f.Ny, f.Nz, f.N = self.Nx, self.Ny, self.Nz, self.N
        f.dtype = np.dtype(dtype).type if dtype is not None else self.dtype
        f.on_gpu, f.device, f.version, f.xp = True, self.device, 0, torch
        f.array = array.to(torch_dtype(f.dtype)).contiguous()
        return f

    def _operand(self, other):
        if isinstance(other, Field):
            return other.array
        if isinstance(other, (int, float)):
            return other
        t = self._as_tensor(other) if not isinstance(other, torch.Tensor) else other.to(self.device)
        if t.dim() > 1:  # (Nx,Ny,Nz) matrix applied to every component
            t = t.permute(2, 1, 0).reshape(-1).repeat(3)
        return t

    def __mul__(self, other, dtype=None):
        return self._new(self.array * self._operand(other), dtype)

    def __truediv__(self, other, dtype=None):
        return self._new(self.array / self._operand(other), dtype)

    __div__ = __truediv__

    def __add__(self, other, dtype=None):
        return self._new(self.array + self._operand(other), dtype)

    def __sub__(self, other, dtype=None):
        return self._new(self.array - self._operand(other), dtype)

    def copy(self):
        return self._new(self.array.clone())

    def compute_ijk(self, n):
        if n > self.N:
            raise IndexError("Lexico-graphic index cannot be higher than product of dimensions")
        k = n // (self.Nx * self.Ny)
        i = (n - k * self.Nx * self.Ny) % self.Nx
        j = (n - k * self.Nx * self.Ny) // self.Nx
        return i, j, k

    def get_abs(self, as_matrix=True):
        """|F| as a host numpy array (field.py:536-570)."""
        if as_matrix:
            a = torch.sqrt(self.to_matrix("x") ** 2 + self.to_matrix("y") ** 2
                           + self.to_matrix("z") ** 2)
        else:
            a = torch.sqrt(self.field_x ** 2 + self.field_y ** 2 + self.field_z ** 2)
        return a.cpu().numpy()

    def __str__(self):
        """The three components as host arrays, one after another (field.py:475-486)."""
        return ("x:\n" + str(self.to_matrix("x").cpu().numpy()) + "\n" + "y:\n"
                + str(self.to_matrix("y").cpu().numpy()) + "\n" + "z:\n"
                + str(self.to_matrix("z").cpu().numpy()))

    def __repr__(self):
        return f"Field({self.Nx}x{self.Ny}x{self.Nz}, {np.dtype(self.dtype).name}, {self.device})"
