"""Run output files: HDF5 when h5py is importable, else a .npz with the same dataset names.

Mirrors what GPU_Wigner2D_FFT.Run writes (WriteHDF5_variables, save_Frame,
/Ehrenfest/* datasets; GPU_Wigner2D_FFT.py:660-663, 844-857, 1936-1957).
"""
from __future__ import annotations

import os

import numpy as np

try:  # pragma: no cover - h5py is absent in this image
    import h5py
except ImportError:  # pragma: no cover
    h5py = None


def output_path(file_name: str) -> str:
    if h5py is not None:
        return file_name
    return file_name if file_name.endswith(".npz") else file_name + ".npz"


class RunFile:
    """dict-like writer: f['dx'] = v; f.create_dataset(name, data=...); close()."""

    def __init__(self, file_name: str):
        self.path = output_path(file_name)
        self._data = {}
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        if os.path.exists(self.path):
            os.remove(self.path)

    def __setitem__(self, key, value):
        self._data[key.lstrip("/")] = np.asarray(value)

    def create_dataset(self, name, data):
        self[name] = data

    def close(self):
        if h5py is not None:  # pragma: no cover
            with h5py.File(self.path, "w") as f:
                for k, v in self._data.items():
                    f[k] = v
        else:
            np.savez(self.path, **self._data)


def read(file_name: str) -> dict:
    """All datasets of a run file (HDF5 or the .npz fallback)."""
    path = output_path(file_name) if not os.path.exists(file_name) else file_name
    if path.endswith(".npz"):
        with np.load(path, allow_pickle=False) as z:
            return {k: z[k] for k in z.files}
    with h5py.File(path, "r") as f:  # pragma: no cover
        out = {}
        f.visititems(lambda k, v: out.__setitem__(k, v[...]) if hasattr(v, "shape") else None)
        return out
