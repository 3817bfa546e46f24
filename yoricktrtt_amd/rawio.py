"""Raw binary data files read by the TRTT scripts.

Yorick's ``raw_read(file, type, d1, d2, ..., offset=n)`` reads
d1*d2*... values of ``type`` in native byte order after n header bytes,
d1 varying fastest.  The scripts use it for
* 3-D projections: ``raw_read(data_file, float, nu, nv, 720, offset=512)``
  then every ``each``-th view (Tomography_3D/Ytrtt3D_script_reconstructions.i:50),
  each view reversed along u and transposed (:52-54) on request;
* 3-D reference volumes: ``raw_read(object_file, float, nx, ny, nz,
  offset=512)`` (same script, :51);
* 2-D real projections: ``raw_read(file, short, 512, 512, offset=100)
  + 2^15`` (Tomography_2D/Ytrtt2D_read_real_data.i:10).
Here the arrays come back in the C order of this package (last Yorick
dimension first: [views][nv][nu], [nz][ny][nx]) as float64, and
``to_device`` moves them to the GPU through one pinned staging buffer.
"""
from __future__ import annotations

import math

import numpy as np

# Yorick type names -> numpy dtypes (native byte order)
TYPES = {"char": np.uint8, "short": np.int16, "int": np.int32, "long": np.int64,
         "float": np.float32, "double": np.float64}


def raw_read(path, ytype: str, dims, offset: int = 0) -> np.ndarray:
    """Yorick's raw_read: shape reversed(dims) in C order (memory-mapped)."""
    dt = np.dtype(TYPES[ytype])
    count = int(math.prod(dims))
    arr = np.memmap(path, dtype=dt, mode="r", offset=int(offset), shape=(count,))
    return arr.reshape(tuple(reversed([int(d) for d in dims])))


def raw_write(path, array, offset: int = 0, ytype: str = None) -> None:
    """Inverse of raw_read: ``offset`` zero header bytes, then the values of
    ``array`` (C order, i.e. the last axis fastest = Yorick's d1)."""
    a = np.ascontiguousarray(array, dtype=TYPES[ytype] if ytype else None)
    with open(path, "wb") as f:
        f.write(b"\0" * int(offset))
        a.tofile(f)


def read_projections_3d(path, nu: int, nv: int, nstored: int = 720, each: int = 1,
                        ndata: int = None, offset: int = 512,
                        reference_orientation: bool = False) -> np.ndarray:
    """[ndata][nv][nu] float64: every ``each``-th of ``nstored`` FP32 views,
    the first ``ndata`` of them (Ytrtt3D_script_reconstructions.i:50).

    With ``reference_orientation`` every view is also reversed along u and
    transposed, as the script does before reconstructing
    (data(..,k) = transpose(data(::-1,,k)), :52-54): in this package's
    [nv][nu] order, out[k][j][i] = stored[k][i][nu - 1 - j].  The script's
    in-place assignment needs a square detector, and so does this option."""
    data = raw_read(path, "float", (nu, nv, nstored), offset)[::each]
    if ndata is not None:
        data = data[:ndata]
    out = np.asarray(data, dtype=np.float64)
    if reference_orientation:
        if nu != nv:
            raise ValueError(f"reference_orientation needs nu == nv (got {nu} x {nv})")
        out = np.ascontiguousarray(out[:, :, ::-1].transpose(0, 2, 1))
    return out


def read_volume(path, nx: int, ny: int, nz: int, offset: int = 512) -> np.ndarray:
    """[nz][ny][nx] float64 from FP32 values (Ytrtt3D_script_reconstructions.i:51)."""
    return np.asarray(raw_read(path, "float", (nx, ny, nz), offset), dtype=np.float64)


def read_projection_2d(path, n: int = 512, offset: int = 100, shift: float = 2.0 ** 15,
                       transpose: bool = False) -> np.ndarray:
    """One 16-bit real projection as float64 plus ``shift``
    (Ytrtt2D_read_real_data.i:10; Ytrtt2D_batch_real_data.i:31 adds
    2^16 - 1 and transposes)."""
    a = np.asarray(raw_read(path, "short", (n, n), offset), dtype=np.float64) + shift
    return a.T.copy() if transpose else a


def to_device(array, device="cuda", dtype=None):
    """Host array -> device tensor through one pinned staging buffer
    (asynchronous copy from page-locked memory)."""
    import torch

    src = torch.from_numpy(np.ascontiguousarray(array))
    if dtype is not None:
        src = src.to(dtype)
    staged = torch.empty(src.shape, dtype=src.dtype, pin_memory=True)
    staged.copy_(src)
    out = torch.empty(src.shape, dtype=src.dtype, device=device)
    out.copy_(staged, non_blocking=True)
    torch.cuda.current_stream(out.device).synchronize()
    return out
