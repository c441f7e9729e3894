"""Checkpoint files in the reference's binary layout.

A BHaH checkpoint (``write_checkpoint``, BHaH/tests/write_checkpoint_write_checkpoint__openmp.c:84-117)
is the raw ``commondata_struct``, then for each grid its raw ``params_struct`` followed by
the payload ``int count; int indices[count]; double data[count * NUM_EVOL_GFS]``
(point-major).  The two structs are generated per project, so they are carried here as
opaque byte strings of the project's ``sizeof``; the payload is produced and consumed on
the device by ``bhah_hip_checkpoint_pack`` / ``_unpack``.
"""

from __future__ import annotations

from typing import List, Sequence, Tuple


def write_checkpoint(path: str, commondata: bytes, grids: Sequence[Tuple[bytes, object, object]]) -> None:
    """grids: (params_struct bytes, Patch, y_n tensor) per grid."""
    with open(path, "wb") as f:
        f.write(commondata)
        for params, patch, y_n in grids:
            f.write(params)
            f.write(patch.checkpoint_payload(y_n))


def read_checkpoint(path: str, commondata_size: int, params_size: int, grids: Sequence[Tuple[object, object]]
                    ) -> Tuple[bytes, List[bytes]]:
    """grids: (Patch, y_n tensor) per grid, filled in place.  Returns the raw commondata and
    params_struct bytes.  Truncated files, bad counts or indices raise (the reference exits)."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < commondata_size:
        raise ValueError("checkpoint truncated in commondata_struct")
    commondata = data[:commondata_size]
    pos = commondata_size
    params = []
    for patch, y_n in grids:
        params.append(data[pos:pos + params_size])
        pos += params_size
        pos += patch.restore_payload(data[pos:], y_n)
    return commondata, params
