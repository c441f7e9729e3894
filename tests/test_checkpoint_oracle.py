"""The oracle's checkpoint payload against the layout the reference's write_checkpoint
produces (a hand-built known answer), and the reader's validation."""

import struct

import numpy as np


def test_payload_layout_known_answer(oracle_lib):
    L = oracle_lib.lib()
    ntot, ngf = 3, 2
    y = np.array([1.0, 2.0, 3.0, 10.0, 20.0, 30.0])  # gf-major: gf0 = 1,2,3; gf1 = 10,20,30
    out = np.zeros(4 * (1 + ntot) + 8 * ntot * ngf, np.uint8)
    n = L.og_checkpoint_pack(ntot, ngf, y.ctypes.data, out.ctypes.data)
    assert n == out.size
    b = out.tobytes()
    assert struct.unpack("<4i", b[:16]) == (3, 0, 1, 2)
    assert struct.unpack("<6d", b[16:]) == (1.0, 10.0, 2.0, 20.0, 3.0, 30.0)
    back = np.zeros_like(y)
    assert L.og_checkpoint_unpack(ntot, ngf, out.ctypes.data, back.ctypes.data) == 0
    assert np.array_equal(back, y)


def test_reader_rejects_bad_indices(oracle_lib):
    L = oracle_lib.lib()
    bad = np.frombuffer(struct.pack("<3i", 2, 1, 1) + struct.pack("<2d", 0.0, 0.0), np.uint8).copy()
    y = np.zeros(2)
    assert L.og_checkpoint_unpack(2, 1, bad.ctypes.data, y.ctypes.data) == -1
    bad = np.frombuffer(struct.pack("<2i", 5, 0), np.uint8).copy()
    assert L.og_checkpoint_unpack(2, 1, bad.ctypes.data, y.ctypes.data) == -1


def test_partial_payload_zeroes_unlisted_points(oracle_lib):
    """A payload listing only some points: the reader zeroes all of y_n first, so unlisted
    points come back as 0 whatever y_n held (read_checkpoint__openmp.c:163-171)."""
    L = oracle_lib.lib()
    ntot, ngf = 5, 2
    payload = struct.pack("<3i", 2, 1, 3) + struct.pack("<4d", 1.5, 2.5, 3.5, 4.5)
    buf = np.frombuffer(payload, np.uint8).copy()
    y = np.full(ntot * ngf, 7.0)
    assert L.og_checkpoint_unpack(ntot, ngf, buf.ctypes.data, y.ctypes.data) == 0
    assert y.tolist() == [0.0, 1.5, 0.0, 3.5, 0.0, 0.0, 2.5, 0.0, 4.5, 0.0]
