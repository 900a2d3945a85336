"""The C-ABI library loads and exports every symbol include/physicell_hip.h
declares (no GPU work is done here)."""
import ctypes

from physicell_amd import hip as H


def test_header_declares_entry_points():
    names = H.declared_functions()
    for must in ("hc_create", "hc_diffusion_sweep", "hc_secretion", "hc_forces_only_pipeline",
                 "hc_integrate", "hc_wait", "pc_sim_run"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = H.load()
    missing = [n for n in H.declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_sizes_are_consistent():
    lib = H.load()
    M = 1000
    assert lib.hc_cells_bytes(M) == M * (23 * 8 + 5 * 4)
    assert lib.hc_cells_field_offset(0, M) == 0
    assert lib.hc_cells_field_offset(22, M) == 22 * 8 * M
    assert lib.hc_cells_field_offset(23, M) == 23 * 8 * M
    assert lib.hc_cells_field_offset(27, M) == 23 * 8 * M + 4 * 4 * M
    assert lib.hc_rates_bytes(M, 3) == 4 * M * 3 * 8
    g = H.Grid(10, 20, 30, 2, 0, 0, 1, 1, 1, 0.01, 0, 0, 0)
    assert lib.hc_thomas_table_bytes(ctypes.byref(g)) >= 2 * 3 * 2 * 30 * 8
    assert lib.hc_hash_bytes(100, M) > 64 * M


class _NoDeviceCtx:
    """Stands in for a Context: any call that reaches the library fails the test."""
    h = None

    def _check(self, rc, what):
        raise AssertionError(f"{what} reached the library")

    def wait(self):
        raise AssertionError("wait reached the library")


def _buffer_without_device(nbytes):
    b = object.__new__(H.Buffer)
    b.ctx, b.nbytes, b.h = _NoDeviceCtx(), nbytes, None
    return b


def test_buffer_rejects_out_of_range_copies_before_the_host_mirror():
    """Buffer.upload/download check offset + size against the buffer before any
    memmove into or out of the pinned mirror (which is exactly nbytes long)."""
    import numpy as np
    import pytest
    b = _buffer_without_device(64)
    for arr, off in ((np.zeros(9), 0), (np.zeros(8), 8), (np.zeros(1), 64), (np.zeros(1), -8)):
        with pytest.raises(H.HipError, match="outside the 64-byte buffer"):
            b.upload(arr, offset=off)
    for dtype, count, off in ((np.float64, 9, 0), (np.float64, 8, 8), (np.uint8, 8, 57), (np.float64, 0, 65),
                              (np.float64, 1, -1)):
        with pytest.raises(H.HipError, match="outside the 64-byte buffer"):
            b.download(dtype, count, offset=off)
