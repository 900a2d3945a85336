"""pc_sim_write_snapshot: the .mat files hold the device state (positions,
velocities, pressure, densities) exactly, in the reference's row schema."""
import numpy as np
import pytest

from physicell_amd import hip as H
from physicell_amd import matio
from physicell_amd import scenarios as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("builder", ["config0", "small3d"])
def test_sim_snapshot_holds_state(tmp_path, builder):
    spec, cells = S.BUILDERS[builder]()
    sim = H.Simulation(spec, cells)
    sim.run_periods(3)
    sim.wait()
    sim.write_snapshot(tmp_path, 12)
    name, m = matio.read_mat(str(tmp_path / "output00000012_cells.mat"))
    assert name == "cells" and m.shape[1] == cells.n
    for row, f in [(matio.CELL_ROW_X, H.F_POS_X), (matio.CELL_ROW_Y, H.F_POS_Y), (matio.CELL_ROW_Z, H.F_POS_Z),
                   (matio.CELL_ROW_VX, H.F_VEL_X), (matio.CELL_ROW_PRESSURE, H.F_PRESSURE),
                   (matio.CELL_ROW_RADIUS, H.F_RADIUS)]:
        assert np.array_equal(m[row], sim.field(f)), row
    assert np.array_equal(m[matio.CELL_ROW_ID], np.arange(cells.n))
    name, me = matio.read_mat(str(tmp_path / "output00000012_microenvironment0.mat"))
    assert name == "multiscale_microenvironment"
    assert me.shape == (4 + len(spec.substrates), spec.n_voxels)
    assert np.array_equal(me[matio.ME_FIRST_SUBSTRATE:].reshape(-1), sim.density())
    # voxel centres (FP32 as the reference computes them)
    x0 = np.float32(spec.x_min) + (np.arange(spec.nx, dtype=np.float32) + np.float32(0.5)) * np.float32(spec.dx)
    assert np.array_equal(me[matio.ME_ROW_X][:spec.nx], x0.astype(np.float64))
    sim.close()
