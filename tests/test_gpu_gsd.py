"""HOOMD GSD frames of a device state (hmc_gsd_append_frame:
HoomdGsdFile::append_frame with configuration/box, configuration/dimensions
and particles/position, hoomd-gsd/src/hoomd.rs:343-1013): the frames read
back hold the sites in tag order rounded to FP32 (Rust's `as f32`), the box
of Hypercuboid::to_gsd_box and particles/N, for spheres, disks (z = 0) and
cubes (with orientations)."""
import sys

import numpy as np
import pytest

import _pkg

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hmc():
    p = _pkg.pkg()
    p.lib()
    return p


@pytest.fixture(scope="module")
def wl():
    return _pkg.workloads()


@pytest.mark.parametrize("dim", [2, 3])
def test_frames_of_a_running_system(hmc, wl, tmp_path, dim):
    if dim == 3:
        L, pos = wl.fcc_spheres((8, 8, 8), packing_fraction=0.45)
    else:
        L, pos = wl.square_lattice_disks(n_side=40, packing_fraction=0.6)
    log(f"GSD frames: D={dim} N={len(pos)}")
    sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), precision="mixed")
    path = tmp_path / "traj.gsd"
    f = hmc.GsdFile.create(path)
    sites = []
    for step in range(4):
        sim.apply(step, 0, 11, 1.0, 0.1)
        f.append_frame(sim, step * 10)
        sites.append(sim.positions()[1])
    f.close()
    g = hmc.GsdFile.open(path)
    assert g.n_frames == 4
    assert g.header() == ("hoomd-rs_amd", "hoomd", (1, 4))
    for k in range(4):
        assert int(g.read(k, "configuration/step")[0, 0]) == k * 10
        assert int(g.read(k, "configuration/dimensions")[0, 0]) == dim
        box = g.read(k, "configuration/box").ravel()
        want = np.array([L[0], L[1], L[2] if dim == 3 else 0.0, 0.0, 0.0, 0.0], np.float32)
        assert np.array_equal(box, want)
        assert int(g.read(k, "particles/N")[0, 0]) == len(pos)
        p = g.read(k, "particles/position")
        want = np.zeros((len(pos), 3), np.float32)
        want[:, :dim] = sites[k].astype(np.float32)
        assert p.dtype == np.float32 and np.array_equal(p, want)
    g.close()
    sim.close()


def test_frames_of_cubes_carry_orientations(hmc, wl, tmp_path):
    L, pos = wl.simple_cubic(n_side=6, packing_fraction=0.3)
    q0 = np.tile([1.0, 0.0, 0.0, 0.0], (len(pos), 1))
    sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), model="convex", vertices=wl.UNIT_CUBE, orientations=q0)
    sim.apply_rotate(0, 0, 5, 1.0, 0.2)
    f = hmc.GsdFile.create(tmp_path / "cubes.gsd")
    f.append_frame(sim, 3)
    f.close()
    g = hmc.GsdFile.open(tmp_path / "cubes.gsd")
    _bq, sq = sim.orientations()
    assert np.array_equal(g.read(0, "particles/orientation"), sq.astype(np.float32))
    assert np.array_equal(g.read(0, "particles/position"), sim.positions()[1].astype(np.float32))
    g.close()
    sim.close()
