"""The backprojector command line (main.cpp:29-187) on synthetic HDF5 NLOS datasets: dataset
parsing, bounce selection, data-order handling and resolution rescale on the CPU (--dry-run);
end-to-end reconstructions on the GPU against the oracle's pipeline on the same arrays."""
import json
import os
import subprocess

import numpy as np
import pytest

from tests import h5lite

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "backprojector_amd", "bin", "backprojector")

pytestmark = pytest.mark.skipif(h5lite.path() is None or not os.path.exists(CLI),
                                reason="HDF5 C library or the built CLI is missing")
f32 = np.float32


def wall(side):
    g = np.linspace(-1, 1, side, dtype=np.float64)
    X, Y = np.meshgrid(g, g, indexing="ij")
    return np.stack([X, Y, np.zeros_like(X)], -1).astype(f32)


def confocal_arrays(side=8, T=512, bounces=3, channels=2, seed=0):
    rng = np.random.default_rng(seed)
    # bounce b of every sample holds (b + 1) * base, channel c adds 1000 c: selection is visible
    base = rng.integers(0, 5, (side, side, 1, T, 1)).astype(f32)
    data = np.concatenate([base * (b + 1) for b in range(bounces)], axis=2)
    data = np.concatenate([data + 1000 * c for c in range(channels)], axis=4)
    return data, wall(side)


def run(args, check=True):
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120)
    if check and r.returncode != 0:
        raise AssertionError(r.stderr + r.stdout)
    return r


def dry(path, *extra):
    return json.loads(run([path, "--dry-run", *extra]).stdout)


def test_dry_run_row_major_confocal_bounces(tmp_path):
    data, grid = confocal_arrays()
    p = str(tmp_path / "scene.hdf5")
    h5lite.nlos_dataset(p, data, grid, grid, [0, -1, 0.5], [0, -1, 0.5], 0.1, 0.01, True,
                        [0, 0, 0.8], [1.0, 0.5, 1.0])
    d = dry(p)
    assert d["confocal"] and d["data_order"] == "row-major" and d["bounces"] == 1
    assert d["transient_shape"] == [8, 8, 512]
    # default third_bounce = 1: bounce 3 is index 1 of the bounce axis, channel 0
    assert d["transient_sum"] == pytest.approx(float(data[:, :, 1, :, 0].sum()))
    # equal resolutions scale with the volume's aspect, truncated (main.cpp:128-131)
    assert d["voxels"] == [32, 16, 32]
    assert d["volume_position"] == [0, 0, pytest.approx(0.8)] and d["t0"] == pytest.approx(0.1)
    assert d["output"] == str(tmp_path / "scene_recon.hdf5") and d["mode"] == "gpu"
    # -b 4 sums bounces 3 and 4 (indices 1 and 2)
    d4 = dry(p, "-b", "4", "-r", "16,8,4", "-c", "0.1,0.2,0.9", "-s2,2,2", "--phasor", "-w", "0.04")
    assert d4["bounces"] == 2
    assert d4["transient_sum"] == pytest.approx(float(data[:, :, 1:3, :, 0].sum()))
    assert d4["voxels"] == [16, 8, 4] and d4["mode"] == "phasor" and d4["wavelength"] == pytest.approx(0.04)
    assert d4["volume_position"] == [pytest.approx(0.1), pytest.approx(0.2), pytest.approx(0.9)]
    assert d4["volume_size"] == [2, 2, 2]


def test_dry_run_third_bounce_attribute(tmp_path):
    data, grid = confocal_arrays()
    p = str(tmp_path / "tb.hdf5")
    h5lite.nlos_dataset(p, data, grid, grid, [0, -1, 0.5], [0, -1, 0.5], 0.0, 0.01, True,
                        [0, 0, 0.8], [1, 1, 1], third_bounce=0)
    d = dry(p)   # third_bounce = 0: bounce 3 is index 0
    assert d["transient_sum"] == pytest.approx(float(data[:, :, 0, :, 0].sum()))


def test_dry_run_column_major_exhaustive(tmp_path):
    rng = np.random.default_rng(1)
    data = rng.integers(0, 7, (3, 4, 5, 2, 2, 64, 1)).astype(f32)   # [lx][ly][cx][cy][b][T][c]
    lg, cg = rng.normal(size=(3, 4, 3)).astype(f32), rng.normal(size=(5, 2, 3)).astype(f32)
    p = str(tmp_path / "ex.hdf5")
    h5lite.nlos_dataset(p, data, cg, lg, [0, -1, 0.5], [0.1, -1, 0.5], 0.0, 0.01, False,
                        [0, 0, 0.8], [1, 1, 1], order="column-major")
    d = dry(p)
    assert not d["confocal"] and d["data_order"] == "column-major"
    assert d["transient_shape"] == [3, 4, 5, 2, 64]
    assert d["camera_grid_shape"] == [5, 2, 3] and d["laser_grid_shape"] == [3, 4, 3]
    assert d["transient_sum"] == pytest.approx(float(data[..., 1, :, 0].sum()))


def test_errors(tmp_path):
    data, grid = confocal_arrays()
    p = str(tmp_path / "bad.hdf5")
    h5lite.write(p, {"data": data}, {"data order": "row-major"})
    r = run([p, "--dry-run"], check=False)
    assert r.returncode == 1 and "missing" in r.stderr
    r = run([str(tmp_path / "nope.hdf5")], check=False)
    assert r.returncode == 1 and "cannot open" in r.stderr
    r = run([p, "--frobnicate"], check=False)
    assert r.returncode == 1 and "unknown flag" in r.stderr
    p2 = str(tmp_path / "ok.hdf5")
    h5lite.nlos_dataset(p2, data, grid, grid, [0, -1, 0.5], [0, -1, 0.5], 0.0, 0.01, True, [0, 0, 0.8], [1, 1, 1])
    r = run([p2, "-b", "9", "--dry-run"], check=False)
    assert r.returncode == 1 and "bounce" in r.stderr


def rel_l2(a, b):
    a, b = np.asarray(a, np.complex128), np.asarray(b, np.complex128)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def oracle_volume(oracle, cap, data, N, confocal=True, use_cpu=False, wavelength=None):
    """The oracle's bp::backproject / bp::phasor_reconstruction (backproject.hpp:762-909,
    932-1039) on the row-major arrays the dataset holds: (reference-order volume, exactly
    accumulated fp32-formula volume or None)."""
    a = (data, cap.camera_grid, cap.laser_grid, cap.camera_position, cap.laser_position, cap.t0,
         cap.deltaT, confocal, cap.volume_position, cap.volume_size, N)
    if wavelength is not None:
        vol, _ = oracle.phasor_pipeline(*a, wavelength, use_cpu=use_cpu)
        return vol, None
    vol, it = oracle.backproject_pipeline(*a, use_cpu=use_cpu)
    if use_cpu:
        return vol, None
    exact = oracle.a1_volume_f64acc(it["H"], it["pairs"], N, it["zero"], it["inc"], float(it["t0"]),
                                    cap.deltaT)
    return vol, exact


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["gpu", "cpu", "phasor", "octree"])
def test_cli_end_to_end(tmp_path, gpu, oracle, mode):
    """main.cpp:112-187 end to end: the written volume against the oracle's pipeline on the same
    arrays.  The capture holds integer photon counts, so every GPU-formula voxel sum is exact
    and the volume must equal the exactly accumulated oracle volume bit for bit."""
    from backprojector_amd import scenes
    cap = scenes.render(scenes.config_scene(1, np.random.default_rng(4)), 16, 4096, seed=4)
    assert np.array_equal(cap.transient, np.round(cap.transient))
    data = cap.transient[:, :, None, :, None]        # one bounce (the third), one channel
    p = str(tmp_path / "cap.hdf5")
    h5lite.nlos_dataset(p, data, cap.camera_grid, cap.laser_grid, cap.camera_position,
                        cap.laser_position, cap.t0, cap.deltaT, True, cap.volume_position,
                        cap.volume_size, third_bounce=0)
    out = str(tmp_path / f"vol_{mode}.hdf5")
    flags = {"gpu": [], "cpu": ["--cpu"], "phasor": ["--phasor", "-w", "0.04"], "octree": ["--octree"]}[mode]
    r = run([p, "-r", "32,32,32", "-o", out] + flags)
    assert f"Backprojected to {out}" in r.stdout
    vol, props = h5lite.read_volume(out)
    assert vol.shape == (32, 32, 32)
    assert props["chunk"] == (32, 32, 32) and props["deflate"] == 4 and np.isnan(props["fill"])
    N = (32, 32, 32)
    if mode == "phasor":
        cref, _ = oracle_volume(oracle, cap, cap.transient, N, wavelength=0.04)
        # the CLI writes std::abs of each complex voxel (main.cpp:162-166)
        mag = np.abs(cref.astype(np.complex128))
        assert mag.max() > 0 and rel_l2(vol, mag) <= 1e-5
    elif mode == "octree":
        _, exact = oracle_volume(oracle, cap, cap.transient, N)
        computed = vol != 0
        assert np.array_equal(vol[computed], exact[computed].astype(f32))
        assert np.all(computed[exact >= 0.25 * exact.max()])
    elif mode == "cpu":
        ref, _ = oracle_volume(oracle, cap, cap.transient, N, use_cpu=True)
        assert np.abs(ref).max() > 0 and rel_l2(vol, ref) <= 1e-5
    else:
        ref, exact = oracle_volume(oracle, cap, cap.transient, N)
        assert np.abs(ref).max() > 0 and np.array_equal(vol, exact.astype(f32))
        assert rel_l2(vol, ref) <= 1e-5


@pytest.mark.gpu
def test_cli_bounce_sum_end_to_end(tmp_path, gpu, oracle):
    """-b 4 sums bounces 3 and 4 (dataset_loader.hpp:87-141): the volume equals the oracle's
    reconstruction of the hand-summed array."""
    from backprojector_amd import scenes
    cap = scenes.render(scenes.config_scene(1, np.random.default_rng(9)), 12, 2048, t0=2.0, seed=9)
    rng = np.random.default_rng(10)
    extra = rng.poisson(0.2, cap.transient.shape).astype(f32)   # a distinct fourth bounce
    data = np.stack([np.zeros_like(cap.transient), cap.transient, extra], axis=2)[..., None]
    p = str(tmp_path / "bounces.hdf5")
    h5lite.nlos_dataset(p, data, cap.camera_grid, cap.laser_grid, cap.camera_position,
                        cap.laser_position, cap.t0, cap.deltaT, True, cap.volume_position,
                        cap.volume_size)            # third_bounce = 1: bounce 3 is index 1
    out = str(tmp_path / "vol_b4.hdf5")
    run([p, "-b", "4", "-r", "16,16,16", "-o", out])
    vol, _ = h5lite.read_volume(out)
    summed = cap.transient + extra
    ref, exact = oracle_volume(oracle, cap, summed, (16, 16, 16))
    assert np.abs(ref).max() > 0 and np.array_equal(vol, exact.astype(f32))
    only3, _ = oracle_volume(oracle, cap, cap.transient, (16, 16, 16))
    assert not np.allclose(vol, only3)


@pytest.mark.gpu
@pytest.mark.parametrize("order", ["row-major", "column-major"])
def test_cli_exhaustive_end_to_end(tmp_path, gpu, oracle, order):
    """A non-confocal dataset (5x5 lasers x 6x6 SPADs) through the command line, stored row- or
    column-major: the volume equals the oracle's pipeline on the row-major arrays."""
    from backprojector_amd import scenes
    cap = scenes.render(scenes.config_scene(2), 6, 2048, t0=2.0, confocal=False, laser_side=5, seed=21)
    data = cap.transient[:, :, :, :, None, :, None]          # [lx][ly][cx][cy][bounce][T][channel]
    p = str(tmp_path / f"ex_{order}.hdf5")
    h5lite.nlos_dataset(p, data, cap.camera_grid, cap.laser_grid, cap.camera_position, cap.laser_position,
                        cap.t0, cap.deltaT, False, cap.volume_position, cap.volume_size, order=order,
                        third_bounce=0)
    out = str(tmp_path / "vol.hdf5")
    r = run([p, "-r", "24,24,24", "-o", out])
    assert f"Backprojected to {out}" in r.stdout
    vol, _ = h5lite.read_volume(out)
    ref, exact = oracle_volume(oracle, cap, cap.transient, (24, 24, 24), confocal=False)
    assert np.abs(ref).max() > 0 and np.array_equal(vol, exact.astype(f32))
    assert rel_l2(vol, ref) <= 1e-5
