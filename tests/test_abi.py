"""CPU checks of the C ABI: the library loads, exports every symbol include/bpx.h declares,
rejects invalid arguments with status codes (no exit, no GPU needed), and the host renderer
produces known answers."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "bpx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bpx_[a-z_0-9]+)\s*\(", src)))


def test_every_declared_symbol_is_exported(bpx):
    from backprojector_amd import _native
    L = _native.lib()
    syms = header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), f"{s} declared in include/bpx.h but not exported"
    assert sorted(_native.EXPORTED) == syms


def test_pair_layout_matches_reference_struct(bpx):
    assert bpx.PAIR_DTYPE.itemsize == 32
    assert [bpx.PAIR_DTYPE.fields[n][1] for n in bpx.PAIR_DTYPE.names] == [0, 12, 24, 28]


def test_version_and_abi(bpx):
    from backprojector_amd import _native
    assert _native.lib().bpx_abi_version() == 1
    assert b"gfx950" in _native.lib().bpx_version()


def _tiny():
    from backprojector_amd import scenes
    cg = scenes.wall_grid(4)
    pr = scenes.wall_pairs(cg, cg, True)
    H = np.ones((16, 64), np.float32)
    return H, pr


@pytest.mark.parametrize("kw,msg", [
    (dict(deltaT=0.0), "deltaT"),
    (dict(deltaT=float("nan")), "deltaT"),
    (dict(t0=float("inf")), "t0"),
    (dict(N=(0, 4, 4)), "voxels_per_side"),
    (dict(zero=(np.nan, 0, 0)), "volume_zero_pos"),
])
def test_invalid_arguments_return_status(bpx, kw, msg):
    H, pr = _tiny()
    args = dict(deltaT=0.001, t0=0.0, N=(4, 4, 4), zero=(0, 0, 0))
    args.update(kw)
    with pytest.raises(bpx.InvalidArgument, match=msg):
        bpx.backprojection(H, pr, args["N"], args["zero"], np.eye(3).ravel() * 0.1, args["t0"],
                           args["deltaT"])


def test_size_mismatch_is_invalid(bpx):
    from backprojector_amd import _native
    H, pr = _tiny()
    N = np.array([4, 4, 4], np.uint32)
    out = np.zeros((4, 4, 4), np.float32)
    z = np.zeros(3, np.float32)
    inc = np.eye(3, dtype=np.float32).ravel()
    st = _native.lib().bpx_backprojection(_native.ptr(H), H.size - 1, 64, _native.ptr(pr), 16,
                                          _native.ptr(out), _native.ptr(N), _native.ptr(z),
                                          _native.ptr(inc), 0.0, 0.001)
    assert st == 1 and b"transient_size" in _native.lib().bpx_last_error()


def test_plan_queries_reject_null(bpx):
    from backprojector_amd import _native
    L = _native.lib()
    buf = ctypes.create_string_buffer(64)
    assert L.bpx_plan_kernel_name(None, buf, len(buf)) == 1
    assert b"NULL" in L.bpx_last_error()
    assert L.bpx_plan_synchronize(None) == 1 and L.bpx_plan_execute(None, None) == 1


def test_nlosbpy_signature_matches_reference(bpx):
    import inspect
    sig = inspect.signature(bpx.nlosbpy.backproject)
    assert list(sig.parameters) == [
        "transient_data", "camera_grid_positions", "laser_grid_positions", "camera_position",
        "laser_position", "t0", "deltaT", "is_confocal", "volume_position", "volume_size",
        "voxels_per_side", "wall_normal", "assume_row_major", "use_cpu"]
    assert sig.parameters["assume_row_major"].default is False
    assert sig.parameters["use_cpu"].default is False
    sig = inspect.signature(bpx.nlosbpy.phasor)
    assert list(sig.parameters)[11] == "wavelength"


def test_top_level_nlosbpy_module_is_the_reference_name(bpx):
    """`import nlosbpy` works as with the reference's pybind11 module (nlosbpy.cpp:81) and
    exposes the same function objects, argument names and defaults (nlosbpy.cpp:89-119)."""
    import inspect
    import nlosbpy
    assert nlosbpy.backproject is bpx.nlosbpy.backproject
    assert nlosbpy.phasor is bpx.nlosbpy.phasor
    assert {"backproject", "phasor"} <= set(nlosbpy.__all__)
    sig = inspect.signature(nlosbpy.phasor)
    positional = [n for n, p in sig.parameters.items() if p.kind is p.POSITIONAL_OR_KEYWORD]
    assert positional == [
        "transient_data", "camera_grid_positions", "laser_grid_positions", "camera_position",
        "laser_position", "t0", "deltaT", "is_confocal", "volume_position", "volume_size",
        "voxels_per_side", "wavelength", "assume_row_major", "use_cpu"]
    assert sig.parameters["assume_row_major"].default is False
    assert sig.parameters["use_cpu"].default is False
    # the one addition is keyword-only with the reference's behaviour as its default
    extra = {n: p for n, p in sig.parameters.items() if p.kind is not p.POSITIONAL_OR_KEYWORD}
    assert list(extra) == ["output"] and extra["output"].kind is extra["output"].KEYWORD_ONLY
    assert extra["output"].default == "complex"


def test_nlosbpy_shape_errors(bpx):
    H = np.zeros((4, 4, 32), np.float32)
    g = np.zeros((4, 5, 3), np.float32)
    with pytest.raises(bpx.InvalidArgument):
        bpx.nlosbpy.backproject(H, g, g, np.zeros(3), np.zeros(3), 0.0, 0.001, True,
                                np.zeros(3), np.ones(3), np.array([4, 4, 4]), np.zeros(3),
                                assume_row_major=True)


# ----------------------------------------------------------------- renderer
def test_renderer_single_point_bin(bpx):
    # isotropic point straight above the confocal wall point: one bin per pair at the path
    # length floor((2 lw + 2 |w - q|) / dT); value = peak for the brightest pair.
    from backprojector_amd import scenes
    cg = scenes.wall_grid(3)
    pr = scenes.wall_pairs(cg, cg, True)
    q = np.array([[0.0, 0.0, 0.75, 0, 0, 0, 1.0]], np.float32)
    H = bpx.render_transient(pr, 8192, 0.0, 0.001, q, peak_count=100.0, poisson=False)
    for p in range(len(pr)):
        w = pr[p]["camera_point"].astype(np.float64)
        L = 2 * float(pr[p]["camera_wall"]) + 2 * np.linalg.norm(w - [0, 0, 0.75])
        b = int(np.floor(L / np.float64(np.float32(0.001))))
        assert H[p].nonzero()[0].tolist() == [b]
    assert H.max() == pytest.approx(100.0)
    assert H[4].argmax() == H[4].nonzero()[0][0]      # centre pair is the closest
    assert H[4].max() == H.max()


def test_renderer_poisson_is_seeded_and_unbiased(bpx):
    from backprojector_amd import scenes
    rng = np.random.default_rng(0)
    hidden = scenes.points(rng, 4)
    a = scenes.render(hidden, 8, 4096, seed=5).transient
    b = scenes.render(hidden, 8, 4096, seed=5).transient
    c = scenes.render(hidden, 8, 4096, seed=6).transient
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    mean = scenes.render(hidden, 8, 4096, poisson=False).transient
    assert np.all(a == np.round(a))
    m = mean > 50
    z = (a[m] - mean[m]) / np.sqrt(mean[m])
    assert abs(z.mean()) < 0.5 and 0.6 < z.std() < 1.4


def test_pointer_arguments_keep_temporaries_alive():
    """ptr(_f32(x)) passes a temporary: the ctypes argument must own a reference to it, since
    the library runs without the GIL and a freed buffer could be reused during the call."""
    import gc
    import weakref
    from backprojector_amd import _native as nat
    a = np.arange(8, dtype=np.uint32)
    r = weakref.ref(a)
    p = nat.ptr(a)
    del a
    gc.collect()
    assert r() is not None
    del p
    gc.collect()
    assert r() is None
