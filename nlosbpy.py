"""`import nlosbpy` — the reference's Python module name (nlosbpy.cpp:81,
`PYBIND11_MODULE(nlosbpy, m)`), so a caller of the reference module runs unchanged:

    import nlosbpy
    vol = nlosbpy.backproject(transient_data, camera_grid_positions, laser_grid_positions,
                              camera_position, laser_position, t0, deltaT, is_confocal,
                              volume_position, volume_size, voxels_per_side, wall_normal,
                              assume_row_major=False, use_cpu=False)      # nlosbpy.cpp:89-103
    cvol = nlosbpy.phasor(..., voxels_per_side, wavelength,
                          assume_row_major=False, use_cpu=False)          # nlosbpy.cpp:105-119

The functions are those of `backprojector_amd.nlosbpy` (the same objects); every computation
runs on the GPU through libbpx.so.  `phasor_multi` and `prepare` are extensions with no
counterpart in the reference module.
"""
from backprojector_amd.nlosbpy import backproject, phasor, phasor_multi, prepare

__all__ = ["backproject", "phasor", "phasor_multi", "prepare"]
