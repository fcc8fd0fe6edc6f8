"""Writes reference_fixtures.json: known-answer vectors transcribed from the
reference's own test suite (inputs and expected outputs only), each tagged
with the test file:line it comes from. Run: python tests/golden/make_fixtures.py
"""

import json
import os

FIX = {
    # CUDA/test/cpu/gpuNUFFT_kernel_tests.cpp:8-26 — loadGrid3Kernel with the
    # defaults osr=1, kw=3. The interior checks are one-sided EXPECT_LT.
    "lut_osf1_kw3": {
        "src": "CUDA/test/cpu/gpuNUFFT_kernel_tests.cpp:11-23",
        "osf": 1.0, "kw": 3, "K": 1365,
        "exact": {"0": 1.0, "1364": 0.0},
        "one_sided_lt": {"1": 0.9940, "401": 0.0621, "665": 0.0041}, "eps": 1e-4,
    },
    # CUDA/test/cpu/gpuNUFFT_precomputation_tests.cpp:39-62
    "sector_count_per_dim": {
        "src": "CUDA/test/cpu/gpuNUFFT_precomputation_tests.cpp:39-62",
        "cases": [[128, 8, 16], [124, 8, 16], [7, 8, 1], [120, 8, 15]],
    },
    # CUDA/test/cpu/gpuNUFFT_operator_factory_tests.cpp:18-78
    "factory_3d": {
        "src": "CUDA/test/cpu/gpuNUFFT_operator_factory_tests.cpp:18-78",
        "img": [16, 16, 16], "osf": 1.5, "sw": 8, "kw": 3,
        "coords_soa": [[-0.5, -0.3, -0.1, 0.1, 0.3, 0.5],
                       [-0.5, -0.5, 0.0, 0.0, 0.5, 0.45],
                       [-0.33, -0.16666, 0.0, 0.0, -0.23, 0.45]],
        "grid_width": 24,
        "sector_count": 27,
        # the test compares only the first sectorDataCount.count() = 28 entries
        "sector_data_count": [0, 1, 1, 1, 1, 1, 1, 1, 1, 2, 3, 3, 3, 3, 5, 5, 5, 5, 5, 5, 5, 5,
                              5, 5, 5, 5, 5, 6],
        "center0": [4, 4, 4],
        "data_indices": [0, 4, 1, 2, 3, 5],
        "sorted_coords_checks": {"0": -0.5, "6": -0.5, "12": -0.33}, "eps": 1e-4,
    },
    # CUDA/test/cpu/gpuNUFFT_operator_factory_tests.cpp:80-109
    "factory_invalid_channels": {
        "src": "CUDA/test/cpu/gpuNUFFT_operator_factory_tests.cpp:80-109",
        "traj_channels": 3, "expect": "invalid_argument",
    },
    # CUDA/test/cpu/gpuNUFFT_operator_factory_tests.cpp:111-167
    "factory_2d": {
        "src": "CUDA/test/cpu/gpuNUFFT_operator_factory_tests.cpp:111-167",
        "img": [16, 16], "osf": 1.5, "sw": 8, "kw": 3,
        "coords_soa": [[-0.5, -0.3, -0.1, 0.1, 0.3, 0.5],
                       [-0.5, -0.5, 0.0, 0.0, 0.5, 0.45]],
        "grid_count": 576, "sector_count": 9,
        "sector_data_count": [0, 2, 2, 2, 2, 4, 4, 4, 4, 6],
        "center0": [4, 4],
    },
    # CUDA/test/cpu/gpuNUFFT_tests.cpp — gpuNUFFT_cpu with hand-built tables.
    # coords interleaved (x,y,z); data interleaved (re,im); osr 1; grid width
    # 10; expected values are real parts at (x,y,z); tolerance 1e-3 at the
    # centre, 1e-2 elsewhere.
    "cpu_gridding": [
        {"src": "CUDA/test/cpu/gpuNUFFT_tests.cpp:10-87 (CPUTest_1SectorKernel3)",
         "kw": 3, "sw": 8, "width": 10, "data": [[1, 1]], "coords": [[0, 0, 0]],
         "sectors": [0, 1], "centers": [[5, 5, 5]],
         "expect": [[5, 5, 5, 1.0, 1e-3], [5, 4, 5, 0.4502, 1e-2], [4, 5, 5, 0.4502, 1e-2],
                    [5, 6, 5, 0.4502, 1e-2], [6, 6, 5, 0.2027, 1e-2], [4, 4, 5, 0.2027, 1e-2],
                    [4, 6, 5, 0.2027, 1e-2]]},
        {"src": "CUDA/test/cpu/gpuNUFFT_tests.cpp:89-170 (CPUTest_1SectorKernel5)",
         "kw": 5, "sw": 10, "width": 10, "data": [[1, 1]], "coords": [[0, 0, 0]],
         "sectors": [0, 1], "centers": [[5, 5, 5]],
         "expect": [[5, 5, 5, 1.0, 1e-3], [3, 3, 5, 0.0049, 1e-2], [4, 4, 5, 0.3218, 1e-2],
                    [5, 4, 5, 0.5673, 1e-2], [5, 7, 5, 0.0697, 1e-2], [5, 3, 5, 0.0697, 1e-2]]},
        {"src": "CUDA/test/cpu/gpuNUFFT_tests.cpp:174-264 (CPUTest_1SectorKernel3nData)",
         "kw": 3, "sw": 10, "width": 10,
         "data": [[1, 1], [1, 1], [0.5, 0.5], [1, 1]],
         "coords": [[0, 0, 0], [0.3, 0.3, 0], [-0.3, -0.3, 0], [0.5, 0, 0]],
         "sectors": [0, 4], "centers": [[5, 5, 5]],
         "expect": [[5, 5, 5, 1.0, 1e-3], [3, 3, 5, 0.1013, 1e-2], [1, 2, 5, 0.2251, 1e-2],
                    [6, 5, 5, 0.4502, 1e-2], [8, 8, 5, 1.0, 1e-2], [9, 9, 5, 0.2027, 1e-2]]},
        {"src": "CUDA/test/cpu/gpuNUFFT_tests.cpp:266-366 (CPUTest_2SectorsKernel3nData)",
         "kw": 3, "sw": 5, "width": 10,
         "data": [[0.5, 0.5], [0.7, 1], [1, 1], [1, 1], [1, 1]],
         "coords": [[-0.3, 0.2, 0], [-0.1, 0, 0], [0, 0, 0], [0.5, 0, 0], [0.3, 0.3, 0]],
         "sectors": [0, 2, 5], "centers": [[2, 7, 5], [7, 7, 5]],
         "expect": [[5, 5, 5, 1.3152, 1e-3], [3, 6, 5, 0.2432, 1e-2], [1, 7, 5, 0.2251, 1e-2],
                    [6, 5, 5, 0.4502, 1e-2], [8, 8, 5, 1.0, 1e-2], [9, 9, 5, 0.2027, 1e-2]]},
    ],
    # CUDA/test/gpu/gpuNUFFT_fft_tests.cpp:49-140 — shift permutations of
    # 0..N-1 on a (x fastest) grid, offsets floor/ceil(G/2) (all even here).
    "fftshift": [
        {"src": "CUDA/test/gpu/gpuNUFFT_fft_tests.cpp:49-72", "dims": [2, 2, 2],
         "expect": [7, 6, 5, 4, 3, 2, 1, 0]},
        {"src": "CUDA/test/gpu/gpuNUFFT_fft_tests.cpp:98-140", "dims": [4, 4, 2],
         "expect": [26, 27, 24, 25, 30, 31, 28, 29, 18, 19, 16, 17, 22, 23, 20, 21, 10, 11, 8,
                    9, 14, 15, 12, 13, 2, 3, 0, 1, 6, 7, 4, 5]},
    ],
    # CUDA/test/gpu/forward/gpuNUFFT_forward_tests.cpp:169-266 — full forward
    # of a constant (1+1i) 32x32 image, 4 coils, kw 5, osf 1.25, sw 8, no
    # dens/sens. Trajectory array is SoA: x = coords[0..2], y = coords[3..5].
    # Expected real part of sample 0, every coil; tolerance 0.02.
    "forward_2d_32_4": {
        "src": "CUDA/test/gpu/forward/gpuNUFFT_forward_tests.cpp:169-266",
        "img": [32, 32], "osf": 1.25, "kw": 5, "sw": 8, "coils": 4,
        "coords_flat": [-0.31719, -0.38650, 0.0, 0.0, 0.31719, 0.38650],
        "expect_nearest": 0.73328, "expect_texture": 0.816, "eps": 0.02,
    },
    # CUDA/test/gpu/forward/gpuNUFFT_forward_tests.cpp:268-390 — sens = (c+1)(1+1i):
    # coil c output = (c+1) x coil 0 output; tolerance 1e-3.
    "forward_2d_32_8_sens": {
        "src": "CUDA/test/gpu/forward/gpuNUFFT_forward_tests.cpp:268-390",
        "img": [32, 32], "osf": 1.25, "kw": 5, "sw": 8, "coils": 8,
        "coords_flat": [-0.31719, -0.38650, 0.0, 0.0, 0.25, 0.25], "eps": 1e-3,
    },
    # CUDA/test/gpu/gpuNUFFT_deapo_tests.cpp:342-438 — adjoint of one sample
    # (1+0.5i) at k=0, 30 coils, sens 1+0.2i, dens 1, 32x32 osf 2 kw 3 sw 8.
    "deapo_2d_30_b1": {
        "src": "CUDA/test/gpu/gpuNUFFT_deapo_tests.cpp:342-438",
        "img": [32, 32], "osf": 2.0, "kw": 3, "sw": 8, "coils": 30,
        "data": [1.0, 0.5], "sens": [1.0, 0.2], "pixel": [0, 0], "expect": 1.03125,
        "eps": 0.01,
    },
    # CUDA/test/gpu/gpuNUFFT_deapo_tests.cpp:161-253 — 3-D 32^3 osf 2, one
    # sample (1+0i) at k=0, 2 coils, nearest lookup: value at (0,0,16),
    # both coils equal.
    "deapo_3d_32_2coils": {
        "src": "CUDA/test/gpu/gpuNUFFT_deapo_tests.cpp:161-253",
        "img": [32, 32, 32], "osf": 2.0, "kw": 3, "sw": 8, "coils": 2,
        "pixel": [0, 0, 16], "expect": 0.005513, "eps": 0.01,
    },
    # CUDA/test/gpu/gpuNUFFT_conv_tests.cpp and gpuNUFFT_2D_conv_tests.cpp —
    # adjoint CONVOLUTION stop (the raw oversampled grid) of a few samples,
    # nearest (const) lookup, one coil. "traj" is the raw trajectory array as
    # the test writes it, read by the factory as SoA (row a = axis a).
    # Checks: [x, y, z, component (0 re / 1 im), expected, tolerance the
    # reference asserts]; grid index x + Gx*(y + Gy*z). The expected values
    # are printed to 4 decimals.
    "gpu_conv": [
        {"src": "CUDA/test/gpu/gpuNUFFT_conv_tests.cpp:108-188", "img": [10, 10, 10], "osf": 1.0,
         "kw": 3, "sw": 5, "data": [[1, 1]], "traj": [[0], [0], [0]],
         "checks": [[5, 5, 5, 0, 1.0, 0.02], [5, 4, 5, 0, 0.4502, 0.2], [4, 5, 5, 0, 0.4502, 0.2],
                    [5, 6, 5, 0, 0.4502, 0.2], [6, 6, 5, 0, 0.2027, 0.2], [4, 4, 5, 0, 0.2027, 0.2],
                    [4, 6, 5, 0, 0.2027, 0.2]]},
        # sector width 9 (not a divisor of the grid), a sample whose taps wrap to x = 0
        {"src": "CUDA/test/gpu/gpuNUFFT_conv_tests.cpp:190-284", "img": [10, 10, 10], "osf": 1.0,
         "kw": 3, "sw": 9, "data": [[1, 1], [1, 1]], "traj": [[0, 0.4], [0, 0], [0, 0]],
         "checks": [[5, 5, 5, 0, 1.0, 0.02], [5, 4, 5, 0, 0.4502, 0.2], [4, 5, 5, 0, 0.4502, 0.2],
                    [5, 6, 5, 0, 0.4502, 0.2], [6, 6, 5, 0, 0.2027, 0.2], [4, 4, 5, 0, 0.2027, 0.2],
                    [4, 6, 5, 0, 0.2027, 0.2], [9, 4, 5, 0, 0.4502, 0.2], [8, 5, 5, 0, 0.4502, 0.2],
                    [9, 6, 5, 0, 0.4502, 0.2], [8, 6, 5, 0, 0.2027, 0.2], [8, 4, 5, 0, 0.2027, 0.2],
                    [0, 6, 5, 0, 0.2027, 0.2]]},
        {"src": "CUDA/test/gpu/gpuNUFFT_conv_tests.cpp:286-372", "img": [10, 10, 10], "osf": 1.0,
         "kw": 5, "sw": 5, "data": [[1, 1]], "traj": [[0], [0], [0]],
         "checks": [[5, 5, 5, 0, 1.0, 0.02], [3, 3, 5, 0, 0.0049, 0.2], [4, 4, 5, 0, 0.3218, 0.2],
                    [5, 4, 5, 0, 0.5673, 0.2], [5, 7, 5, 0, 0.0697, 0.2], [5, 3, 5, 0, 0.0697, 0.2]]},
        {"src": "CUDA/test/gpu/gpuNUFFT_conv_tests.cpp:374-501 (and 8SectorsKernel3nData :503-609)",
         "img": [10, 10, 10], "osf": 1.0, "kw": 3, "sw": 5, "data": [[0.5, 0.5], [0.7, 1], [1, 1], [1, 1], [1, 1]], "traj": [[-0.3, -0.1, 0, 0.5, 0.3], [0.2, 0, 0, 0, 0.3], [0, 0, 0, 0, 0]],
         "checks": [[5, 5, 5, 0, 1.3152, 0.02], [3, 6, 5, 0, 0.2432, 0.2], [1, 7, 5, 0, 0.2251, 0.2],
                    [6, 5, 5, 0, 0.4502, 0.2], [8, 8, 5, 0, 1.0, 0.2], [9, 9, 5, 0, 0.2027, 0.2]]},
        {"src": "CUDA/test/gpu/gpuNUFFT_conv_tests.cpp:611-712", "img": [10, 10, 10], "osf": 1.0,
         "kw": 4, "sw": 5, "data": [[0.5, 0.5], [0.7, 1], [1, 1], [1, 1], [1, 1]], "traj": [[-0.3, -0.1, 0, 0.5, 0.3], [0.2, 0, 0, 0, 0.3], [0, 0, 0, 0, 0]],
         "checks": [[5, 5, 5, 0, 1.3558, 0.02], [3, 6, 5, 0, 0.3101, 0.2], [1, 7, 5, 0, 0.2542, 0.2],
                    [6, 5, 5, 0, 0.5084, 0.2], [8, 8, 5, 0, 1.0, 0.2], [9, 9, 5, 0, 0.2585, 0.2]]},
        {"src": "CUDA/test/gpu/gpuNUFFT_conv_tests.cpp:714-819", "img": [10, 10, 10], "osf": 1.0,
         "kw": 5, "sw": 5, "data": [[0.5, 0.5], [0.7, 1], [1, 1], [1, 1], [1, 1]], "traj": [[-0.3, -0.1, 0, 0.5, 0.3], [0.2, 0, 0, 0, 0.3], [0, 0, 0, 0, 0]],
         "checks": [[5, 5, 5, 0, 1.3976, 0.02], [3, 6, 5, 0, 0.4268, 0.2], [6, 3, 5, 0, 0.0430, 0.2],
                    [8, 6, 5, 0, 0.1093, 0.2]]},
        {"src": "CUDA/test/gpu/gpuNUFFT_conv_tests.cpp:915-1006", "img": [16, 16, 16], "osf": 1.0,
         "kw": 3, "sw": 8, "data": [[0.5, 0.5], [0.5, 0.5], [1, 1], [1, 1], [1, 1]],
         "traj": [[0, 0, 0, 0.5, 0.3], [0, 0, 0, 0, 0.3], [0, 0, 0, 0, 0]],
         "checks": [[8, 8, 8, 0, 2.0, 0.02]]},
        # complex values
        {"src": "CUDA/test/gpu/gpuNUFFT_conv_tests.cpp:1008-1121", "img": [32, 32, 32], "osf": 1.0,
         "kw": 3, "sw": 8, "data": [[0.5, 0], [0.7, 0], [-0.2, 0.8], [-0.2, 0.8], [1, 0]],
         "traj": [[-0.3, -0.1, 0, 0, 0.5], [0.2, 0, 0, 0, 0], [0, 0, 0, 0, 0]],
         "checks": [[12, 16, 16, 0, 0.4289, 0.02], [13, 16, 16, 0, 0.6803, 0.02],
                    [14, 16, 16, 0, 0.2065, 0.02], [15, 16, 16, 0, -0.1801, 0.02],
                    [15, 16, 16, 1, 0.7206, 0.02], [16, 16, 16, 0, -0.4, 0.02],
                    [16, 16, 16, 1, 1.6, 0.02], [17, 16, 16, 0, -0.1801, 0.02],
                    [17, 16, 16, 1, 0.7206, 0.02], [12, 15, 16, 0, 0.1932, 0.02],
                    [14, 17, 16, 0, 0.0930, 0.02]]},
        # one sample of a MATLAB reconstruction; the raw array holds one
        # (x, y, z) triplet
        {"src": "CUDA/test/gpu/gpuNUFFT_conv_tests.cpp:1123-1213", "img": [32, 32, 32], "osf": 1.0,
         "kw": 3, "sw": 8, "data": [[0.0046, -0.0021]], "traj": [[0.25], [-0.4330], [0]],
         "checks": [[23, 3, 16, 0, 0.0012, 0.02], [23, 2, 16, 0, 0.0020, 0.02],
                    [23, 1, 16, 0, 0.0007, 0.02], [24, 3, 16, 0, 0.0026, 0.02],
                    [24, 3, 16, 1, -0.0012, 0.02], [24, 2, 16, 0, 0.0045, 0.02],
                    [24, 1, 16, 0, 0.0016, 0.02], [25, 3, 16, 0, 0.0012, 0.02],
                    [25, 2, 16, 0, 0.0020, 0.02], [25, 2, 16, 1, -0.0009, 0.02],
                    [25, 1, 16, 0, 0.0007, 0.02]]},
        # anisotropic grids
        {"src": "CUDA/test/gpu/gpuNUFFT_conv_tests.cpp:1215-1320", "img": [10, 10, 5], "osf": 1.0,
         "kw": 4, "sw": 5, "data": [[0.5, 0.5], [0.7, 1], [1, 1], [1, 1], [1, 1]], "traj": [[-0.3, -0.1, 0, 0.5, 0.3], [0.2, 0, 0, 0, 0.3], [0, 0, 0, 0, 0]],
         "checks": [[5, 5, 2, 0, 1.1548, 0.02], [3, 6, 2, 0, 0.2636, 0.2], [1, 7, 2, 0, 0.2161, 0.2],
                    [6, 5, 2, 0, 0.4335, 0.2], [8, 8, 2, 0, 0.8518, 0.2], [9, 9, 2, 0, 0.2197, 0.2]]},
        {"src": "CUDA/test/gpu/gpuNUFFT_conv_tests.cpp:1322-1429", "img": [20, 20, 10], "osf": 1.0,
         "kw": 4, "sw": 5, "data": [[0.5, 0.5], [0.7, 1], [1, 1], [1, 1], [1, 1]], "traj": [[-0.3, -0.1, 0, 0.5, 0.3], [0.2, 0, 0, 0, 0.3], [0, 0, 0, 0, 0]],
         "checks": [[10, 10, 5, 0, 1.0, 0.02], [9, 10, 5, 0, 0.8639, 0.2], [1, 11, 5, 0, 0.2582, 0.2],
                    [7, 11, 5, 0, 0.1808, 0.2], [9, 9, 4, 0, 0.2231, 0.2], [9, 9, 6, 0, 0.2231, 0.2]]},
        # osf 1.5 (grid 30 x 30 x 15), kernel width 4; the balanced operator
        # of :1517-1605 asserts the same values
        {"src": "CUDA/test/gpu/gpuNUFFT_conv_tests.cpp:1431-1515", "img": [20, 20, 10], "osf": 1.5,
         "kw": 4, "sw": 5, "data": [[1.0, 0.5], [3.0, 1], [3.0, 1], [4.0, 1], [5.0, 1]],
         "traj": [[-0.3, 0.3, 0.0, 0.5, 0.3], [0.2, 0.0, 0.0, 0.0, 0.3],
                  [-0.25, -0.005, 0.0, 0.1, 0.22]],
         "checks": [[6, 21, 3, 0, 0.5853, 0.02], [6, 21, 4, 0, 0.9445, 0.02],
                    [24, 24, 11, 0, 4.8931, 0.2]]},
        # 2-D
        {"src": "CUDA/test/gpu/gpuNUFFT_2D_conv_tests.cpp:106-183", "img": [10, 10], "osf": 1.0,
         "kw": 3, "sw": 5, "data": [[1, 1]], "traj": [[0], [0]],
         "checks": [[5, 5, 0, 0, 1.0, 0.02], [5, 4, 0, 0, 0.4502, 0.2], [4, 5, 0, 0, 0.4502, 0.2],
                    [5, 6, 0, 0, 0.4502, 0.2], [6, 6, 0, 0, 0.2027, 0.2], [4, 4, 0, 0, 0.2027, 0.2],
                    [4, 6, 0, 0, 0.2027, 0.2]]},
        {"src": "CUDA/test/gpu/gpuNUFFT_2D_conv_tests.cpp:185-255", "img": [10, 10], "osf": 1.0,
         "kw": 5, "sw": 5, "data": [[1, 1]], "traj": [[0], [0]],
         "checks": [[5, 5, 0, 0, 1.0, 0.02], [3, 3, 0, 0, 0.0049, 0.2], [4, 4, 0, 0, 0.3218, 0.2],
                    [5, 4, 0, 0, 0.5673, 0.2], [5, 7, 0, 0, 0.0697, 0.2], [5, 3, 0, 0, 0.0697, 0.2]]},
        {"src": "CUDA/test/gpu/gpuNUFFT_2D_conv_tests.cpp:257-356 (and :358-455)", "img": [10, 10],
         "osf": 1.0, "kw": 3, "sw": 5, "data": [[0.5, 0.5], [0.7, 1], [1, 1], [1, 1], [1, 1]], "traj": [[-0.3, -0.1, 0, 0.5, 0.3], [0.2, 0, 0, 0, 0.3]],
         "checks": [[5, 5, 0, 0, 1.3152, 0.02], [3, 6, 0, 0, 0.2432, 0.2], [1, 7, 0, 0, 0.2251, 0.2],
                    [6, 5, 0, 0, 0.4502, 0.2], [8, 8, 0, 0, 1.0, 0.2], [9, 9, 0, 0, 0.2027, 0.2]]},
        {"src": "CUDA/test/gpu/gpuNUFFT_2D_conv_tests.cpp:457-554", "img": [10, 10], "osf": 1.0,
         "kw": 4, "sw": 5, "data": [[0.5, 0.5], [0.7, 1], [1, 1], [1, 1], [1, 1]], "traj": [[-0.3, -0.1, 0, 0.5, 0.3], [0.2, 0, 0, 0, 0.3]],
         "checks": [[5, 5, 0, 0, 1.3558, 0.02], [3, 6, 0, 0, 0.3101, 0.2], [1, 7, 0, 0, 0.2542, 0.2],
                    [6, 5, 0, 0, 0.5084, 0.2], [8, 8, 0, 0, 1.0, 0.2], [9, 9, 0, 0, 0.2585, 0.2]]},
        {"src": "CUDA/test/gpu/gpuNUFFT_2D_conv_tests.cpp:556-651", "img": [10, 10], "osf": 1.0,
         "kw": 5, "sw": 5, "data": [[0.5, 0.5], [0.7, 1], [1, 1], [1, 1], [1, 1]], "traj": [[-0.3, -0.1, 0, 0.5, 0.3], [0.2, 0, 0, 0, 0.3]],
         "checks": [[5, 5, 0, 0, 1.3976, 0.02], [3, 6, 0, 0, 0.4268, 0.2], [6, 3, 0, 0, 0.0430, 0.2],
                    [8, 6, 0, 0, 0.1093, 0.2]]},
        {"src": "CUDA/test/gpu/gpuNUFFT_2D_conv_tests.cpp:819-924", "img": [32, 32], "osf": 1.0,
         "kw": 3, "sw": 8, "data": [[0.5, 0], [0.7, 0], [-0.2, 0.8], [-0.2, 0.8], [1, 0]],
         "traj": [[-0.3, -0.1, 0, 0, 0.5], [0.2, 0, 0, 0, 0]],
         "checks": [[12, 16, 0, 0, 0.4289, 0.02], [13, 16, 0, 0, 0.6803, 0.02],
                    [14, 16, 0, 0, 0.2065, 0.02], [15, 16, 0, 0, -0.1801, 0.02],
                    [15, 16, 0, 1, 0.7206, 0.02], [16, 16, 0, 0, -0.4, 0.02],
                    [16, 16, 0, 1, 1.6, 0.02], [17, 16, 0, 0, -0.1801, 0.02],
                    [17, 16, 0, 1, 0.7206, 0.02], [12, 15, 0, 0, 0.1932, 0.02],
                    [14, 17, 0, 0, 0.0930, 0.02]]},
        {"src": "CUDA/test/gpu/gpuNUFFT_2D_conv_tests.cpp:1013-1091 (and 4 channels, every factory "
                "flavour, :1093-1204)", "img": [256, 128], "osf": 1.0, "kw": 3, "sw": 8,
         "data": [[1.0, 0.5]], "traj": [[0.0], [0.0]],
         "checks": [[128, 64, 0, 0, 1.0, 0.02], [128, 64, 0, 1, 0.5, 0.02],
                    [127, 64, 0, 0, 0.45, 0.02]]},
    ],
    # CUDA/test/gpu/gpuNUFFT_fft_tests.cpp — adjoint FFT stop (after the
    # inverse FFT, shifts, crop and 1/sqrt(Ng) scaling, before the
    # deapodization), nearest lookup; same check layout as gpu_conv.
    "gpu_fft_stop": [
        {"src": "CUDA/test/gpu/gpuNUFFT_fft_tests.cpp:142-215", "img": [10, 10, 10], "osf": 1.0,
         "kw": 3, "sw": 5, "data": [[1, 1]], "traj": [[0], [0], [0]],
         "checks": [[5, 5, 5, 0, 0.2172, 0.02], [5, 4, 5, 0, 0.1975, 0.2], [4, 5, 5, 0, 0.1975, 0.2],
                    [0, 1, 5, 0, 0.001619, 0.2]]},
        {"src": "CUDA/test/gpu/gpuNUFFT_fft_tests.cpp:217-288", "img": [64, 64, 64], "osf": 1.0,
         "kw": 5, "sw": 8, "data": [[1, 0]], "traj": [[0], [0], [0]],
         "checks": [[57, 33, 32, 0, 0.0001, 0.2]]},
        {"src": "CUDA/test/gpu/gpuNUFFT_fft_tests.cpp:290-388", "img": [16, 16, 16], "osf": 1.0,
         "kw": 3, "sw": 8, "data": [[0.5, 0.5], [0.5, 0.5], [1, 1], [1, 1], [1, 1]],
         "traj": [[0, 0, 0, 0.5, 0.3], [0, 0, 0, 0, 0.3], [0, 0, 0, 0, 0]],
         "checks": [[8, 8, 5, 0, 0.28397, 0.02], [2, 2, 5, 0, 0.0039668, 0.02],
                    [4, 3, 5, 0, 0.0159846, 0.02]]},
    ],
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_fixtures.json")
    with open(out, "w") as f:
        json.dump(FIX, f, indent=1)
    print("wrote", out)
