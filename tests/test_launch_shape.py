"""Host logic of the advance's default strip height (hx::host::default_rows in
hydroexa_amd/csrc/hx_host.hpp), compiled for the host with hipcc and run on
the CPU: the rule's values at the sizes its measurements cover."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC") or shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

SRC = r"""
#include <cstdio>
#include "hx_host.hpp"
int main(int argc, char** argv)
{
    (void)argc;
    const int solver = atoi(argv[1]), dtype = atoi(argv[2]), nx = atoi(argv[3]);
    const int ny = atoi(argv[4]), cus = atoi(argv[5]), waves = atoi(argv[6]);
    printf("%d\n", hx::host::default_rows(solver, dtype, nx, ny, cus, waves));
    return 0;
}
"""

ROE, HYD = 0, 1
F64, F32 = 0, 1


@pytest.fixture(scope="module")
def rule(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    d = tmp_path_factory.mktemp("launch_shape")
    src, exe = d / "rule.cpp", d / "rule"
    src.write_text(SRC)
    subprocess.run([HIPCC, "-std=c++17", f"-I{ROOT}/include", f"-I{ROOT}/hydroexa_amd/csrc",
                    str(src), "-o", str(exe)], check=True, capture_output=True)

    def run(solver, dtype, nx, ny, cus=256, waves=0):
        out = subprocess.run([str(exe), *map(str, (solver, dtype, nx, ny, cus, waves))],
                             check=True, capture_output=True, text=True).stdout
        return int(out)
    return run


def test_streaming_marches_keep_16_rows_when_the_level_fills_the_device(rule):
    # 16-row strips give at least 3 wavefronts per SIMD on 256 CUs
    assert rule(ROE, F64, 4096, 4096) == 16
    assert rule(ROE, F64, 2048, 2048) == 16
    assert rule(ROE, F32, 2048, 2048) == 16
    assert rule(HYD, F32, 4096, 4096) == 16


def test_small_levels_take_6_row_strips(rule):
    for n in (256, 512, 1024):
        assert rule(ROE, F64, n, n) == 6
        assert rule(ROE, F32, n, n) == 6
    assert rule(HYD, F32, 1024, 1024) == 6


def test_hydrostatic_fp64_height_minimises_rounds_times_rows(rule):
    # 3 resident wavefronts per SIMD, 256 CUs: 3072 slots
    assert rule(HYD, F64, 2048, 2048, 256, 3) == 24     # 3010 wavefronts: one round
    assert rule(HYD, F64, 4096, 4096, 256, 3) == 31     # 9177 wavefronts: three rounds
    assert rule(HYD, F64, 1024, 1024, 256, 3) == 8      # below one round: the shortest
    # brute force of the stated cost over [8, 32], larger height on ties
    for (nx, ny) in ((512, 512), (1000, 3000), (4096, 1024), (2048, 2048)):
        nsx = -(-nx // 60)
        slots = 256 * 4 * 3
        best = min(range(8, 33), key=lambda r: (-(-nsx * -(-ny // r) // slots) * (r + 4), -r))
        assert rule(HYD, F64, nx, ny, 256, 3) == best


def test_unknown_sizes_fall_back_to_16(rule):
    assert rule(ROE, F64, 0, 0) == 16
    assert rule(HYD, F64, 2048, 2048, 256, 0) == 16      # occupancy query failed
