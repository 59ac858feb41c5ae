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


OFFSETS_SRC = r"""
#include <cstdio>
#include <cstdlib>
#include "hx_host.hpp"
// Largest byte offsets a strip of the fused advance forms from its base
// (2 rows and 2 columns before the strip's first cell): input rows up to
// rows + 3 (the hydrostatic march's second halo row above the strip),
// input columns up to the last stored column + 2, outputs up to row
// rows - 1 and column nx - 1.
template <typename R>
static void run(int solver, long nx, long pitch, long opitch, int ghost_hi)
{
    hx::StepArgs<R> a{};
    a.p.nx = (int)nx;
    a.p.ny = 64;
    a.p.pitch = pitch;
    a.opitch = opitch;
    a.p.bh0 = ghost_hi ? HX_BC_GHOST : HX_BC_TRANSMIT;
    hx::host::fill_geometry(a, solver, 16, 1.0, 1.0);
    const long sz = (long)sizeof(R);
    const long in_max = (long)(a.rows + 3) * pitch * sz + (nx - 1 + ghost_hi + 2) * sz;
    const long out_max = (long)(a.rows - 1) * opitch * sz + (nx - 1) * sz;
    printf("%d %ld %ld\n", a.rows, in_max, out_max);
}
int main(int argc, char** argv)
{
    (void)argc;
    const int f64 = atoi(argv[1]), solver = atoi(argv[2]), ghost_hi = atoi(argv[6]);
    const long nx = atol(argv[3]), pitch = atol(argv[4]), opitch = atol(argv[5]);
    if (f64) run<double>(solver, nx, pitch, opitch, ghost_hi);
    else run<float>(solver, nx, pitch, opitch, ghost_hi);
    return 0;
}
"""


@pytest.fixture(scope="module")
def offsets(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    d = tmp_path_factory.mktemp("strip_offsets")
    src, exe = d / "offsets.cpp", d / "offsets"
    src.write_text(OFFSETS_SRC)
    subprocess.run([HIPCC, "-std=c++17", f"-I{ROOT}/include", f"-I{ROOT}/hydroexa_amd/csrc",
                    str(src), "-o", str(exe)], check=True, capture_output=True)

    def run(f64, solver, nx, pitch, opitch, ghost_hi):
        out = subprocess.run([str(exe), *map(str, (int(f64), solver, nx, pitch, opitch, ghost_hi))],
                             check=True, capture_output=True, text=True).stdout.split()
        return tuple(int(v) for v in out)
    return run


WIDEST = 1 << 26      # the largest pitch check_patch and hx_advance_level accept


@pytest.mark.parametrize("f64", [True, False])
@pytest.mark.parametrize("solver", [ROE, HYD])
@pytest.mark.parametrize("nx,pitch,opitch,ghost_hi", [
    (WIDEST, WIDEST, WIDEST, 0),              # pitch == nx, physical sides
    (WIDEST - 2, WIDEST, WIDEST, 2),          # a 2-wide stored high ghost fills the pitch
    (WIDEST - 2, WIDEST, WIDEST - 2, 2),      # output pitch narrower than the input's
    (1000, 1000, WIDEST, 0),                  # output pitch is the wide one
    ((1 << 25) + 7, (1 << 25) + 8, (1 << 25) + 7, 1),
])
def test_strip_offsets_stay_below_4_gib(offsets, f64, solver, nx, pitch, opitch, ghost_hi):
    rows, in_max, out_max = offsets(f64, solver, nx, pitch, opitch, ghost_hi)
    assert rows >= 1
    assert in_max < 2 ** 32 and out_max < 2 ** 32, (rows, in_max, out_max)


def test_strip_clamp_leaves_ordinary_levels_alone(offsets):
    assert offsets(True, ROE, 4096, 4096, 4096, 0)[0] == 16
    assert offsets(True, HYD, 4100, 4104, 4100, 2)[0] == 16
    # at a pitch of 2^26 FP64 elements three rows fit
    assert offsets(True, HYD, WIDEST, WIDEST, WIDEST, 0)[0] == 3
