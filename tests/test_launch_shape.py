"""Host logic of the advance's default strip height (hx::host::default_rows in
hydroexa_amd/csrc/hx_host.hpp), compiled for the host with hipcc and run on
the CPU: the rule's values at the sizes its measurements cover, and the
stated cost restated here for other shapes, CU counts and occupancies."""
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
    const bool exact = argc > 7 && atoi(argv[7]) != 0;
    printf("%d\n", hx::host::default_rows(solver, dtype, nx, ny, cus, waves, exact));
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

    def run(solver, dtype, nx, ny, cus=256, waves=0, exact=False):
        out = subprocess.run([str(exe), *map(str, (solver, dtype, nx, ny, cus, waves, int(exact)))],
                             check=True, capture_output=True, text=True).stdout
        return int(out)
    return run


# waves per SIMD the advance kernels hold on MI355X (their register counts)
WAVES = {(ROE, F64): 3, (ROE, F32): 6, (HYD, F64): 3, (HYD, F32): 5}
# per-row cost to the busiest CU holding k = 1..8 workgroups, hundredths
ROW_COST = {(ROE, F64): (81, 121, 169, 244, 303, 365, 427, 499),
            (ROE, F32): (73, 84, 108, 135, 166, 196, 234, 260),
            (HYD, F64): (164, 212, 270, 352, 418, 472, 558, 613),
            (HYD, F32): (139, 162, 204, 249, 293, 341, 382, 422)}


def busiest_cu(nx, ny, r, cols, cus):
    """Workgroups (4 strips each, dealt round-robin to 8 XCDs) the busiest CU
    holds at strip height r."""
    wgs = -(-(-(-nx // cols) * -(-ny // r)) // 4)
    return -(-(-(-wgs // 8)) // max(1, cus // 8))


def banded_at_16(solver, dtype, nx, ny, cus, waves):
    cols = 60 if solver == HYD else 62
    return -(-nx // cols) * -(-ny // 16) >= 2 * 8 * max(1, cus // 8) * 4 * waves


def small_level_cost(solver, dtype, nx, ny, r, cus):
    """The stated cost of an r-row launch on a level below the banded
    tiling: (r + prologue rows) x the busiest CU's per-row cost."""
    cols, prologue = (60, 4) if solver == HYD else (62, 2)
    g = ROW_COST[(solver, dtype)]
    k = busiest_cu(nx, ny, r, cols, cus)
    cost = g[k - 1] if k <= 8 else g[7] + (k - 8) * (g[7] - g[3]) // 4
    return (r + prologue) * cost


def rounds_cost(nx, ny, r, cus, waves):
    """The stated cost of an r-row hydrostatic FP64 launch on a banded level:
    the busiest CU's workgroups run in rounds of `waves`, a row costs
    0.78/1.10/1.60 with 1/2/3 resident."""
    k = busiest_cu(nx, ny, r, 60, cus)
    unit = {0: 0, 1: 78, 2: 110, 3: 160}
    rounds = [waves] * (k // waves) + [k % waves]
    return sum(unit[m] if m <= 3 else 160 * m // 3 for m in rounds) * (r + 4)


def test_levels_that_fill_the_device_twice_keep_16_rows(rule):
    assert rule(ROE, F64, 4096, 4096, 256, 3) == 16     # the banded headline launch
    assert rule(ROE, F64, 3072, 3072, 256, 3) == 16
    assert rule(ROE, F32, 4096, 4096, 256, 6) == 16
    assert rule(HYD, F32, 4096, 4096, 256, 5) == 16


def test_hydrostatic_fp64_on_banded_levels_weighs_rounds_of_resident_wavefronts(rule):
    assert rule(HYD, F64, 4096, 4096, 256, 3) == 31     # three rounds of 3
    for (nx, ny) in ((4096, 4096), (3072, 3072), (8192, 2048), (6000, 4100)):
        for cus, waves in ((256, 3), (256, 2), (304, 3)):
            assert banded_at_16(HYD, F64, nx, ny, cus, waves)
            best = min(range(8, 33), key=lambda r: (rounds_cost(nx, ny, r, cus, waves), -r))
            assert rule(HYD, F64, nx, ny, cus, waves) == best, (nx, ny, cus, waves)


def test_smaller_levels_take_the_least_modelled_busiest_cu_time(rule):
    # heights timed as the defaults on one MI355X (at or below the best of
    # the swept heights 4-32): the smallest height of a band of equal
    # workgroups per CU
    assert rule(ROE, F64, 384, 384, 256, 3) == 4        # one workgroup per CU
    assert rule(ROE, F64, 1280, 1280, 256, 3) == 9      # 751 workgroups: 3 a CU
    assert rule(ROE, F64, 2048, 2048, 256, 3) == 23
    assert rule(ROE, F32, 2048, 2048, 256, 6) == 14
    assert rule(HYD, F64, 384, 384, 256, 3) == 4        # 13.0 us against 19.4 at 8 rows
    assert rule(HYD, F64, 512, 512, 256, 3) == 5
    assert rule(HYD, F64, 1024, 1024, 256, 3) == 10
    assert rule(HYD, F32, 1024, 1024, 256, 5) == 5
    # brute force of the stated cost, larger height on ties
    shapes = ((384, 384), (512, 512), (1000, 3000), (4096, 1024), (2048, 2048), (3000, 700),
              (130, 9000))
    for solver, dtype in WAVES:
        for (nx, ny) in shapes:
            for cus in (256, 304, 80):
                waves = WAVES[(solver, dtype)]
                if banded_at_16(solver, dtype, nx, ny, cus, waves):
                    continue
                best = min(range(4, 33),
                           key=lambda r: (small_level_cost(solver, dtype, nx, ny, r, cus), -r))
                assert rule(solver, dtype, nx, ny, cus, waves) == best, (solver, dtype, nx, ny, cus)


def test_exact_fp64_plain_roe_on_banded_levels_takes_the_arithmetic_bound_rule(rule):
    # the exact build's FP64 march is arithmetic-bound: 30 rows at 4096^2
    # (measured 293.7-295.4 us at 32 rows against 307-309 at 16)
    assert rule(ROE, F64, 4096, 4096, 256, 3, exact=True) == 30
    assert rule(ROE, F64, 4096, 4096, 256, 3, exact=False) == 16
    for (nx, ny) in ((4096, 4096), (3072, 3072), (8192, 2048)):
        best = min(range(8, 33), key=lambda r: (rounds_cost_cols(nx, ny, r, 256, 3, 62, 2), -r))
        assert rule(ROE, F64, nx, ny, 256, 3, exact=True) == best
    # FP32 and levels below the banded tiling keep their rules
    assert rule(ROE, F32, 4096, 4096, 256, 6, exact=True) == 16
    for n in (384, 1024, 2048):
        assert rule(ROE, F64, n, n, 256, 3, exact=True) == rule(ROE, F64, n, n, 256, 3)


def rounds_cost_cols(nx, ny, r, cus, waves, cols, prologue):
    k = busiest_cu(nx, ny, r, cols, cus)
    unit = {0: 0, 1: 78, 2: 110, 3: 160}
    rounds = [waves] * (k // waves) + [k % waves]
    return sum(unit[m] if m <= 3 else 160 * m // 3 for m in rounds) * (r + prologue)


def test_unknown_sizes_fall_back_to_16(rule):
    assert rule(ROE, F64, 0, 0, 256, 3) == 16
    assert rule(HYD, F64, 2048, 2048, 256, 0) == 16      # occupancy query failed
    assert rule(ROE, F32, 512, 512, 256, 0) == 16


MULTI_SRC = r"""
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "hx_host.hpp"
// argv: solver dtype cus waves n nx0 ny0 nx1 ny1 ...
int main(int argc, char** argv)
{
    (void)argc;
    const int solver = atoi(argv[1]), dtype = atoi(argv[2]), cus = atoi(argv[3]);
    const int waves = atoi(argv[4]), n = atoi(argv[5]);
    std::vector<int> nx(n), ny(n);
    for (int q = 0; q < n; ++q) { nx[q] = atoi(argv[6 + 2 * q]); ny[q] = atoi(argv[7 + 2 * q]); }
    printf("%d\n", hx::host::default_rows_multi(solver, dtype, nx.data(), ny.data(), n, cus, waves));
    return 0;
}
"""


@pytest.fixture(scope="module")
def multi_rule(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    d = tmp_path_factory.mktemp("launch_shape_multi")
    src, exe = d / "multi.cpp", d / "multi"
    src.write_text(MULTI_SRC)
    subprocess.run([HIPCC, "-std=c++17", f"-I{ROOT}/include", f"-I{ROOT}/hydroexa_amd/csrc",
                    str(src), "-o", str(exe)], check=True, capture_output=True)

    def run(solver, dtype, sizes, cus=256, waves=3):
        flat = [str(v) for wh in sizes for v in wh]
        out = subprocess.run([str(exe), str(solver), str(dtype), str(cus), str(waves),
                              str(len(sizes)), *flat], check=True, capture_output=True,
                             text=True).stdout
        return int(out)
    return run


def batch_rule(solver, dtype, sizes, cus, waves):
    """The stated rule over a batch: strips summed over the patches."""
    cols = 60 if solver == HYD else 62

    def strips(r):
        return sum(-(-nx // cols) * -(-ny // r) for nx, ny in sizes)

    def k_of(r):
        return -(-(-(-(-(-strips(r) // 4)) // 8)) // max(1, cus // 8))
    if strips(16) >= 2 * 8 * max(1, cus // 8) * 4 * waves:
        if (solver, dtype) != (HYD, F64):
            return 16
        unit = {0: 0, 1: 78, 2: 110, 3: 160}

        def rounds(r):
            k = k_of(r)
            parts = [waves] * (k // waves) + [k % waves]
            return sum(unit[m] if m <= 3 else 160 * m // 3 for m in parts) * (r + 4)
        return min(range(8, 33), key=lambda r: (rounds(r), -r))
    g = ROW_COST[(solver, dtype)]
    prologue = 4 if solver == HYD else 2

    def cost(r):
        k = k_of(r)
        return (r + prologue) * (g[k - 1] if k <= 8 else g[7] + (k - 8) * (g[7] - g[3]) // 4)
    return min(range(4, 33), key=lambda r: (cost(r), -r))


def test_one_patch_batch_takes_the_level_height(rule, multi_rule):
    for solver, dtype in WAVES:
        for (nx, ny) in ((384, 384), (1280, 1280), (2048, 2048), (4096, 4096), (3000, 700)):
            w = WAVES[(solver, dtype)]
            assert multi_rule(solver, dtype, [(nx, ny)], 256, w) == rule(solver, dtype, nx, ny, 256, w)


def test_a_batch_takes_one_height_for_its_summed_strips(multi_rule):
    import random
    rnd = random.Random(128)
    batch = [(rnd.randint(128, 512), rnd.randint(128, 512)) for _ in range(128)]
    small = batch[:6]
    for solver, dtype in WAVES:
        w = WAVES[(solver, dtype)]
        for sizes in (batch, small):
            assert multi_rule(solver, dtype, sizes, 256, w) == batch_rule(solver, dtype, sizes, 256, w)
    # 128 scenarios of 128^2-512^2 fill the device twice at 16 rows: the
    # hydrostatic FP64 batch takes the banded rule, not a small scenario's 4-5 rows
    assert multi_rule(HYD, F64, batch, 256, 3) >= 8


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


BANDS_SRC = r"""
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "hx_host.hpp"
// Plans the banded tiling of an nx x ny level (plan_bands) and enumerates
// every item of the grid through band_strip, as advance_kernel does: prints
// "uniform", or "banded <workgroups> <tail rows> <tail band> <max height>"
// followed by "ok" when every cell of every strip column is marched exactly
// once and no strip is taller than the planned height.
int main(int argc, char** argv)
{
    (void)argc;
    const int solver = atoi(argv[1]), nx = atoi(argv[2]), ny = atoi(argv[3]);
    const int rows = atoi(argv[4]), cus = atoi(argv[5]), waves = atoi(argv[6]);
    hx::StepArgs<double> a{};
    a.p.nx = nx;
    a.p.ny = ny;
    a.p.pitch = nx;
    a.opitch = nx;
    hx::host::fill_geometry(a, solver, rows, 1.0, 1.0);
    hx::host::plan_bands(a, solver, cus, waves);
    if (!a.band_wgs) {
        printf("uniform\n");
        return 0;
    }
    std::vector<int> hits((size_t)a.nsx * ny, 0);
    int tallest = 0;
    bool ok = true;
    for (long b = 0; b < 8L * a.band_wgs; ++b)
        for (int w = 0; w < hx::kWaves; ++w) {
            int sxi, y0, y1;
            hx::band_strip(a, (unsigned)(b & 7), (int)((b >> 3) * hx::kWaves + w), sxi, y0, y1);
            if (y0 >= y1) continue;
            if (sxi < 0 || sxi >= a.nsx || y0 < 0 || y1 > ny) { ok = false; continue; }
            tallest = y1 - y0 > tallest ? y1 - y0 : tallest;
            for (int y = y0; y < y1; ++y) hits[(size_t)y * a.nsx + sxi]++;
        }
    for (int v : hits) ok = ok && (v == 1);
    printf("banded %d %d %d %d %s\n", a.band_wgs, a.tail_rows, a.tail_band, tallest,
           ok && tallest <= a.rows ? "ok" : "bad");
    return 0;
}
"""


@pytest.fixture(scope="module")
def bands(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    d = tmp_path_factory.mktemp("bands")
    src, exe = d / "bands.cpp", d / "bands"
    src.write_text(BANDS_SRC)
    subprocess.run([HIPCC, "-std=c++17", f"-I{ROOT}/include", f"-I{ROOT}/hydroexa_amd/csrc",
                    str(src), "-o", str(exe)], check=True, capture_output=True)

    def run(solver, nx, ny, rows, cus=256, waves=3):
        return subprocess.run([str(exe), *map(str, (solver, nx, ny, rows, cus, waves))],
                              check=True, capture_output=True, text=True).stdout.split()
    return run


@pytest.mark.parametrize("solver,nx,ny,rows", [
    (ROE, 4096, 4096, 16), (ROE, 4096, 4103, 16), (ROE, 1000, 8192, 16), (ROE, 4096, 4096, 32),
    (HYD, 4096, 4096, 31), (HYD, 4100, 4097, 16), (HYD, 1000, 8192, 16), (ROE, 8192, 8192, 16),
])
def test_banded_tiling_marches_every_cell_once(bands, solver, nx, ny, rows):
    out = bands(solver, nx, ny, rows)
    assert out[0] == "banded", out
    assert out[-1] == "ok", out
    wgs, tail_rows, tail_band, tallest = map(int, out[1:5])
    assert tail_rows == (4 if solver == ROE else 8) and tail_band >= tail_rows
    assert tallest == rows


def test_banded_tiling_at_the_headline_size(bands):
    # 4096^2 plain Roe, 16-row strips, 3 wavefronts per SIMD on 256 CUs: one
    # round of an XCD's 384 slots in 4-row strips is 23 rows of 67 strips,
    # rounded to 24; the 512-row bands keep 30 strips of 16 rows and 8 of 4
    out = bands(ROE, 4096, 4096, 16)
    assert out == ["banded", str((67 * 38 + 3) // 4), "4", "24", "16", "ok"]


def test_levels_that_fill_the_device_less_than_twice_stay_uniform(bands):
    assert bands(ROE, 2048, 2048, 16) == ["uniform"]        # 4352 strips < 2 x 3072
    assert bands(HYD, 2048, 2048, 24) == ["uniform"]
    assert bands(ROE, 4096, 4096, 4) == ["uniform"]         # no shorter tail strip
    assert bands(HYD, 4096, 4096, 8) == ["uniform"]
    assert bands(ROE, 4096, 4096, 16, 256, 0) == ["uniform"]  # occupancy unknown
