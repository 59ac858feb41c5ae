// hx_host.hpp -- host-side internals shared by the library's translation
// units: the context behind the opaque hx_ctx handle, patch views, launch
// geometry and the fused-advance launcher (defined in hx_kernels.hip).
#pragma once

#include <algorithm>
#include <cstdint>

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/hydroexa.h"
#include "hx_types.hpp"

namespace hx {
// Steps of dt / max|vel| history the context holds; longer runs are done in
// chunks of this many steps, with one synchronisation per chunk.
constexpr int kHistCap = 4096;
struct TwoLevelWork {
    CflSlot slots[4];   // coarse read/write, fine read/write
    double dt[2];
    int err;
    int pad;
};
// Where the last successful hx_run_level / hx_run_level_resume on a context
// left its rollout: the final state's buffers and geometry, and the slot that
// holds the CFL estimate of that state (reduced by the last fused step).
struct RunMark {
    int valid;
    int solver, dtype, nx, ny, slot;
    int bc[4];
    int64_t pitch;
    double dx, dy;
    const void *h, *hu, *hv, *z, *n;
};
}  // namespace hx

struct hx_ctx {
    int device;
    hx::CflSlot* slots;  // 4 slots: 3 for the fused loop, 1 for hx_level_dt
    int* err;
    double* hist;        // 2 * hist_cap doubles (dt, vel), allocated with the context
    double* hpin;        // pinned host staging of a run chunk's error flag and history
    int hist_cap;
    hx::TwoLevelWork* tlw;   // hx_two_level_run's slots, dt pair and error flag
    hx::CflSlot* amr_slots;  // hx_amr_run's per-level CFL slots (HX_AMR_MAX_LEVELS)
    hx::AmrPatch* amr_tab;   // hx_amr_run's patch tables (amr_cap entries, grown on regrid)
    hx::AmrPatch* amr_host;  // pinned staging of the tables; holds what amr_tab holds
    int amr_cap;
    int amr_n;               // entries amr_tab holds (-1: none)
    hx::RunMark mark;
    void* scratch;           // this device's static scratch (hx_run_level_noalloc)
    uint32_t* pflags;    // prune flags
    int32_t* pboxes;
    int pcap;
    int rows[2][2];      // [solver][dtype]; 0: the default of the arithmetic build
    int occ[2][2];       // [solver][dtype] minimum waves per SIMD requested; 0: default
    int arith;           // HX_ARITH_EXACT or HX_ARITH_FAST
    int cus;             // compute units of the device
    double* cons;        // conservation partials (kConsBlocks x 9) + result (5)
};

namespace hx {
namespace host {

// Launch shape of the fused advance for (solver, dtype): the value set by
// hx_set_tuning, or the measured default (one MI355X):
// * 3 waves per SIMD for FP64, 1 requested for FP32 (the FP32 kernels hold
//   6 (plain Roe) and 5 (hydrostatic) by their register counts);
// * the strip height from the level size:
//   - levels that fill the device twice at 16-row strips (the banded tiling
//     marches them, plan_bands): plain Roe and FP32 stream HBM at 16 rows
//     (14-20 rows within 1-3 % on the 4096^2 dam break, 24-128 rows 5-30 %
//     slower; 2560^2 and 3072^2 FP64 best at 14-16); the FP64 hydrostatic
//     march, and the exact build's FP64 plain-Roe march, are arithmetic-bound
//     and end on a tail of idle SIMDs, so their height is the r in [8, 32] of
//     least (sum over rounds of resident wavefronts of the round's per-row
//     cost, round_units) x (r + prologue rows): 31 rows at 4096^2 hydrostatic
//     (5 % faster than 16), 30 plain Roe exact (4.5 % faster);
//   - smaller levels: a launch lasts as long as its busiest CU.  Workgroups
//     are dealt round-robin to the 8 XCDs, so that CU holds
//     k = ceil(ceil(workgroups / 8) / (CUs / 8)) of them; a strip of r rows
//     marches r + 2 (plain Roe) or r + 4 (hydrostatic) rows at a per-row cost
//     row_cost(k) fitted to heights 4-32 on 384^2 to 3072^2 levels of each
//     march.  The height is the r in [4, 32] of least (r + prologue) x
//     row_cost(k), the larger on ties.  On those
//     levels its choice measured 0.1-0.6 % above the best height on average
//     (at most 5 %), where the rule it replaces (6 rows below a fill of
//     3 wavefronts per SIMD, else 16) was 2-11 % above (at most 22 %).
inline int default_occ(int solver, int dtype)
{
    (void)solver;
    return (dtype == HX_F64) ? 3 : 1;
}
// Cost of one row to a SIMD holding m wavefronts of the FP64 hydrostatic
// march, in hundredths (0 for none), on levels that fill the device twice.
inline int64_t round_units(int64_t m)
{
    return m <= 0 ? 0 : m == 1 ? 78 : m == 2 ? 110 : m == 3 ? 160 : 160 * m / 3;
}
// Per-row cost of a strip to the busiest CU when it holds k workgroups of a
// march, in hundredths of a microsecond, k = 1..8 (least squares over the
// measured step times, made non-decreasing); beyond 8 the cost grows by the
// mean step of k = 4..8.
constexpr int kRowCost[2][2][8] = {
    {{81, 121, 169, 244, 303, 365, 427, 499},      // plain Roe, FP64
     {73, 84, 108, 135, 166, 196, 234, 260}},      // plain Roe, FP32
    {{164, 212, 270, 352, 418, 472, 558, 613},     // hydrostatic, FP64
     {139, 162, 204, 249, 293, 341, 382, 422}}};   // hydrostatic, FP32
inline int64_t row_cost(int solver, int dtype, int64_t k)
{
    const int* g = kRowCost[solver == HX_SOLVER_HYDROSTATIC_ROE ? 1 : 0][dtype == HX_F64 ? 0 : 1];
    if (k <= 8) return g[k < 1 ? 0 : k - 1];
    return g[7] + (k - 8) * (g[7] - g[3]) / 4;
}
// The rule of default_rows over the strips one launch deals out:
// strips(r) is their number at height r (summed over the patches of a
// launch that covers several).  exact: the exact arithmetic build, whose FP64
// plain-Roe march is arithmetic-bound too (VALU 92 % busy at 4096^2): on
// banded levels it takes the hydrostatic FP64 rule (30 rows at 4096^2:
// 293.7-295.4 us per step at 32 rows, 307-309 at 16).
template <typename F>
inline int rows_rule(int solver, int dtype, F strips, int cus, int simd_waves, bool exact = false)
{
    const bool hydro = (solver == HX_SOLVER_HYDROSTATIC_ROE);
    const bool arith_bound = dtype == HX_F64 && (hydro || exact);
    const int prologue = hydro ? 4 : 2;
    const int64_t cus_xcd = cus >= 8 ? cus / 8 : 1;
    const int64_t slots_xcd = cus_xcd * 4 * simd_waves;
    int best = 32;
    int64_t cost = INT64_MAX;
    if (strips(16) >= 2 * 8 * slots_xcd) {
        if (!arith_bound) return 16;
        for (int r = 8; r <= 32; ++r) {
            const int64_t wgs = (strips(r) + 3) / 4;
            const int64_t k = ((wgs + 7) / 8 + cus_xcd - 1) / cus_xcd;
            const int64_t units = (k / simd_waves) * round_units(simd_waves) + round_units(k % simd_waves);
            const int64_t c = units * (r + prologue);
            if (c <= cost) { cost = c; best = r; }
        }
        return best;
    }
    for (int r = 4; r <= 32; ++r) {
        const int64_t wgs = (strips(r) + 3) / 4;
        const int64_t k = ((wgs + 7) / 8 + cus_xcd - 1) / cus_xcd;
        const int64_t c = (r + prologue) * row_cost(solver, dtype, k);
        if (c <= cost) { cost = c; best = r; }
    }
    return best;
}
inline int default_rows(int solver, int dtype, int nx, int ny, int cus, int simd_waves,
                        bool exact = false)
{
    if (nx <= 0 || ny <= 0 || cus <= 0 || simd_waves <= 0) return 16;
    const int cols = (solver == HX_SOLVER_HYDROSTATIC_ROE) ? kHydroCols : kRoeCols;
    const int64_t nsx = (nx + cols - 1) / cols;
    return rows_rule(solver, dtype, [&](int r) { return nsx * ((ny + r - 1) / r); }, cus, simd_waves,
                     exact);
}
// One height for a launch over n patches (a batch of scenarios): their
// strips are dealt together, so the rule applies to their sum.
inline int default_rows_multi(int solver, int dtype, const int* nx, const int* ny, int n, int cus,
                              int simd_waves, bool exact = false)
{
    if (n <= 0 || cus <= 0 || simd_waves <= 0) return 16;
    const int cols = (solver == HX_SOLVER_HYDROSTATIC_ROE) ? kHydroCols : kRoeCols;
    for (int q = 0; q < n; ++q)
        if (nx[q] <= 0 || ny[q] <= 0) return 16;
    return rows_rule(solver, dtype, [&](int r) {
        int64_t t = 0;
        for (int q = 0; q < n; ++q) t += (int64_t)((nx[q] + cols - 1) / cols) * ((ny[q] + r - 1) / r);
        return t;
    }, cus, simd_waves, exact);
}
// Whether the fused loop's steps fold their CFL partials into every
// sub-slot of a slot (1) or into sub-slot 0 (0), the next step reading the
// same.  Spreading takes the queue of same-line atomics away from the
// workgroups that finish together (levels up to 2048^2: up to 2x faster per
// step); on levels that fill the device twice the partials arrive spread
// over the launch, and reading 64 lines in every wavefront cost more than
// the queue (4096^2: 221.4 against 219.7 us per step FP64, 124.7 against
// 122.8 FP32).  Below 2^23 cells: spread.
inline int cfl_spread_for(int nx, int ny)
{
    return ((int64_t)nx * (int64_t)ny < ((int64_t)1 << 23)) ? 1 : 0;
}
inline int eff_occ(const hx_ctx* c, int solver, int dtype)
{
    const int o = c->occ[solver][dtype];
    return o ? o : default_occ(solver, dtype);
}
// Defined in hx_kernels.hip: waves per SIMD of the advance kernel in use.
int advance_simd_waves(const hx_ctx* c, int solver, int dtype);
// nx, ny: the level the launch covers (0 when the launch covers many).
inline int eff_rows(const hx_ctx* c, int solver, int dtype, int nx, int ny)
{
    const int r = c->rows[solver][dtype];
    if (r) return r;
    if (nx <= 0 || ny <= 0) return 16;
    return default_rows(solver, dtype, nx, ny, c->cus, advance_simd_waves(c, solver, dtype),
                        c->arith == HX_ARITH_EXACT);
}
inline int eff_rows_multi(const hx_ctx* c, int solver, int dtype, const int* nx, const int* ny, int n)
{
    const int r = c->rows[solver][dtype];
    if (r) return r;
    return default_rows_multi(solver, dtype, nx, ny, n, c->cus, advance_simd_waves(c, solver, dtype),
                              c->arith == HX_ARITH_EXACT);
}

}  // namespace host
}  // namespace hx


#define HX_TRY(x)                                  \
    do {                                           \
        hipError_t _e = (x);                       \
        if (_e != hipSuccess) return HX_E_HIP;     \
    } while (0)

namespace hx {
namespace host {

inline int hip_status(hipError_t e) { return (e == hipSuccess) ? HX_OK : HX_E_HIP; }

// 1/d is exact when d is a power of two (mantissa bits all zero).
inline bool is_pow2(double d)
{
    int e;
    return d > 0.0 && std::frexp(d, &e) == 0.5;
}

template <typename R>
inline PatchView<R> view_of(const hx_patch* in)
{
    PatchView<R> v;
    v.h = (const R*)in->h;
    v.hu = (const R*)in->hu;
    v.hv = (const R*)in->hv;
    v.eta = (const R*)in->eta;
    v.z = (const R*)in->z;
    v.n = (const R*)in->n;
    v.pitch = in->pitch;
    v.nx = in->nx;
    v.ny = in->ny;
    v.bl0 = in->bc_lo[0];
    v.bl1 = in->bc_lo[1];
    v.bh0 = in->bc_hi[0];
    v.bh1 = in->bc_hi[1];
    return v;
}

// A strip is addressed as base + 32-bit byte offsets.  The base sits 2 rows
// and 2 columns before the strip's first cell; the hydrostatic march reads
// rows up to rows + 3 of that frame and columns up to nx + 1 + (high ghosts)
// <= pitch + 1 (a stored high ghost makes pitch == nx + ghosts possible), so
// the largest input offset is (rows + 3) span + (pitch + 1) sizeof(R) with
// span = max(pitch, opitch) sizeof(R); outputs stay below rows * span.
// Keeping (rows + 5) span within 4 GiB keeps every offset below 2^32
// (span >= 2 sizeof(R)).  Both pitches are at most 2^26 elements
// (check_patch, hx_advance_level), which leaves at least 3 FP64 rows.
template <typename R>
inline int64_t max_strip_rows(int64_t pitch, int64_t opitch)
{
    const int64_t span = std::max<int64_t>(pitch, opitch) * (int64_t)sizeof(R);
    return ((int64_t)1 << 32) / span - 5;
}
template <typename R>
inline int clamp_strip_rows(int rows, int64_t pitch, int64_t opitch)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>(rows, max_strip_rows<R>(pitch, opitch)));
}

template <typename R>
inline void fill_geometry(StepArgs<R>& a, int solver, int rows, double dx, double dy)
{
    a.dx = (R)dx;
    a.dy = (R)dy;
    a.rdx = (R)(1.0 / dx);
    a.rdy = (R)(1.0 / dy);
    a.dx_pow2 = is_pow2(dx) && is_pow2(dy) ? 1 : 0;
    const int cols = (solver == HX_SOLVER_ROE) ? kRoeCols : kHydroCols;
    rows = clamp_strip_rows<R>(rows, a.p.pitch, a.opitch);
    a.rows = rows;
    a.nsx = (a.p.nx + cols - 1) / cols;
    const int nsy = (a.p.ny + rows - 1) / rows;
    a.nstrips = a.nsx * nsy;
}

#ifndef HX_TAIL_ROWS_ROE
#define HX_TAIL_ROWS_ROE 4
#endif
#ifndef HX_TAIL_ROWS_HYDRO
#define HX_TAIL_ROWS_HYDRO 8
#endif
#ifndef HX_TAIL_PCT
#define HX_TAIL_PCT 100
#endif

// Banded tiling of a level's strips for advance_kernel.  A launch ends when
// its last wavefronts do; with equal strips the SIMDs go idle one by one
// over the last wavefront's lifetime.  The kernel gives XCD x the rows
// [x ny / 8, (x + 1) ny / 8) and marches them bottom to top in dispatch
// order, so the strips that start last are the ones of the band's top
// `tail_band` rows, cut to tail_rows rows: about one round of the XCD's
// wavefront slots (CUs / 8 x 4 SIMDs x waves per SIMD).  Levels that fill
// the device fewer than two times keep the uniform tiling.
template <typename R>
inline void plan_bands(StepArgs<R>& a, int solver, int cus, int simd_waves)
{
    a.band_wgs = 0;
    const int ht = (solver == HX_SOLVER_ROE) ? HX_TAIL_ROWS_ROE : HX_TAIL_ROWS_HYDRO;
    if (ht <= 0 || a.rows <= ht || cus < 8 || simd_waves <= 0) return;
    const int64_t slots_xcd = (int64_t)(cus / 8) * 4 * simd_waves;
    if ((int64_t)a.nstrips < 2 * 8 * slots_xcd) return;
    int64_t tail = (HX_TAIL_PCT * slots_xcd * ht / 100 + a.nsx - 1) / a.nsx;
    tail = std::max<int64_t>(ht, ((tail + ht - 1) / ht) * ht);
    const int bmin = a.p.ny / 8, bmax = (a.p.ny + 7) / 8;
    if (tail + a.rows > bmin) return;
    auto strips = [&](int b) {
        const int64_t nm = (b - tail) / a.rows;
        const int64_t rest = b - nm * a.rows;
        return (int64_t)a.nsx * (nm + (rest + ht - 1) / ht);
    };
    a.tail_rows = ht;
    a.tail_band = (int)tail;
    a.band_wgs = (int)((std::max(strips(bmin), strips(bmax)) + kWaves - 1) / kWaves);
}

// Defined in hx_kernels.hip.
int check_patch(const hx_patch* in, int need_ghost);
int launch_step_f64(const hx_ctx* c, int solver, const StepArgs<double>& a, hipStream_t st);
int launch_adj_batch_f64(const hx_ctx* c, int solver, const AdjBatch& b, hipStream_t st);
int reduce_cfl_f64(const hx_patch* in, CflSlot* slot, hipStream_t st);
int clear_slots(CflSlot* s, int n, int* err, hipStream_t st);

}  // namespace host
}  // namespace hx
